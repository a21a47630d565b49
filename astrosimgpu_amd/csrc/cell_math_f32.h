// Per-cell math of the FP32 step (device only).
//
// State, conductances and rings are FP32; divisions by constants become
// multiplications by FP32 reciprocals and exp/log are the hardware's fast
// FP32 approximations. FP64 is kept where the dynamics are sensitive:
//  * an AdEx substep whose membrane potential is within kNearThresholdSlopes
//    slope factors of V_th (or above) runs entirely in FP64 -- near
//    threshold the exponential term grows like e^((V - V_th)/Delta_T) and
//    decides the spike time;
//  * the astrocyte's Ca-threshold test of the SIC, (Ca - SIC_th) * 1000 > 1,
//    is evaluated in FP64 (Ca is a small difference above a threshold).
// Synaptic input is summed with Neumaier's compensated summation.
#pragma once

#include <hip/hip_runtime.h>

#include "cell_math.h"

namespace asg {

constexpr float kNearThresholdSlopes = 2.0f;

struct NeuronKf {
    float inv_C_m, neg_g_L, E_L, V_th, inv_Delta_T, a, inv_tau_w, V_peak, E_ex, E_in, I_e;
    float gL_DeltaT, psc_ex, psc_in, ex_decay, in_decay;
    float near_V;  // V_th - kNearThresholdSlopes * Delta_T
};

struct AstroKf {
    float inv_ratio_ER_cyt, Kd_IP3_1, Kd_IP3_2, Kd_act, ratio_rate_IP3R, ratio_rate_L;
    float rate_SERCA, Km_SERCA_sq, k_Kd_inh, k_IP3R;
};

// Neumaier's compensated sum: s + c carries the running total.
__device__ inline void comp_add(float& s, float& c, float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x)) {
        c += (s - t) + x;
    } else {
        c += (x - t) + s;
    }
    s = t;
}

__device__ inline void adex_rhs_f32(const NeuronKf& p, float V, float w, float g_ex, float g_in,
                                    float I_ext, float& dV, float& dw) {
    const float Vc = fminf(V, p.V_peak);
    const float I_spike = p.gL_DeltaT * __expf((Vc - p.V_th) * p.inv_Delta_T);
    dV = (p.neg_g_L * (Vc - p.E_L) + I_spike - g_ex * (Vc - p.E_ex) - g_in * (Vc - p.E_in) - w +
          p.I_e + I_ext) * p.inv_C_m;
    dw = (p.a * (Vc - p.E_L) - w) * p.inv_tau_w;
}

// One substep of the AdEx + alpha-conductance update in FP32, or in FP64 via
// the exact cell math when the cell is near threshold. Returns true on a
// spike (the caller keeps at most one per step).
__device__ inline bool adex_substep_mixed(const NeuronKf& pf, const NeuronK& pd, float V_reset,
                                          float b, float I_ext, float hs, float half_h,
                                          float h_sixth, double hs_d, double half_h_d,
                                          double h_sixth_d, float& V, float& w, float& g_ex,
                                          float& dg_ex, float& g_in, float& dg_in, int& refr) {
    if (V > pf.near_V && refr == 0) {
        double Vd = V, wd = w, ge = g_ex, dge = dg_ex, gi = g_in, dgi = dg_in;
        int r = refr;
        const bool spiked = adex_step(pd, static_cast<double>(V_reset), static_cast<double>(b),
                                      static_cast<double>(I_ext), hs_d, half_h_d, h_sixth_d, 1,
                                      Vd, wd, ge, dge, gi, dgi, r);
        V = static_cast<float>(Vd);
        w = static_cast<float>(wd);
        g_ex = static_cast<float>(ge);
        dg_ex = static_cast<float>(dge);
        g_in = static_cast<float>(gi);
        dg_in = static_cast<float>(dgi);
        refr = r;
        return spiked;
    }
    const float g_ex_next = pf.ex_decay * (g_ex + hs * dg_ex);
    const float dg_ex_next = pf.ex_decay * dg_ex;
    const float g_in_next = pf.in_decay * (g_in + hs * dg_in);
    const float dg_in_next = pf.in_decay * dg_in;
    if (refr > 0) {
        V = V_reset;
        float dV, dw;
        adex_rhs_f32(pf, V, w, g_ex, g_in, I_ext, dV, dw);
        w += hs * dw;
    } else {
        const float g_ex_mid = 0.5f * (g_ex + g_ex_next);
        const float g_in_mid = 0.5f * (g_in + g_in_next);
        float k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
        adex_rhs_f32(pf, V, w, g_ex, g_in, I_ext, k1V, k1w);
        adex_rhs_f32(pf, V + half_h * k1V, w + half_h * k1w, g_ex_mid, g_in_mid, I_ext, k2V, k2w);
        adex_rhs_f32(pf, V + half_h * k2V, w + half_h * k2w, g_ex_mid, g_in_mid, I_ext, k3V, k3w);
        adex_rhs_f32(pf, V + hs * k3V, w + hs * k3w, g_ex_next, g_in_next, I_ext, k4V, k4w);
        V += h_sixth * (k1V + 2.0f * k2V + 2.0f * k3V + k4V);
        w += h_sixth * (k1w + 2.0f * k2w + 2.0f * k3w + k4w);
    }
    g_ex = g_ex_next;
    dg_ex = dg_ex_next;
    g_in = g_in_next;
    dg_in = dg_in_next;
    if (refr > 0) {
        --refr;
    } else if (V >= pf.V_peak) {
        V = V_reset;
        w += b;
        refr = pd.refr_steps;
        return true;
    }
    return false;
}

__device__ inline void astro_rhs_f32(const AstroKf& c, float Ca_tot, float IP3_0, float inv_tau,
                                     float Ca, float IP3, float h, float noise, float& dCa,
                                     float& dIP3, float& dh) {
    const float Ca_ER = (Ca_tot - Ca) * c.inv_ratio_ER_cyt;
    const float m_inf = __fdividef(IP3, IP3 + c.Kd_IP3_1);
    const float n_inf = __fdividef(Ca, Ca + c.Kd_act);
    const float m3 = m_inf * m_inf * m_inf;
    const float n3 = n_inf * n_inf * n_inf;
    const float h3 = h * h * h;
    const float J_channel = c.ratio_rate_IP3R * m3 * n3 * h3 * (Ca_ER - Ca);
    const float J_pump = __fdividef(c.rate_SERCA * Ca * Ca, c.Km_SERCA_sq + Ca * Ca);
    const float J_leak = c.ratio_rate_L * (Ca_ER - Ca);
    const float alpha_h = __fdividef(c.k_Kd_inh * (IP3 + c.Kd_IP3_1), IP3 + c.Kd_IP3_2);
    const float beta_h = c.k_IP3R * Ca;
    dCa = J_channel - J_pump + J_leak + noise;
    dIP3 = (IP3_0 - IP3) * inv_tau;
    dh = alpha_h * (1.0f - h) - beta_h * h;
}

__device__ inline void astro_step_f32(const AstroKf& c, float Ca_tot, float IP3_0, float inv_tau,
                                      float delta_IP3, float ip3_input, float noise, float hs,
                                      float half_h, float h_sixth, int substeps, float& Ca,
                                      float& IP3, float& h) {
    if (ip3_input != 0.0f) {
        IP3 += delta_IP3 * ip3_input;
    }
    for (int s = 0; s < substeps; ++s) {
        float k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca, IP3, h, noise, k1C, k1I, k1h);
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca + half_h * k1C, IP3 + half_h * k1I,
                      h + half_h * k1h, noise, k2C, k2I, k2h);
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca + half_h * k2C, IP3 + half_h * k2I,
                      h + half_h * k2h, noise, k3C, k3I, k3h);
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca + hs * k3C, IP3 + hs * k3I, h + hs * k3h,
                      noise, k4C, k4I, k4h);
        Ca += h_sixth * (k1C + 2.0f * k2C + 2.0f * k3C + k4C);
        IP3 += h_sixth * (k1I + 2.0f * k2I + 2.0f * k3I + k4I);
        h += h_sixth * (k1h + 2.0f * k2h + 2.0f * k3h + k4h);
        Ca = Ca < 0.0f ? 0.0f : (Ca > Ca_tot ? Ca_tot : Ca);
        h = h < 0.0f ? 0.0f : (h > 1.0f ? 1.0f : h);
        IP3 = IP3 < 0.0f ? 0.0f : IP3;
    }
}

// The SIC factor with the threshold test in FP64 (astrocyte.cpp:331-339).
__device__ inline float sic_factor_mixed(float Ca, double SIC_th, double SIC_scale) {
    const double y = (static_cast<double>(Ca) - SIC_th) * 1000.0;
    if (y <= 1.0) {
        return 0.0f;
    }
    return static_cast<float>(SIC_scale) * __logf(static_cast<float>(y));
}

}  // namespace asg

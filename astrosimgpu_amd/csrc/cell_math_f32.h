// Per-cell math of the FP32 step (device only).
//
// Conductances, the SIC current, parameters and the drive are stored in
// FP32; the slow state -- V and w of a neuron, Ca, IP3 and h of an astrocyte
// -- is carried in FP64 between steps. A neuron's step is integrated in FP64
// registers -- FP64 vector instructions issue at the FP32 rate on this GPU,
// and near threshold the exponential term e^((V - V_th)/Delta_T) decides the
// spike time -- with the divisions by C_m, Delta_T and tau_w taken as
// products with FP64 reciprocals; the conductances are rounded to FP32 at
// the end of every step, so results do not depend on window boundaries.
// Astrocytes evaluate their right-hand sides in FP32 with hardware
// reciprocals and add each step's increment to the FP64 state; their
// Ca-threshold test of the SIC, (Ca - SIC_th) * 1000 > 1, is evaluated in
// FP64 (Ca is a small difference above a threshold). Random draws use the
// FP64 step's counter streams: identical Poisson counts, and normals from the
// same two uniforms through FP32 hardware functions. Synaptic input arrives through fixed-point rings
// (exact integer sums); the SIC gather sums with Neumaier's compensation.
#pragma once

#include <hip/hip_runtime.h>

#include "cell_math.h"
#include "rng.h"

namespace asg {

// Neuron constants of the FP32 step's FP64 integration, folded so that the
// right-hand side needs no division and fewer operations:
//   (V - V_th) / Delta_T * log2(e) = V * exp_a + exp_b, and
//   -g_L (V - E_L) - g_ex (V - E_ex) - g_in (V - E_in) + I_e
//       = -V (g_L + g_ex + g_in) + (g_L E_L + I_e) + g_ex E_ex + g_in E_in.
struct NeuronKf {
    double inv_C_m, inv_tau_w;
    double exp_a, exp_b;
    double gL_EL_Ie;  // g_L E_L + I_e
    double a_EL;      // a E_L
    double g_L;
};

struct AstroKf {
    float inv_ratio_ER_cyt, Kd_IP3_1, Kd_IP3_2, Kd_act, ratio_rate_IP3R, ratio_rate_L;
    float rate_SERCA, Km_SERCA_sq, k_Kd_inh, k_IP3R;
};

inline AstroKf make_astro_kf(const asg_astro_constants& c) {
    AstroKf k{};
    k.inv_ratio_ER_cyt = static_cast<float>(1.0 / c.ratio_ER_cyt);
    k.Kd_IP3_1 = static_cast<float>(c.Kd_IP3_1);
    k.Kd_IP3_2 = static_cast<float>(c.Kd_IP3_2);
    k.Kd_act = static_cast<float>(c.Kd_act);
    k.ratio_rate_IP3R = static_cast<float>(c.ratio_ER_cyt * c.rate_IP3R);
    k.ratio_rate_L = static_cast<float>(c.ratio_ER_cyt * c.rate_L);
    k.rate_SERCA = static_cast<float>(c.rate_SERCA);
    k.Km_SERCA_sq = static_cast<float>(c.Km_SERCA * c.Km_SERCA);
    k.k_Kd_inh = static_cast<float>(c.k_IP3R * c.Kd_inh);
    k.k_IP3R = static_cast<float>(c.k_IP3R);
    return k;
}

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

// 2^y of the AdEx spike current: y rounded to FP32 and the hardware
// v_exp_f32 (about 1 ulp), with no range reduction, at half the issue cycles
// of an FP64 polynomial. Its relative error, ln2 |y| 2^-24 from rounding y
// plus the instruction's own, is about 1e-7 near threshold (|y| < 2), where
// the current is g_L Delta_T ~ 40 pA: a change of V of ~1e-9 mV per step,
// three orders below an FP32 ulp of V.
__device__ __forceinline__ double spike_exp2(double y) {
    return static_cast<double>(__builtin_amdgcn_exp2f(static_cast<float>(y)));
}

// The conductance terms of the AdEx right-hand side for one (g_ex, g_in):
// gsum = g_L + g_ex + g_in and gc = (g_L E_L + I_e + I_ext) + g_ex E_ex +
// g_in E_in, shared by the RK4 stages that use the same conductances.
struct AdexG {
    double gsum, gc;
};
__device__ inline AdexG adex_g(const NeuronK& p, const NeuronKf& pf, double g_ex, double g_in,
                               double I_ext) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    return AdexG{pf.g_L + g_ex + g_in,
                 fma(g_in, p.E_in, fma(g_ex, p.E_ex, pf.gL_EL_Ie + I_ext))};
}

// The AdEx right-hand side (neuron.cpp:85-99) in FP64 on the folded
// constants: dV = (gc - w - V gsum + g_L Delta_T 2^y) / C_m,
// dw = (a V - a E_L - w) / tau_w, with V clamped at V_peak.
__device__ inline void adex_rhs_folded(const NeuronK& p, const NeuronKf& pf, double V, double w,
                                       const AdexG& g, double& dV, double& dw) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    const double Vc = fmin(V, p.V_peak);  // one v_min_f64 (V is never NaN)
    const double I_spike = p.gL_DeltaT * spike_exp2(fma(Vc, pf.exp_a, pf.exp_b));
    dV = (fma(-Vc, g.gsum, g.gc - w) + I_spike) * pf.inv_C_m;
    dw = (fma(p.a, Vc, -pf.a_EL) - w) * pf.inv_tau_w;
}

// x rounded to FP32 and back.
__device__ __forceinline__ double round_f32(double x) {
    return static_cast<double>(static_cast<float>(x));
}

// One substep of the AdEx + alpha-conductance update (neuron.cpp:101-141)
// on FP64 registers: RK4 for a non-refractory cell, the clamped Euler update
// of w for a refractory one. Returns true on a spike (the caller keeps at
// most one per step).
__device__ inline bool adex_substep_reg(const NeuronK& pd, const NeuronKf& pf, double V_reset,
                                        double b, double I_ext, double hs, double half_h,
                                        double h_sixth, double& V, double& w, double& g_ex,
                                        double& dg_ex, double& g_in, double& dg_in, int& refr) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    const double ge_next = pd.ex_decay * (g_ex + hs * dg_ex);
    const double gi_next = pd.in_decay * (g_in + hs * dg_in);
    if (refr > 0) {
        V = V_reset;
        const double Vc = fmin(V, pd.V_peak);
        w += hs * ((fma(pd.a, Vc, -pf.a_EL) - w) * pf.inv_tau_w);
    } else {
        const AdexG g1 = adex_g(pd, pf, g_ex, g_in, I_ext);
        const AdexG g4 = adex_g(pd, pf, ge_next, gi_next, I_ext);
        // the midpoint conductances' terms: gsum and gc are affine in
        // (g_ex, g_in), so they are the means of the end points'
        const AdexG g2{0.5 * (g1.gsum + g4.gsum), 0.5 * (g1.gc + g4.gc)};
        double k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
        adex_rhs_folded(pd, pf, V, w, g1, k1V, k1w);
        adex_rhs_folded(pd, pf, V + half_h * k1V, w + half_h * k1w, g2, k2V, k2w);
        adex_rhs_folded(pd, pf, V + half_h * k2V, w + half_h * k2w, g2, k3V, k3w);
        adex_rhs_folded(pd, pf, V + hs * k3V, w + hs * k3w, g4, k4V, k4w);
        V += h_sixth * (k1V + 2.0 * k2V + 2.0 * k3V + k4V);
        w += h_sixth * (k1w + 2.0 * k2w + 2.0 * k3w + k4w);
    }
    g_ex = ge_next;
    dg_ex = pd.ex_decay * dg_ex;
    g_in = gi_next;
    dg_in = pd.in_decay * dg_in;
    if (refr > 0) {
        --refr;
        return false;
    }
    if (V >= pd.V_peak) {
        V = V_reset;
        w += b;
        refr = pd.refr_steps;
        return true;
    }
    return false;
}

// Box-Muller on the two uniforms the FP64 draw uses, with the hardware
// log2, square root and cos(2 pi x) (v_log_f32, v_sqrt_f32, v_cos_f32, about
// 1 ulp each; v_cos_f32 takes its argument in revolutions, so no range
// reduction). u1 <= 1e-300 (probability 2^-53) takes log2(1e-300), as the
// clamped FP64 draw does.
__device__ __forceinline__ float box_muller_f32(double u1, double u2) {
    constexpr float kLog2Of1em300 = -996.578428466208704f;
    constexpr float kMinus2Ln2 = -1.386294361119890618f;
    const float l2 = u1 <= 1e-300 ? kLog2Of1em300 : __builtin_amdgcn_logf(static_cast<float>(u1));
    return __builtin_amdgcn_sqrtf(kMinus2Ln2 * l2) * __builtin_amdgcn_cosf(static_cast<float>(u2));
}

// rng_normal_fresh's draw (u1 redrawn while <= 1e-300) in FP32, from the
// stream's key (rng_key).
__device__ __forceinline__ float rng_normal_fresh_f32(std::uint64_t key) {
    std::uint64_t counter = 0;
    double u1;
    do {
        u1 = rng_uniform_keyed(key, counter++);
    } while (u1 <= 1e-300);
    return box_muller_f32(u1, rng_uniform_keyed(key, counter));
}

// rng_normal_clamped's draw (u1 clamped at 1e-300) in FP32, from the
// stream's key.
__device__ __forceinline__ float rng_normal_clamped_f32(std::uint64_t key) {
    return box_muller_f32(rng_uniform_keyed(key, 0), rng_uniform_keyed(key, 1));
}


__device__ inline void astro_rhs_f32(const AstroKf& c, float Ca_tot, float IP3_0, float inv_tau,
                                     float Ca, float IP3, float h, float noise, float& dCa,
                                     float& dIP3, float& dh) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    // The Li-Rinzel terms of astro_derivatives, restated for fewer
    // instructions: m_inf^3 n_inf^3 h^3 = (m_inf n_inf h)^3; the four
    // quotients take two hardware reciprocals, one per pair of denominators
    // (d1 d2 and d3 d4, all positive); J_channel + J_leak share the factor
    // (Ca_ER - Ca).
    const float d1 = IP3 + c.Kd_IP3_1;
    const float d2 = Ca + c.Kd_act;
    const float Ca2 = Ca * Ca;
    const float d3 = c.Km_SERCA_sq + Ca2;
    const float d4 = IP3 + c.Kd_IP3_2;
    const float r12 = __builtin_amdgcn_rcpf(d1 * d2);
    const float r34 = __builtin_amdgcn_rcpf(d3 * d4);
    const float mnh = IP3 * Ca * r12 * h;  // m_inf n_inf h
    const float gap = (Ca_tot - Ca) * c.inv_ratio_ER_cyt - Ca;  // Ca_ER - Ca
    const float J_ch_leak = (c.ratio_rate_IP3R * (mnh * mnh * mnh) + c.ratio_rate_L) * gap;
    const float J_pump = c.rate_SERCA * Ca2 * (r34 * d4);
    const float alpha_h = c.k_Kd_inh * d1 * (r34 * d3);
    const float beta_h = c.k_IP3R * Ca;
    dCa = J_ch_leak - J_pump + noise;
    dIP3 = (IP3_0 - IP3) * inv_tau;
    dh = alpha_h * (1.0f - h) - beta_h * h;
}

// One astrocyte step with FP32 right-hand sides. S is the type the state
// (Ca, IP3, h) is carried in between steps: float for the FP32-state
// experiment of the isolated update, double for the FP32 step of the
// simulation. Each RK4 stage is evaluated in FP32 from the state rounded to
// FP32; the step's increment h/6 (k1 + 2 k2 + 2 k3 + k4), formed in FP32, is
// then added to the state in S. With S = double no increment is lost to the
// state's rounding: near equilibrium an increment can fall below half an
// FP32 ulp of the state (h relaxes with a time constant of seconds), which
// FP32 state would drop at every step.
template <typename S>
__device__ inline void astro_step_f32(const AstroKf& c, float Ca_tot, float IP3_0, float inv_tau,
                                      float delta_IP3, float ip3_input, float noise, float hs,
                                      float half_h, float h_sixth, int substeps, S& Ca, S& IP3,
                                      S& h) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    if (ip3_input != 0.0f) {
        IP3 += static_cast<S>(delta_IP3) * static_cast<S>(ip3_input);
    }
    for (int s = 0; s < substeps; ++s) {
        const float Ca0 = static_cast<float>(Ca), IP30 = static_cast<float>(IP3),
                    h0 = static_cast<float>(h);
        float k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca0, IP30, h0, noise, k1C, k1I, k1h);
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca0 + half_h * k1C, IP30 + half_h * k1I,
                      h0 + half_h * k1h, noise, k2C, k2I, k2h);
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca0 + half_h * k2C, IP30 + half_h * k2I,
                      h0 + half_h * k2h, noise, k3C, k3I, k3h);
        astro_rhs_f32(c, Ca_tot, IP3_0, inv_tau, Ca0 + hs * k3C, IP30 + hs * k3I, h0 + hs * k3h,
                      noise, k4C, k4I, k4h);
        Ca += static_cast<S>(h_sixth * (k1C + 2.0f * k2C + 2.0f * k3C + k4C));
        IP3 += static_cast<S>(h_sixth * (k1I + 2.0f * k2I + 2.0f * k3I + k4I));
        h += static_cast<S>(h_sixth * (k1h + 2.0f * k2h + 2.0f * k3h + k4h));
        const S Ct = static_cast<S>(Ca_tot);
        Ca = Ca < S(0) ? S(0) : (Ca > Ct ? Ct : Ca);
        h = h < S(0) ? S(0) : (h > S(1) ? S(1) : h);
        IP3 = IP3 < S(0) ? S(0) : IP3;
    }
}

// The SIC factor with the threshold test in FP64 (astrocyte.cpp:331-339).
__device__ inline float sic_factor_mixed(double Ca, double SIC_th, double SIC_scale) {
    const double y = (Ca - SIC_th) * 1000.0;
    if (y <= 1.0) {
        return 0.0f;
    }
    return static_cast<float>(SIC_scale) * __logf(static_cast<float>(y));
}

}  // namespace asg

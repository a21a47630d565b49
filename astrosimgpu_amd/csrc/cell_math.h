// Per-cell FP64 model math of the device step: the Li-Rinzel astrocyte RK4
// update and the AdEx neuron RK4 update with exact alpha-conductance
// propagation. Operation order is the reference's, term for term, and the
// file is compiled with -ffp-contract=off, so each step rounds exactly as the
// reference's host build does. Products of constants that the reference
// evaluates first in a left-to-right chain (ratio_ER_cyt * rate_IP3R, ...)
// are precomputed once; that is the same rounding.
#pragma once

#include "amath.h"
#include "asg.h"

namespace asg {

// a / b rounded to nearest, from y = 1/b rounded to nearest (b a positive
// normal constant with a normal reciprocal). q0 = a y is within 1.5 ulp of
// a/b; one correction q1 = q0 + (a - b q0) y makes it faithful, and for a
// faithful q the remainder a - b q is exact and q + (a - b q) y rounds to
// RN(a/b) (Markstein's theorem: y within half an ulp of 1/b, q faithful,
// no overflow or underflow; here |a| is 0 or far from both). A zero or
// non-finite a takes q0 = a y, which is a / b's +-0, +-inf or NaN. Eight
// full-rate instructions in place of the eleven of an IEEE division with
// its quarter-rate reciprocal.
ASG_HD double div_rn(double a, double b, double y) {
    const double q0 = a * y;
    const double q1 = fma(fma(-b, q0, a), y, q0);
    const double q2 = fma(fma(-b, q1, a), y, q1);
    return (a == 0.0 || !__builtin_isfinite(a)) ? q0 : q2;
}

// AstroConstants (astrocyte_kernel.hpp:29-40) with the constant products of
// astro_derivatives folded in.
struct AstroK {
    double ratio_ER_cyt;
    double inv_ratio_ER_cyt;  // 1 / ratio_ER_cyt rounded to nearest
    double Kd_IP3_1, Kd_IP3_2, Kd_act;
    double ratio_rate_IP3R;   // ratio_ER_cyt * rate_IP3R
    double ratio_rate_L;      // ratio_ER_cyt * rate_L
    double rate_SERCA;
    double Km_SERCA_sq;       // Km_SERCA * Km_SERCA
    double k_Kd_inh;          // k_IP3R * Kd_inh
    double k_IP3R;
};

inline AstroK make_astro_k(const asg_astro_constants& c) {
    AstroK k{};
    k.ratio_ER_cyt = c.ratio_ER_cyt;
    k.inv_ratio_ER_cyt = 1.0 / c.ratio_ER_cyt;
    k.Kd_IP3_1 = c.Kd_IP3_1;
    k.Kd_IP3_2 = c.Kd_IP3_2;
    k.Kd_act = c.Kd_act;
    k.ratio_rate_IP3R = c.ratio_ER_cyt * c.rate_IP3R;
    k.ratio_rate_L = c.ratio_ER_cyt * c.rate_L;
    k.rate_SERCA = c.rate_SERCA;
    k.Km_SERCA_sq = c.Km_SERCA * c.Km_SERCA;
    k.k_Kd_inh = c.k_IP3R * c.Kd_inh;
    k.k_IP3R = c.k_IP3R;
    return k;
}

// astro_derivatives, astrocyte_kernel.hpp:47-68, with its divisions by
// ratio_ER_cyt and tau_IP3 as div_rn (inv_tau = 1 / tau_IP3 rounded to
// nearest; the same roundings).
ASG_HD void astro_rhs(const AstroK& c, double Ca_tot, double IP3_0, double tau_IP3,
                      double inv_tau, double Ca, double IP3, double h, double noise, double& dCa,
                      double& dIP3, double& dh) {
    const double Ca_ER = div_rn(Ca_tot - Ca, c.ratio_ER_cyt, c.inv_ratio_ER_cyt);
    const double m_inf = IP3 / (IP3 + c.Kd_IP3_1);
    const double n_inf = Ca / (Ca + c.Kd_act);
    const double m3 = m_inf * m_inf * m_inf;
    const double n3 = n_inf * n_inf * n_inf;
    const double h3 = h * h * h;
    const double J_channel = c.ratio_rate_IP3R * m3 * n3 * h3 * (Ca_ER - Ca);
    const double J_pump = c.rate_SERCA * Ca * Ca / (c.Km_SERCA_sq + Ca * Ca);
    const double J_leak = c.ratio_rate_L * (Ca_ER - Ca);
    const double alpha_h = c.k_Kd_inh * (IP3 + c.Kd_IP3_1) / (IP3 + c.Kd_IP3_2);
    const double beta_h = c.k_IP3R * Ca;
    dCa = J_channel - J_pump + J_leak + noise;
    dIP3 = div_rn(IP3_0 - IP3, tau_IP3, inv_tau);
    dh = alpha_h * (1.0 - h) - beta_h * h;
}

// astro_advance, astrocyte_kernel.hpp:71-103. half_h = 0.5 * h_step and
// h_sixth = h_step / 6.0, evaluated as the reference does.
ASG_HD void astro_step(const AstroK& c, double Ca_tot, double IP3_0, double tau_IP3,
                       double delta_IP3, double ip3_input, double noise, double h_step,
                       double half_h, double h_sixth, int substeps, double& Ca, double& IP3,
                       double& h) {
    if (ip3_input != 0.0) {
        IP3 += delta_IP3 * ip3_input;
    }
    const double inv_tau = 1.0 / tau_IP3;
    for (int s = 0; s < substeps; ++s) {
        double k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, inv_tau, Ca, IP3, h, noise, k1C, k1I, k1h);
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, inv_tau, Ca + half_h * k1C, IP3 + half_h * k1I,
                  h + half_h * k1h, noise, k2C, k2I, k2h);
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, inv_tau, Ca + half_h * k2C, IP3 + half_h * k2I,
                  h + half_h * k2h, noise, k3C, k3I, k3h);
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, inv_tau, Ca + h_step * k3C, IP3 + h_step * k3I,
                  h + h_step * k3h, noise, k4C, k4I, k4h);
        Ca += h_sixth * (k1C + 2.0 * k2C + 2.0 * k3C + k4C);
        IP3 += h_sixth * (k1I + 2.0 * k2I + 2.0 * k3I + k4I);
        h += h_sixth * (k1h + 2.0 * k2h + 2.0 * k3h + k4h);
        Ca = Ca < 0.0 ? 0.0 : (Ca > Ca_tot ? Ca_tot : Ca);
        h = h < 0.0 ? 0.0 : (h > 1.0 ? 1.0 : h);
        IP3 = IP3 < 0.0 ? 0.0 : IP3;
    }
}

// AstrocytePopulation::sic_factor, astrocyte.cpp:331-339.
ASG_HD double sic_factor(double Ca, double SIC_th, double SIC_scale) {
    const double y = (Ca - SIC_th) * 1000.0;
    if (y <= 1.0) {
        return 0.0;
    }
    return SIC_scale * asg_log(y);
}

// One neuron population's constants (NeuronParams, parameters.hpp:40-58),
// with the step-invariant values the reference recomputes per call
// precomputed by the same functions: exp(-h/tau_syn), e/tau_syn, g_L*Delta_T.
struct NeuronK {
    double C_m, neg_g_L, E_L, V_th, Delta_T, a, tau_w, V_peak, E_ex, E_in, I_e;
    double inv_C_m, inv_Delta_T, inv_tau_w;  // 1 / C_m, ... rounded to nearest
    double gL_DeltaT;           // g_L * Delta_T
    double psc_ex, psc_in;      // e / tau_syn_ex, e / tau_syn_in
    double ex_decay, in_decay;  // exp(-h / tau_syn_ex), exp(-h / tau_syn_in)
    int refr_steps;             // int(t_ref / h + 0.5)
};

// NeuronPopulation::derivatives, neuron.cpp:85-99, with its three divisions
// by constants as div_rn (the same roundings); exp's table from exp_tab (a
// copy of ASG_EXP_TAB).
ASG_HD void adex_rhs(const NeuronK& p, double V, double w, double g_ex, double g_in, double I_ext,
                     const std::uint64_t* exp_tab, double& dV, double& dw) {
    const double Vc = (p.V_peak < V) ? p.V_peak : V;
    const double I_spike =
        p.gL_DeltaT * asg_exp_with(div_rn(Vc - p.V_th, p.Delta_T, p.inv_Delta_T), ExpTabAt{exp_tab});
    const double I_syn_ex = g_ex * (Vc - p.E_ex);
    const double I_syn_in = g_in * (Vc - p.E_in);
    dV = div_rn(p.neg_g_L * (Vc - p.E_L) + I_spike - I_syn_ex - I_syn_in - w + p.I_e + I_ext,
                p.C_m, p.inv_C_m);
    dw = div_rn(p.a * (Vc - p.E_L) - w, p.tau_w, p.inv_tau_w);
}

// The substep loop of NeuronPopulation::update, neuron.cpp:171-229 (exp's
// table from exp_tab, a copy of ASG_EXP_TAB). Returns
// true if the cell spiked in this step (at most once, as the reference's
// `spiked` flag).
ASG_HD bool adex_step(const NeuronK& p, double V_reset, double b, double I_ext, double h_step,
                      double half_h, double h_sixth, int substeps, double& V, double& w,
                      double& g_ex, double& dg_ex, double& g_in, double& dg_in, int& refr,
                      const std::uint64_t* exp_tab) {
    bool spiked = false;
    for (int s = 0; s < substeps; ++s) {
        const double g_ex_next = p.ex_decay * (g_ex + h_step * dg_ex);
        const double dg_ex_next = p.ex_decay * dg_ex;
        const double g_in_next = p.in_decay * (g_in + h_step * dg_in);
        const double dg_in_next = p.in_decay * dg_in;
        if (refr > 0) {
            V = V_reset;
            double dV, dw;
            adex_rhs(p, V, w, g_ex, g_in, I_ext, exp_tab, dV, dw);
            w += h_step * dw;
        } else {
            const double g_ex_mid = 0.5 * (g_ex + g_ex_next);
            const double g_in_mid = 0.5 * (g_in + g_in_next);
            double k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
            adex_rhs(p, V, w, g_ex, g_in, I_ext, exp_tab, k1V, k1w);
            adex_rhs(p, V + half_h * k1V, w + half_h * k1w, g_ex_mid, g_in_mid, I_ext, exp_tab,
                     k2V, k2w);
            adex_rhs(p, V + half_h * k2V, w + half_h * k2w, g_ex_mid, g_in_mid, I_ext, exp_tab,
                     k3V, k3w);
            adex_rhs(p, V + h_step * k3V, w + h_step * k3w, g_ex_next, g_in_next, I_ext, exp_tab,
                     k4V, k4w);
            V += h_sixth * (k1V + 2.0 * k2V + 2.0 * k3V + k4V);
            w += h_sixth * (k1w + 2.0 * k2w + 2.0 * k3w + k4w);
        }
        g_ex = g_ex_next;
        dg_ex = dg_ex_next;
        g_in = g_in_next;
        dg_in = dg_in_next;
        if (refr > 0) {
            --refr;
        } else if (V >= p.V_peak) {
            V = V_reset;
            w += b;
            refr = p.refr_steps;
            spiked = true;
        }
    }
    return spiked;
}

}  // namespace asg

// TEST INFRASTRUCTURE ONLY -- the CPU oracle of the simulation step.
//
// A plain restatement of the reference's time step, written to follow the
// reference's own loops rather than the device design: per-synapse
// Tsodyks-Markram state, scatter of every spike's rows into the rings in
// source order, scatter of the SIC into ring_sic, pending flags, arrival
// application over the sink list. It consumes the same asg_net_desc the
// device does (the connectivity is built once on the host), so the device
// and the oracle step identical networks.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
// this library, and only as the checker or the timed CPU baseline.
//
// Transcendentals: libm_mode 0 calls the host C library (what the reference
// build calls), libm_mode 1 calls amath.h's asg_exp/asg_log/asg_sin/asg_cos
// (what the device calls). The two modes differ only where amath differs
// from the C library (sin/cos); exp and log agree bit for bit.
//
// Reference map (file:line in the reference tree):
//   rng_bits / uniform / normal / poisson  include/astrosimgpu/rng.hpp:26-124
//   astro_derivatives / astro_advance      include/astrosimgpu/astrocyte_kernel.hpp:47-103
//   AstrocytePopulation::update            src/astrocyte.cpp:233-329
//   AstrocytePopulation::sic_factor        src/astrocyte.cpp:331-339
//   NeuronPopulation::add_synaptic_input   src/neuron.cpp:77-83
//   NeuronPopulation::derivatives          src/neuron.cpp:85-99
//   NeuronPopulation::update               src/neuron.cpp:101-250
//   Network::stp_weight                    src/network.cpp:230-247
//   Network::deliver_spikes                src/network.cpp:249-279
//   Network::deliver_sic                   src/network.cpp:281-299
//   Network::apply_arrivals                src/network.cpp:301-338
//   Network::drive_astrocytes              src/network.cpp:340-355
//   Network::run (step order)              src/network.cpp:385-433
//   ring sizing, sic_sources, sinks        src/network.cpp:73-107
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "amath.h"
#include "asg.h"

namespace {

using u64 = std::uint64_t;

struct GlibcMath {
    static double exp(double x) { return std::exp(x); }
    static double log(double x) { return std::log(x); }
    static double sqrt(double x) { return std::sqrt(x); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
};

struct AMath {
    static double exp(double x) { return asg::asg_exp(x); }
    static double log(double x) { return asg::asg_log(x); }
    static double sqrt(double x) { return std::sqrt(x); }
    static double sin(double x) { return asg::asg_sin(x); }
    static double cos(double x) { return asg::asg_cos(x); }
};

u64 mix(u64 seed, u64 stream, u64 counter) {
    u64 z = seed;
    z += stream * 0x9E3779B97F4A7C15ULL;
    z += counter * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double unif(u64 seed, u64 stream, u64 counter) {
    return static_cast<double>(mix(seed, stream, counter) >> 11) * (1.0 / 9007199254740992.0);
}

template <class M>
double free_normal(u64 seed, u64 stream) {
    double u1 = unif(seed, stream, 0);
    if (u1 <= 1e-300) u1 = 1e-300;
    const double u2 = unif(seed, stream, 1);
    return M::sqrt(-2.0 * M::log(u1)) * M::cos(6.283185307179586476925286766559 * u2);
}

template <class M>
struct Rng {
    u64 seed, stream, counter = 0;
    bool has_spare = false;
    double spare = 0.0;
    Rng(u64 s, u64 st) : seed(s), stream(st) {}
    double uniform() { return unif(seed, stream, counter++); }
    double normal() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u1, u2;
        do {
            u1 = uniform();
        } while (u1 <= 1e-300);
        u2 = uniform();
        const double r = M::sqrt(-2.0 * M::log(u1));
        const double theta = 6.283185307179586476925286766559 * u2;
        spare = r * M::sin(theta);
        has_spare = true;
        return r * M::cos(theta);
    }
    double normal(double mean, double sd) { return mean + sd * normal(); }
    int poisson(double lambda) {
        if (lambda <= 0.0) return 0;
        if (lambda > 30.0) {
            const double v = normal(lambda, M::sqrt(lambda));
            return v < 0.0 ? 0 : static_cast<int>(v + 0.5);
        }
        const double limit = M::exp(-lambda);
        double product = uniform();
        int count = 0;
        while (product > limit) {
            ++count;
            product *= uniform();
        }
        return count;
    }
};

struct Set {
    std::vector<uint32_t> row, tgt;
    double weight = 0.0;
    int delay = 1;
    std::vector<double> x, u, t_last;  // per synapse
    void load(const asg_csr& c, const asg_stp& stp, bool plastic) {
        row.assign(c.row_start, c.row_start + c.n_rows + 1);
        tgt.assign(c.target, c.target + row.back());
        weight = c.weight;
        delay = c.delay_steps;
        if (plastic && stp.enabled) {
            x.assign(tgt.size(), 1.0);
            u.assign(tgt.size(), stp.U);
            t_last.assign(tgt.size(), -1e12);
        }
    }
};

struct Spike {
    int64_t step;
    uint32_t source;
};

}  // namespace

struct oracle_run {
    asg_net_desc d{};
    int libm_mode = 0;
    uint32_t NE = 0, NI = 0, N = 0, NA = 0;
    // neurons
    std::vector<double> V, w, gex, dgex, gin, dgin, Isic, exc_in, inh_in, psc_ex, psc_in, Vreset, b;
    std::vector<int> refr;
    // astrocytes
    std::vector<double> Ca, IP3, hh, ip3_in, Ca_tot, IP3_0, tau, delta;
    // connectivity and rings
    Set exc, inh, n2a, a2n;
    int S = 1, sic_delay = 1;
    std::vector<std::vector<double>> ring_exc, ring_inh, ring_sic, ring_astro;
    std::vector<char> pending;
    std::vector<uint32_t> sic_sources, sinks;
    int64_t step = 0;
    std::vector<Spike> raster, spikes;
    double seconds[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

template <class M>
void astro_rhs(const asg_astro_constants& c, double Ca_tot, double IP3_0, double tau, double Ca,
               double IP3, double h, double noise, double& dCa, double& dIP3, double& dh) {
    const double Ca_ER = (Ca_tot - Ca) / c.ratio_ER_cyt;
    const double m_inf = IP3 / (IP3 + c.Kd_IP3_1);
    const double n_inf = Ca / (Ca + c.Kd_act);
    const double m3 = m_inf * m_inf * m_inf;
    const double n3 = n_inf * n_inf * n_inf;
    const double h3 = h * h * h;
    const double J_channel = c.ratio_ER_cyt * c.rate_IP3R * m3 * n3 * h3 * (Ca_ER - Ca);
    const double J_pump = c.rate_SERCA * Ca * Ca / (c.Km_SERCA * c.Km_SERCA + Ca * Ca);
    const double J_leak = c.ratio_ER_cyt * c.rate_L * (Ca_ER - Ca);
    const double alpha_h = c.k_IP3R * c.Kd_inh * (IP3 + c.Kd_IP3_1) / (IP3 + c.Kd_IP3_2);
    const double beta_h = c.k_IP3R * Ca;
    dCa = J_channel - J_pump + J_leak + noise;
    dIP3 = (IP3_0 - IP3) / tau;
    dh = alpha_h * (1.0 - h) - beta_h * h;
}

template <class M>
void astro_advance(const asg_astro_constants& c, double Ca_tot, double IP3_0, double tau,
                   double delta, double input, double noise, double hs, int substeps, double& Ca,
                   double& IP3, double& h) {
    if (input != 0.0) IP3 += delta * input;
    for (int s = 0; s < substeps; ++s) {
        double k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs<M>(c, Ca_tot, IP3_0, tau, Ca, IP3, h, noise, k1C, k1I, k1h);
        astro_rhs<M>(c, Ca_tot, IP3_0, tau, Ca + 0.5 * hs * k1C, IP3 + 0.5 * hs * k1I,
                     h + 0.5 * hs * k1h, noise, k2C, k2I, k2h);
        astro_rhs<M>(c, Ca_tot, IP3_0, tau, Ca + 0.5 * hs * k2C, IP3 + 0.5 * hs * k2I,
                     h + 0.5 * hs * k2h, noise, k3C, k3I, k3h);
        astro_rhs<M>(c, Ca_tot, IP3_0, tau, Ca + hs * k3C, IP3 + hs * k3I, h + hs * k3h, noise,
                     k4C, k4I, k4h);
        Ca += (hs / 6.0) * (k1C + 2.0 * k2C + 2.0 * k3C + k4C);
        IP3 += (hs / 6.0) * (k1I + 2.0 * k2I + 2.0 * k3I + k4I);
        h += (hs / 6.0) * (k1h + 2.0 * k2h + 2.0 * k3h + k4h);
        Ca = Ca < 0.0 ? 0.0 : (Ca > Ca_tot ? Ca_tot : Ca);
        h = h < 0.0 ? 0.0 : (h > 1.0 ? 1.0 : h);
        IP3 = IP3 < 0.0 ? 0.0 : IP3;
    }
}

template <class M>
void neuron_rhs(const asg_neuron_params& p, double V, double w, double g_ex, double g_in,
                double I_ext, double& dV, double& dw) {
    const double Vc = std::min(V, p.V_peak);
    const double I_spike = p.g_L * p.Delta_T * M::exp((Vc - p.V_th) / p.Delta_T);
    const double I_syn_ex = g_ex * (Vc - p.E_ex);
    const double I_syn_in = g_in * (Vc - p.E_in);
    dV = (-p.g_L * (Vc - p.E_L) + I_spike - I_syn_ex - I_syn_in - w + p.I_e + I_ext) / p.C_m;
    dw = (p.a * (Vc - p.E_L) - w) / p.tau_w;
}

int noise_steps(double noise_dt, double dt) {
    return std::max(1, static_cast<int>(std::llround(noise_dt / dt)));
}

void add_input(oracle_run& r, uint32_t i, double weight) {
    if (weight > 0.0) {
        r.exc_in[i] += weight;
    } else if (weight < 0.0) {
        r.inh_in[i] -= weight;
    }
}

template <class M>
double stp_weight(oracle_run& r, Set& s, uint32_t k, double t_now) {
    if (!r.d.stp.enabled) return s.weight;
    const double dt = t_now - s.t_last[k];
    const double x_decay = M::exp(-dt / r.d.stp.tau_rec);
    const double u_decay = r.d.stp.tau_fac < 1e-10 ? 0.0 : M::exp(-dt / r.d.stp.tau_fac);
    double& x = s.x[k];
    double& u = s.u[k];
    x = 1.0 + (x - x * u - 1.0) * x_decay;
    u = r.d.stp.U + u * (1.0 - r.d.stp.U) * u_decay;
    s.t_last[k] = t_now;
    return s.weight * x * u;
}

template <class M>
void one_step(oracle_run& r) {
    using clock = std::chrono::steady_clock;
    const asg_net_desc& d = r.d;
    const int64_t step = r.step;
    const double dt = d.dt;
    const double hs = dt / static_cast<double>(d.substeps);
    auto t0 = clock::now();

    // apply_arrivals
    {
        const int slot = static_cast<int>(step % r.S);
        const bool sic_arrives = r.pending[slot] != 0;
        r.pending[slot] = 0;
        std::vector<double>& ex = r.ring_exc[slot];
        std::vector<double>& in = r.ring_inh[slot];
        std::vector<double>& sic = r.ring_sic[slot];
        for (uint32_t i = 0; i < r.N; ++i) {
            if (ex[i] != 0.0) {
                add_input(r, i, ex[i]);
                ex[i] = 0.0;
            }
            if (in[i] != 0.0) {
                add_input(r, i, -in[i]);
                in[i] = 0.0;
            }
            if (sic_arrives) {
                r.Isic[i] = sic[i];
                sic[i] = 0.0;
            }
        }
        std::vector<double>& ai = r.ring_astro[slot];
        for (const uint32_t a : r.sinks) {
            if (ai[a] != 0.0) {
                r.ip3_in[a] += ai[a];
                ai[a] = 0.0;
            }
        }
    }
    auto t1 = clock::now();
    // drive_astrocytes (serial, as the reference)
    if (d.input_astro.poiss_rate > 0.0) {
        const double lambda = d.input_astro.poiss_rate * dt * 1e-3;
        for (uint32_t a = 0; a < r.NA; ++a) {
            Rng<M> g(d.seed ^ 0xC2B2AE3D27D4EB4FULL, static_cast<u64>(step) * 1000003ULL + a);
            const int ev = g.poisson(lambda);
            if (ev > 0) r.ip3_in[a] += d.input_astro.poiss_weight * static_cast<double>(ev);
        }
    }
    auto t2 = clock::now();
    // astrocyte update
    {
        const asg_input_params& in = d.input_astro;
        const int spn = noise_steps(in.noise_dt, dt);
        const u64 nidx = static_cast<u64>(step / spn);
        const u64 nseed = d.seed ^ 0x9E3779B97F4A7C15ULL;
        const double nstd = in.gauss_noise_std;
        const bool indep = in.independent_noise != 0;
        const double shared = (nstd > 0.0 && !indep) ? nstd * free_normal<M>(nseed, nidx) : 0.0;
        const int64_t n = r.NA;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const double noise =
                (nstd > 0.0 && indep)
                    ? nstd * free_normal<M>(nseed, nidx * 1000003ULL + static_cast<u64>(i))
                    : shared;
            double ca = r.Ca[i], ip3 = r.IP3[i], h = r.hh[i];
            astro_advance<M>(d.astro.c, r.Ca_tot[i], r.IP3_0[i], r.tau[i], r.delta[i], r.ip3_in[i],
                             noise, hs, d.substeps, ca, ip3, h);
            r.Ca[i] = ca;
            r.IP3[i] = ip3;
            r.hh[i] = h;
            r.ip3_in[i] = 0.0;
        }
    }
    auto t3 = clock::now();
    // neuron update
    r.spikes.clear();
    {
        const asg_input_params* ins[2] = {&d.input_exc, &d.input_inh};
        const int spn_e = noise_steps(d.input_exc.noise_dt, dt);
        const int spn_i = noise_steps(d.input_inh.noise_dt, dt);
        const u64 nidx_e = static_cast<u64>(step / spn_e);
        const u64 nidx_i = static_cast<u64>(step / spn_i);
        double noise_e = 0.0, noise_i = 0.0;
        if (d.input_exc.gauss_noise_std > 0.0 && !d.input_exc.independent_noise) {
            Rng<M> g(d.seed ^ 0xD1B54A32D192ED03ULL, nidx_e);
            noise_e = g.normal(0.0, d.input_exc.gauss_noise_std);
        }
        if (d.input_inh.gauss_noise_std > 0.0 && !d.input_inh.independent_noise) {
            Rng<M> g(d.seed ^ 0xA24BAED4963EE407ULL, nidx_i);
            noise_i = g.normal(0.0, d.input_inh.gauss_noise_std);
        }
        std::vector<Spike> local;
#pragma omp parallel
        {
            std::vector<Spike> mine;
#pragma omp for schedule(static) nowait
            for (int64_t idx = 0; idx < static_cast<int64_t>(r.N); ++idx) {
                const uint32_t cell = static_cast<uint32_t>(idx);
                const bool excit = cell < r.NE;
                const asg_neuron_params& p = d.neuron[excit ? 0 : 1];
                const asg_input_params& in = *ins[excit ? 0 : 1];
                if (in.poiss_rate > 0.0) {
                    Rng<M> g(d.seed ^ 0x2545F4914F6CDD1DULL, static_cast<u64>(step) * 1000003ULL + cell);
                    const int ev = g.poisson(in.poiss_rate * dt * 1e-3);
                    if (ev > 0) r.exc_in[cell] += in.poiss_weight * static_cast<double>(ev);
                }
                double I_ext = r.Isic[cell];
                if (in.gauss_noise_std > 0.0) {
                    if (in.independent_noise) {
                        const u64 ni = excit ? nidx_e : nidx_i;
                        Rng<M> g(d.seed ^ 0x9E3779B185EBCA87ULL, ni * 1000033ULL + cell);
                        I_ext += g.normal(0.0, in.gauss_noise_std);
                    } else {
                        I_ext += excit ? noise_e : noise_i;
                    }
                }
                if (r.exc_in[cell] != 0.0) {
                    r.dgex[cell] += r.exc_in[cell] * r.psc_ex[cell];
                    r.exc_in[cell] = 0.0;
                }
                if (r.inh_in[cell] != 0.0) {
                    r.dgin[cell] += r.inh_in[cell] * r.psc_in[cell];
                    r.inh_in[cell] = 0.0;
                }
                double V = r.V[cell], w = r.w[cell], ge = r.gex[cell], dge = r.dgex[cell],
                       gi = r.gin[cell], dgi = r.dgin[cell];
                bool spiked = false;
                for (int s = 0; s < d.substeps; ++s) {
                    const double ex_decay = M::exp(-hs / p.tau_syn_ex);
                    const double in_decay = M::exp(-hs / p.tau_syn_in);
                    const double ge_n = ex_decay * (ge + hs * dge);
                    const double dge_n = ex_decay * dge;
                    const double gi_n = in_decay * (gi + hs * dgi);
                    const double dgi_n = in_decay * dgi;
                    if (r.refr[cell] > 0) {
                        V = r.Vreset[cell];
                        double dV, dw;
                        neuron_rhs<M>(p, V, w, ge, gi, I_ext, dV, dw);
                        w += hs * dw;
                    } else {
                        const double ge_m = 0.5 * (ge + ge_n);
                        const double gi_m = 0.5 * (gi + gi_n);
                        double k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
                        neuron_rhs<M>(p, V, w, ge, gi, I_ext, k1V, k1w);
                        neuron_rhs<M>(p, V + 0.5 * hs * k1V, w + 0.5 * hs * k1w, ge_m, gi_m, I_ext, k2V, k2w);
                        neuron_rhs<M>(p, V + 0.5 * hs * k2V, w + 0.5 * hs * k2w, ge_m, gi_m, I_ext, k3V, k3w);
                        neuron_rhs<M>(p, V + hs * k3V, w + hs * k3w, ge_n, gi_n, I_ext, k4V, k4w);
                        V += (hs / 6.0) * (k1V + 2.0 * k2V + 2.0 * k3V + k4V);
                        w += (hs / 6.0) * (k1w + 2.0 * k2w + 2.0 * k3w + k4w);
                    }
                    ge = ge_n;
                    dge = dge_n;
                    gi = gi_n;
                    dgi = dgi_n;
                    if (r.refr[cell] > 0) {
                        --r.refr[cell];
                    } else if (V >= p.V_peak) {
                        V = r.Vreset[cell];
                        w += r.b[cell];
                        r.refr[cell] = static_cast<int>(p.t_ref / hs + 0.5);
                        spiked = true;
                    }
                }
                r.V[cell] = V;
                r.w[cell] = w;
                r.gex[cell] = ge;
                r.dgex[cell] = dge;
                r.gin[cell] = gi;
                r.dgin[cell] = dgi;
                if (spiked) mine.push_back(Spike{step, cell});
            }
#pragma omp critical
            local.insert(local.end(), mine.begin(), mine.end());
        }
        std::sort(local.begin(), local.end(),
                  [](const Spike& a, const Spike& b) { return a.source < b.source; });
        r.spikes = local;
    }
    auto t4 = clock::now();
    // deliver_spikes
    {
        const double t_now = static_cast<double>(step) * dt;
        for (const Spike& s : r.spikes) {
            if (s.source < r.NE) {
                const uint32_t src = s.source;
                for (uint32_t k = r.exc.row[src]; k < r.exc.row[src + 1]; ++k) {
                    const int slot = static_cast<int>((step + r.exc.delay) % r.S);
                    r.ring_exc[slot][r.exc.tgt[k]] += stp_weight<M>(r, r.exc, k, t_now);
                }
                for (uint32_t k = r.n2a.row[src]; k < r.n2a.row[src + 1]; ++k) {
                    const int slot = static_cast<int>((step + r.n2a.delay) % r.S);
                    r.ring_astro[slot][r.n2a.tgt[k]] += stp_weight<M>(r, r.n2a, k, t_now);
                }
            } else {
                const uint32_t src = s.source - r.NE;
                for (uint32_t k = r.inh.row[src]; k < r.inh.row[src + 1]; ++k) {
                    const int slot = static_cast<int>((step + r.inh.delay) % r.S);
                    r.ring_inh[slot][r.inh.tgt[k]] -= stp_weight<M>(r, r.inh, k, t_now);
                }
            }
        }
    }
    auto t5 = clock::now();
    // deliver_sic
    if (step % d.sic_interval == 0) {
        r.pending[(step + r.sic_delay) % r.S] = 1;
        for (const uint32_t a : r.sic_sources) {
            const double y = (r.Ca[a] - d.astro.SIC_th) * 1000.0;
            const double factor = y <= 1.0 ? 0.0 : d.astro.SIC_scale * M::log(y);
            if (factor == 0.0) continue;
            for (uint32_t k = r.a2n.row[a]; k < r.a2n.row[a + 1]; ++k) {
                const int slot = static_cast<int>((step + r.a2n.delay) % r.S);
                r.ring_sic[slot][r.a2n.tgt[k]] += r.a2n.weight * factor;
            }
        }
    }
    auto t6 = clock::now();
    r.raster.insert(r.raster.end(), r.spikes.begin(), r.spikes.end());
    auto sec = [](clock::time_point a, clock::time_point b) {
        return std::chrono::duration<double>(b - a).count();
    };
    r.seconds[0] += sec(t0, t1);  // arrival application
    r.seconds[1] += sec(t1, t2);  // background input
    r.seconds[2] += sec(t2, t3);  // update astrocytes
    r.seconds[3] += sec(t3, t4);  // update neurons
    r.seconds[4] += sec(t4, t5);  // spike delivery
    r.seconds[5] += sec(t5, t6);  // SIC gather + deliver
    ++r.step;
}

}  // namespace

extern "C" {

oracle_run* oracle_create(const asg_net_desc* d, int libm_mode) {
    auto* r = new oracle_run;
    r->d = *d;
    r->libm_mode = libm_mode;
    r->NE = d->n_exc;
    r->NI = d->n_inh;
    r->N = d->n_exc + d->n_inh;
    r->NA = d->n_astro;
    const uint32_t N = r->N, NA = r->NA;
    r->V.resize(N);
    r->Vreset.assign(d->V_reset, d->V_reset + N);
    r->b.assign(d->b, d->b + N);
    r->psc_ex.resize(N);
    r->psc_in.resize(N);
    for (uint32_t i = 0; i < N; ++i) {
        const asg_neuron_params& p = d->neuron[i < d->n_exc ? 0 : 1];
        r->V[i] = p.V_m_init;
        r->psc_ex[i] = 2.7182818284590452353602874713527 / p.tau_syn_ex;
        r->psc_in[i] = 2.7182818284590452353602874713527 / p.tau_syn_in;
    }
    for (auto* v : {&r->w, &r->gex, &r->dgex, &r->gin, &r->dgin, &r->Isic, &r->exc_in, &r->inh_in}) {
        v->assign(N, 0.0);
    }
    r->refr.assign(N, 0);
    r->Ca.assign(NA, d->astro.Ca_init);
    r->IP3.assign(NA, d->astro.IP3_init);
    r->hh.assign(NA, d->astro.h_init);
    r->ip3_in.assign(NA, 0.0);
    r->Ca_tot.assign(d->Ca_tot, d->Ca_tot + NA);
    r->IP3_0.assign(d->IP3_0, d->IP3_0 + NA);
    r->tau.assign(d->tau_IP3, d->tau_IP3 + NA);
    r->delta.assign(d->delta_IP3, d->delta_IP3 + NA);
    r->exc.load(d->exc, d->stp, true);
    r->inh.load(d->inh, d->stp, true);
    r->n2a.load(d->n2a, d->stp, true);
    r->a2n.load(d->a2n, d->stp, false);
    int max_delay = std::max(std::max(1, d->exc.delay_steps),
                             std::max(d->inh.delay_steps, d->a2n.delay_steps));
    max_delay = std::max(max_delay, d->n2a.delay_steps);
    r->S = max_delay + 1;
    r->sic_delay = d->a2n.delay_steps;
    r->ring_exc.assign(r->S, std::vector<double>(N, 0.0));
    r->ring_inh.assign(r->S, std::vector<double>(N, 0.0));
    r->ring_sic.assign(r->S, std::vector<double>(N, 0.0));
    r->ring_astro.assign(r->S, std::vector<double>(NA, 0.0));
    r->pending.assign(r->S, 0);
    for (uint32_t a = 0; a < NA; ++a) {
        if (r->a2n.row[a] < r->a2n.row[a + 1]) r->sic_sources.push_back(a);
    }
    std::vector<char> seen(NA, 0);
    for (const uint32_t t : r->n2a.tgt) {
        if (!seen[t]) {
            seen[t] = 1;
            r->sinks.push_back(t);
        }
    }
    std::sort(r->sinks.begin(), r->sinks.end());
    return r;
}

void oracle_destroy(oracle_run* r) { delete r; }

void oracle_step(oracle_run* r, int64_t n_steps) {
    for (int64_t s = 0; s < n_steps; ++s) {
        if (r->libm_mode == 0) {
            one_step<GlibcMath>(*r);
        } else {
            one_step<AMath>(*r);
        }
    }
}

int64_t oracle_current_step(const oracle_run* r) { return r->step; }

int64_t oracle_spike_count(const oracle_run* r) { return static_cast<int64_t>(r->raster.size()); }

// Copies raster entries [first, first + cap) to the host arrays.
int64_t oracle_spikes(const oracle_run* r, int64_t first, int64_t* step, uint32_t* src, int64_t cap) {
    const int64_t total = static_cast<int64_t>(r->raster.size());
    int64_t n = 0;
    for (int64_t i = first; i < total && n < cap; ++i, ++n) {
        step[n] = r->raster[i].step;
        src[n] = r->raster[i].source;
    }
    return n;
}

// Full field copy; field ids are asg.h's ASG_FIELD_*. STP fields report the
// first synapse of each source's exc row (or inh row), NaN for empty rows.
int oracle_read_state(const oracle_run* r, int field, double* out) {
    const std::vector<double>* v = nullptr;
    switch (field) {
        case ASG_FIELD_V: v = &r->V; break;
        case ASG_FIELD_W: v = &r->w; break;
        case ASG_FIELD_G_EX: v = &r->gex; break;
        case ASG_FIELD_DG_EX: v = &r->dgex; break;
        case ASG_FIELD_G_IN: v = &r->gin; break;
        case ASG_FIELD_DG_IN: v = &r->dgin; break;
        case ASG_FIELD_I_SIC: v = &r->Isic; break;
        case ASG_FIELD_CA: v = &r->Ca; break;
        case ASG_FIELD_IP3: v = &r->IP3; break;
        case ASG_FIELD_H: v = &r->hh; break;
        default: break;
    }
    if (v != nullptr) {
        std::memcpy(out, v->data(), v->size() * sizeof(double));
        return 0;
    }
    if (field == ASG_FIELD_STP_X || field == ASG_FIELD_STP_U || field == ASG_FIELD_STP_TLAST) {
        for (uint32_t s = 0; s < r->N; ++s) {
            const Set& set = s < r->NE ? r->exc : r->inh;
            const uint32_t row = s < r->NE ? s : s - r->NE;
            if (!r->d.stp.enabled || set.row[row] == set.row[row + 1]) {
                out[s] = std::nan("");
                continue;
            }
            const uint32_t k = set.row[row];
            out[s] = field == ASG_FIELD_STP_X ? set.x[k] : field == ASG_FIELD_STP_U ? set.u[k] : set.t_last[k];
        }
        return 0;
    }
    return -1;
}

void oracle_phase_seconds(const oracle_run* r, double* out6) {
    for (int i = 0; i < 6; ++i) out6[i] = r->seconds[i];
}

}  // extern "C"

// Element-wise exp/log/sin/cos/sqrt through the host C library (mode 0) or
// amath.h (mode 1), for checking amath against the C library.
// The isolated astrocyte update at scale (asg_astro_scaling), FP64 on the
// CPU: the same initial state, per-cell parameters and spike counts, drawn
// from the same counter streams, then astro_advance per cell and step.
namespace {
double fresh_normal(u64 seed, u64 stream, u64& ctr) {
    double u1;
    do {
        u1 = unif(seed, stream, ctr++);
    } while (u1 <= 1e-300);
    const double u2 = unif(seed, stream, ctr++);
    return AMath::sqrt(-2.0 * AMath::log(u1)) * AMath::cos(6.283185307179586476925286766559 * u2);
}

double scaled(u64 seed, u64 stream, double mean, double lower, double upper, double var) {
    const double lo = mean < 0.0 ? mean * upper : mean * lower;
    const double hi = mean < 0.0 ? mean * lower : mean * upper;
    const double sd = var * std::fabs(mean);
    u64 ctr = 0;
    for (int tries = 0; tries < 256; ++tries) {
        const double v = mean + sd * fresh_normal(seed, stream, ctr);
        if (v >= lo && v <= hi) return v;
    }
    return mean;
}
}  // namespace

extern "C" void oracle_astro_scaling(uint32_t n, int32_t steps, uint64_t seed,
                                     const asg_astro_constants* c, const double* nominal4,
                                     double dt, double lambda, int randomized, double lower,
                                     double upper, double var, double* Ca, double* IP3,
                                     double* h) {
    const u64 s_init = seed ^ 0x6A09E667F3BCC909ULL;
    const u64 s_par = seed ^ 0xBB67AE8584CAA73BULL;
    std::vector<double> p4(4 * static_cast<size_t>(n));
    for (uint32_t i = 0; i < n; ++i) {
        Ca[i] = 0.05 + 0.35 * unif(s_init, i, 0);
        IP3[i] = 0.01 + 0.49 * unif(s_init, i, 1);
        h[i] = 0.50 + 0.45 * unif(s_init, i, 2);
        for (int q = 0; q < 4; ++q) {
            p4[4 * i + q] = randomized ? scaled(s_par, 4 * static_cast<u64>(i) + q, nominal4[q],
                                                lower, upper, var)
                                       : nominal4[q];
        }
    }
    for (int32_t s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
            Rng<AMath> rng(seed, static_cast<u64>(s) * n + static_cast<u64>(i));
            const int k = rng.poisson(lambda);
            astro_advance<AMath>(*c, p4[4 * i], p4[4 * i + 1], p4[4 * i + 2], p4[4 * i + 3],
                                 static_cast<double>(k), 0.0, dt, 1, Ca[i], IP3[i], h[i]);
        }
    }
}

extern "C" void oracle_math(int fn, int mode, const double* x, double* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const double v = x[i];
        double r;
        if (mode == 0) {
            r = fn == 0 ? std::exp(v) : fn == 1 ? std::log(v) : fn == 2 ? std::sin(v)
                : fn == 3 ? std::cos(v) : std::sqrt(v);
        } else {
            r = fn == 0 ? asg::asg_exp(v) : fn == 1 ? asg::asg_log(v) : fn == 2 ? asg::asg_sin(v)
                : fn == 3 ? asg::asg_cos(v) : asg::asg_sqrt(v);
        }
        y[i] = r;
    }
}

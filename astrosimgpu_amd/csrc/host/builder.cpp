#include "builder.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_set>

#include "../rng.h"

namespace asg_host {

double CounterRng::uniform() {
    return static_cast<double>(asg::rng_bits(seed_, stream_, counter_++) >> 11) *
           (1.0 / 9007199254740992.0);
}

double CounterRng::normal() {
    if (has_spare_) {
        has_spare_ = false;
        return spare_;
    }
    double u1;
    do {
        u1 = uniform();
    } while (u1 <= 1e-300);
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double theta = asg::kTwoPi * u2;
    spare_ = r * std::sin(theta);
    has_spare_ = true;
    return r * std::cos(theta);
}

double CounterRng::normal_redraw(double mean, double stddev, double lo, double hi, int max_tries) {
    for (int i = 0; i < max_tries; ++i) {
        const double v = normal(mean, stddev);
        if (v >= lo && v <= hi) return v;
    }
    return std::min(std::max(mean, lo), hi);
}

int delay_to_steps(double delay_ms, double dt) {
    const int steps = static_cast<int>(delay_ms / dt + 0.5);
    return steps < 1 ? 1 : steps;
}

namespace {

// Gaussian around the nominal value, std = var * |mean|, redrawn until it
// lies in [lower, upper] * mean (bounds swap for a negative mean);
// astrocyte.cpp:17-24, neuron.cpp:10-17.
double draw_scaled(double mean, const RandomizeSpec& spec, CounterRng& rng) {
    if (!spec.enabled || spec.var <= 0.0) return mean;
    const double lo = mean < 0.0 ? mean * spec.upper : mean * spec.lower;
    const double hi = mean < 0.0 ? mean * spec.lower : mean * spec.upper;
    return rng.normal_redraw(mean, spec.var * std::abs(mean), lo, hi);
}

void flatten(const std::vector<std::vector<uint32_t>>& rows, Csr& out) {
    out.row_start.assign(rows.size() + 1, 0);
    size_t total = 0;
    for (size_t r = 0; r < rows.size(); ++r) {
        out.row_start[r] = static_cast<uint32_t>(total);
        total += rows[r].size();
    }
    out.row_start[rows.size()] = static_cast<uint32_t>(total);
    out.target.clear();
    out.target.reserve(total);
    for (const auto& r : rows) out.target.insert(out.target.end(), r.begin(), r.end());
}

// network.cpp:110-228.
void build_connections(Model& m, CounterRng& rng) {
    const ModelConfig& c = m.cfg;
    const uint32_t n_exc = c.N_exc, n_inh = c.N_inh, n_neurons = n_exc + n_inh;
    const uint32_t n_astro = c.N_astro;
    std::vector<std::vector<uint32_t>> exc_t(n_exc), astro_t(n_exc), a2n_t(n_astro), inh_t(n_inh);

    // Pools belong to the postsynaptic neuron, per block (exc, then inh).
    std::vector<std::vector<uint32_t>> pool(n_neurons);
    std::vector<uint32_t> candidates;
    auto assign_pools = [&](uint32_t offset, uint32_t count) {
        for (uint32_t local = 0; local < count; ++local) {
            std::vector<uint32_t>& pp = pool[offset + local];
            if (c.conn.pool_type == PoolType::Block) {
                if (n_astro == 0) continue;  // no astrocytes: nothing to pool
                const uint32_t start = static_cast<uint32_t>(
                    (static_cast<uint64_t>(local) * n_astro) / std::max<uint32_t>(count, 1));
                for (int k = 0; k < c.conn.pool_size; ++k) {
                    pp.push_back(static_cast<uint32_t>((start + static_cast<uint32_t>(k)) % n_astro));
                }
            } else {
                // Partial Fisher-Yates without replacement over a fresh list.
                candidates.resize(n_astro);
                for (uint32_t i = 0; i < n_astro; ++i) candidates[i] = i;
                const int take = std::min<int>(c.conn.pool_size, static_cast<int>(n_astro));
                for (int k = 0; k < take; ++k) {
                    const size_t pick =
                        static_cast<size_t>(rng.uniform() * static_cast<double>(candidates.size() - k)) +
                        static_cast<size_t>(k);
                    std::swap(candidates[k], candidates[std::min(pick, candidates.size() - 1)]);
                    pp.push_back(candidates[k]);
                }
            }
        }
    };
    assign_pools(0, n_exc);
    assign_pools(n_exc, n_inh);

    std::unordered_set<uint64_t> a2n_seen;
    auto connect_block = [&](uint32_t post_offset, uint32_t post_count) {
        for (uint32_t pre = 0; pre < n_exc; ++pre) {
            for (uint32_t local = 0; local < post_count; ++local) {
                const uint32_t post = post_offset + local;
                if (!c.conn.allow_autapses && pre == post) continue;
                if (rng.uniform() >= c.conn.p_primary) continue;
                exc_t[pre].push_back(post);
                if (n_astro == 0 || rng.uniform() >= c.conn.p_third_if_primary) continue;
                const std::vector<uint32_t>& pp = pool[post];
                const uint32_t astro =
                    pp[static_cast<size_t>(rng.uniform() * static_cast<double>(pp.size())) % pp.size()];
                astro_t[pre].push_back(astro);
                const uint64_t key = (static_cast<uint64_t>(astro) << 32) | post;
                if (!c.conn.unique_third_out || a2n_seen.insert(key).second) {
                    a2n_t[astro].push_back(post);
                }
            }
        }
    };
    connect_block(0, n_exc);
    connect_block(n_exc, n_inh);

    for (uint32_t pre = 0; pre < n_inh; ++pre) {
        const uint32_t pre_global = n_exc + pre;
        for (uint32_t post = 0; post < n_neurons; ++post) {
            if (!c.conn.allow_autapses && pre_global == post) continue;
            if (rng.uniform() < c.conn.p_primary) inh_t[pre].push_back(post);
        }
    }

    flatten(exc_t, m.exc);
    flatten(inh_t, m.inh);
    flatten(astro_t, m.n2a);
    flatten(a2n_t, m.a2n);
}

// Per-row generator: every row draws from its own counter stream, so rows are
// independent and built in parallel; the gaps between connected targets are
// geometric (Bernoulli(p) per pair, sampled by skipping), so the cost is
// proportional to the connections made rather than to N_pre * N_post.
// Distributionally the same network family as build_connections; not the
// same realisation.
constexpr uint64_t kSaltRowExc = 0x6A09E667F3BCC908ULL;
constexpr uint64_t kSaltRowInh = 0xBB67AE8584CAA73BULL;
constexpr uint64_t kSaltPool = 0x3C6EF372FE94F82BULL;

// Next connected index after `at` among [0, n) with probability p each.
struct GapSampler {
    double log_q;  // log(1 - p)
    double p;
    explicit GapSampler(double prob) : log_q(prob < 1.0 ? std::log1p(-prob) : 0.0), p(prob) {}
    int64_t next(int64_t at, CounterRng& rng) const {
        if (p >= 1.0) return at + 1;
        if (p <= 0.0) return INT64_MAX;
        const double u = 1.0 - rng.uniform();  // (0, 1]
        const double gap = std::floor(std::log(u) / log_q);
        if (gap > 9e18) return INT64_MAX;
        return at + 1 + static_cast<int64_t>(gap);
    }
};

void build_connections_per_row(Model& m) {
    const ModelConfig& c = m.cfg;
    const uint32_t n_exc = c.N_exc, n_inh = c.N_inh, n_neurons = n_exc + n_inh;
    const uint32_t n_astro = c.N_astro;
    std::vector<std::vector<uint32_t>> pool(n_neurons);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t post = 0; post < static_cast<int64_t>(n_neurons); ++post) {
        const bool exc_block = post < n_exc;
        const uint32_t local = exc_block ? static_cast<uint32_t>(post) : static_cast<uint32_t>(post - n_exc);
        const uint32_t count = exc_block ? n_exc : n_inh;
        std::vector<uint32_t>& pp = pool[static_cast<size_t>(post)];
        if (n_astro == 0) continue;
        if (c.conn.pool_type == PoolType::Block) {
            const uint32_t start = static_cast<uint32_t>(
                (static_cast<uint64_t>(local) * n_astro) / std::max<uint32_t>(count, 1));
            for (int k = 0; k < c.conn.pool_size; ++k) {
                pp.push_back(static_cast<uint32_t>((start + static_cast<uint32_t>(k)) % n_astro));
            }
        } else {
            // Uniform subset of pool_size astrocytes (Floyd's algorithm).
            CounterRng rng(c.seed ^ kSaltPool, static_cast<uint64_t>(post));
            const uint32_t take = std::min<uint32_t>(static_cast<uint32_t>(c.conn.pool_size), n_astro);
            for (uint32_t j = n_astro - take; j < n_astro; ++j) {
                const uint32_t r = static_cast<uint32_t>(rng.uniform() * static_cast<double>(j + 1));
                const uint32_t pick = std::min(r, j);
                if (std::find(pp.begin(), pp.end(), pick) == pp.end()) {
                    pp.push_back(pick);
                } else {
                    pp.push_back(j);
                }
            }
        }
    }
    std::vector<std::vector<uint32_t>> exc_t(n_exc), astro_t(n_exc), inh_t(n_inh);
    std::vector<std::vector<uint64_t>> pairs_t(n_exc);  // (astro << 32 | post) per row
    const GapSampler primary(c.conn.p_primary);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pre = 0; pre < static_cast<int64_t>(n_exc); ++pre) {
        CounterRng rng(c.seed ^ kSaltRowExc, static_cast<uint64_t>(pre));
        int64_t post = primary.next(-1, rng);
        while (post < static_cast<int64_t>(n_neurons)) {
            if (c.conn.allow_autapses || post != pre) {
                exc_t[pre].push_back(static_cast<uint32_t>(post));
                if (n_astro > 0 && rng.uniform() < c.conn.p_third_if_primary) {
                    const std::vector<uint32_t>& pp = pool[static_cast<size_t>(post)];
                    const uint32_t astro =
                        pp[static_cast<size_t>(rng.uniform() * static_cast<double>(pp.size())) % pp.size()];
                    astro_t[pre].push_back(astro);
                    pairs_t[pre].push_back((static_cast<uint64_t>(astro) << 32) | static_cast<uint64_t>(post));
                }
            }
            post = primary.next(post, rng);
        }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pre = 0; pre < static_cast<int64_t>(n_inh); ++pre) {
        CounterRng rng(c.seed ^ kSaltRowInh, static_cast<uint64_t>(pre));
        const int64_t pre_global = static_cast<int64_t>(n_exc) + pre;
        int64_t post = primary.next(-1, rng);
        while (post < static_cast<int64_t>(n_neurons)) {
            if (c.conn.allow_autapses || post != pre_global) {
                inh_t[pre].push_back(static_cast<uint32_t>(post));
            }
            post = primary.next(post, rng);
        }
    }
    // Return connections, grouped by astrocyte; with unique_third_out one per
    // (astrocyte, target) pair.
    std::vector<std::vector<uint32_t>> a2n_t(n_astro);
    std::unordered_set<uint64_t> seen;
    for (uint32_t pre = 0; pre < n_exc; ++pre) {
        for (const uint64_t key : pairs_t[pre]) {
            if (!c.conn.unique_third_out || seen.insert(key).second) {
                a2n_t[key >> 32].push_back(static_cast<uint32_t>(key & 0xFFFFFFFFULL));
            }
        }
    }
    flatten(exc_t, m.exc);
    flatten(inh_t, m.inh);
    flatten(astro_t, m.n2a);
    flatten(a2n_t, m.a2n);
}

}  // namespace

void Model::refresh_desc(int32_t precision) {
    const ModelConfig& c = cfg;
    asg_net_desc& d = desc;
    d = asg_net_desc{};
    d.n_exc = c.N_exc;
    d.n_inh = c.N_inh;
    d.n_astro = c.N_astro;
    d.dt = c.dt;
    d.substeps = c.substeps;
    d.seed = c.seed;
    d.neuron[0] = c.neuron_exc;
    d.neuron[1] = c.neuron_inh;
    d.V_reset = V_reset.data();
    d.b = b.data();
    d.input_exc = c.input_exc;
    d.input_inh = c.input_inh;
    d.input_astro = c.input_astro;
    d.astro = c.astro;
    d.Ca_tot = Ca_tot.data();
    d.IP3_0 = IP3_0.data();
    d.tau_IP3 = tau_IP3.data();
    d.delta_IP3 = delta_IP3.data();
    auto set = [](asg_csr& out, const Csr& in, uint32_t rows, double w, int32_t delay) {
        out.n_rows = rows;
        out.row_start = in.row_start.data();
        out.target = in.target.data();
        out.weight = w;
        out.delay_steps = delay;
    };
    set(d.exc, exc, c.N_exc, c.syn.w_e, d_exc);
    set(d.inh, inh, c.N_inh, c.syn.w_i, d_inh);
    set(d.n2a, n2a, c.N_exc, c.syn.w_n2a, d_exc);
    set(d.a2n, a2n, c.N_astro, c.syn.w_a2n, d_a2n);
    d.stp.enabled = c.syn.stp_enabled ? 1 : 0;
    d.stp.U = c.syn.U;
    d.stp.tau_rec = c.syn.tau_rec;
    d.stp.tau_fac = c.syn.tau_fac;
    d.sic_interval = c.syn.sic_interval;
    d.precision = precision;
}

void sequential_build(Model& m);
void finish_model(Model& m);

void build_model(Model& m) {
    const ModelConfig& c = m.cfg;
    if (c.conn.per_row) {
        // Per-cell parameter draws from per-cell streams (same distribution).
        constexpr uint64_t kSaltParamA = 0x510E527FADE682D1ULL;
        constexpr uint64_t kSaltParamN = 0x9B05688C2B3E6C1FULL;
        m.Ca_tot.resize(c.N_astro);
        m.IP3_0.resize(c.N_astro);
        m.tau_IP3.resize(c.N_astro);
        m.delta_IP3.resize(c.N_astro);
        for (uint32_t i = 0; i < c.N_astro; ++i) {
            CounterRng rng(c.seed ^ kSaltParamA, i);
            m.Ca_tot[i] = draw_scaled(c.astro_Ca_tot, c.randomize_astro, rng);
            m.IP3_0[i] = draw_scaled(c.astro_IP3_0, c.randomize_astro, rng);
            m.tau_IP3[i] = draw_scaled(c.astro_tau_IP3, c.randomize_astro, rng);
            m.delta_IP3[i] = draw_scaled(c.astro_delta_IP3, c.randomize_astro, rng);
        }
        const uint32_t n = c.N_exc + c.N_inh;
        m.V_reset.resize(n);
        m.b.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            CounterRng rng(c.seed ^ kSaltParamN, i);
            const asg_neuron_params& src = i < c.N_exc ? c.neuron_exc : c.neuron_inh;
            m.V_reset[i] = draw_scaled(src.V_reset, c.randomize_neuron, rng);
            m.b[i] = draw_scaled(src.b, c.randomize_neuron, rng);
        }
        build_connections_per_row(m);
    } else {
        sequential_build(m);
    }
    finish_model(m);
}

namespace {
}  // namespace

void sequential_build(Model& m) {
    const ModelConfig& c = m.cfg;
    CounterRng rng(c.seed, 0x5EED);

    // AstrocytePopulation::build, astrocyte.cpp:28-53.
    m.Ca_tot.resize(c.N_astro);
    m.IP3_0.resize(c.N_astro);
    m.tau_IP3.resize(c.N_astro);
    m.delta_IP3.resize(c.N_astro);
    for (uint32_t i = 0; i < c.N_astro; ++i) {
        m.Ca_tot[i] = draw_scaled(c.astro_Ca_tot, c.randomize_astro, rng);
        m.IP3_0[i] = draw_scaled(c.astro_IP3_0, c.randomize_astro, rng);
        m.tau_IP3[i] = draw_scaled(c.astro_tau_IP3, c.randomize_astro, rng);
        m.delta_IP3[i] = draw_scaled(c.astro_delta_IP3, c.randomize_astro, rng);
    }
    // NeuronPopulation::build, neuron.cpp:23-75.
    const uint32_t n = c.N_exc + c.N_inh;
    m.V_reset.resize(n);
    m.b.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const asg_neuron_params& src = i < c.N_exc ? c.neuron_exc : c.neuron_inh;
        m.V_reset[i] = draw_scaled(src.V_reset, c.randomize_neuron, rng);
        m.b[i] = draw_scaled(src.b, c.randomize_neuron, rng);
    }
    build_connections(m, rng);
}

void finish_model(Model& m) {
    const ModelConfig& c = m.cfg;
    m.d_exc = delay_to_steps(c.syn.d_e, c.dt);
    m.d_inh = delay_to_steps(c.syn.d_i, c.dt);
    m.d_a2n = delay_to_steps(c.syn.d_a2n, c.dt);
    m.stats.n_astro = c.N_astro;
    m.stats.n_exc = c.N_exc;
    m.stats.n_inh = c.N_inh;
    m.stats.n_primary_exc = static_cast<uint32_t>(m.exc.target.size());
    m.stats.n_primary_inh = static_cast<uint32_t>(m.inh.target.size());
    m.stats.n_neuron_to_astro = static_cast<uint32_t>(m.n2a.target.size());
    m.stats.n_astro_to_neuron = static_cast<uint32_t>(m.a2n.target.size());
    m.refresh_desc(ASG_FP64);
}

}  // namespace asg_host

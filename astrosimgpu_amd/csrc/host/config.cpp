#include "config.h"

#include <sstream>
#include <stdexcept>
#include <vector>

#include "json.h"

namespace asg_host {

namespace {

asg_neuron_params neuron_defaults() {  // NEST aeif_cond_alpha_astro, parameters.hpp:40-58
    asg_neuron_params p{};
    p.C_m = 281.0;
    p.g_L = 30.0;
    p.E_L = -70.6;
    p.V_th = -50.4;
    p.Delta_T = 2.0;
    p.a = 4.0;
    p.b = 80.5;
    p.tau_w = 144.0;
    p.V_reset = -60.0;
    p.V_peak = 0.0;
    p.t_ref = 0.0;
    p.E_ex = 0.0;
    p.E_in = -85.0;
    p.tau_syn_ex = 0.2;
    p.tau_syn_in = 2.0;
    p.I_e = 0.0;
    p.V_m_init = -70.6;
    return p;
}

asg_input_params input_defaults() {  // parameters.hpp:132-138
    asg_input_params p{};
    p.poiss_rate = 0.0;
    p.poiss_weight = 1.0;
    p.gauss_noise_std = 0.0;
    p.noise_dt = 0.0;
    p.independent_noise = 1;
    return p;
}

void load_astro(const JsonValue& j, ModelConfig& c) {  // parameters.cpp:13-35
    if (!j.is_object()) return;
    asg_astro_params& a = c.astro;
    j.read("Ca_tot", c.astro_Ca_tot);
    j.read("IP3_0", c.astro_IP3_0);
    j.read("Kd_IP3_1", a.c.Kd_IP3_1);
    j.read("Kd_IP3_2", a.c.Kd_IP3_2);
    j.read("Kd_act", a.c.Kd_act);
    j.read("Kd_inh", a.c.Kd_inh);
    j.read("Km_SERCA", a.c.Km_SERCA);
    j.read("SIC_scale", a.SIC_scale);
    j.read("SIC_th", a.SIC_th);
    j.read("delta_IP3", c.astro_delta_IP3);
    j.read("k_IP3R", a.c.k_IP3R);
    j.read("rate_IP3R", a.c.rate_IP3R);
    j.read("rate_L", a.c.rate_L);
    j.read("rate_SERCA", a.c.rate_SERCA);
    j.read("ratio_ER_cyt", a.c.ratio_ER_cyt);
    j.read("tau_IP3", c.astro_tau_IP3);
    j.read("Ca_init", a.Ca_init);
    j.read("h_init", a.h_init);
    j.read("IP3", a.IP3_init);        // alias, then the long name wins
    j.read("IP3_init", a.IP3_init);
}

void load_neuron(const JsonValue& j, asg_neuron_params& p) {  // parameters.cpp:37-57
    if (!j.is_object()) return;
    j.read("C_m", p.C_m);
    j.read("g_L", p.g_L);
    j.read("E_L", p.E_L);
    j.read("V_th", p.V_th);
    j.read("Delta_T", p.Delta_T);
    j.read("a", p.a);
    j.read("b", p.b);
    j.read("tau_w", p.tau_w);
    j.read("V_reset", p.V_reset);
    j.read("V_peak", p.V_peak);
    j.read("t_ref", p.t_ref);
    j.read("E_ex", p.E_ex);
    j.read("E_in", p.E_in);
    j.read("tau_syn_ex", p.tau_syn_ex);
    j.read("tau_syn_in", p.tau_syn_in);
    j.read("I_e", p.I_e);
    j.read("V_m", p.V_m_init);
}

void load_input(const JsonValue& j, asg_input_params& p) {  // parameters.cpp:59-69
    if (!j.is_object()) return;
    j.read("poiss_rate", p.poiss_rate);
    j.read("poiss_weight", p.poiss_weight);
    // Both names set the standard deviation; the second wins if both exist.
    j.read("gauss_noise_var", p.gauss_noise_std);
    j.read("gauss_noise_std", p.gauss_noise_std);
    j.read("noise_dt", p.noise_dt);
    bool indep = p.independent_noise != 0;
    j.read("independent_noise", indep);
    p.independent_noise = indep ? 1 : 0;
}

void load_randomize(const JsonValue& j, RandomizeSpec& r) {  // parameters.cpp:71-81
    if (!j.is_object()) return;
    r.enabled = true;
    j.read("enabled", r.enabled);
    j.read("lower", r.lower);
    j.read("upper", r.upper);
    j.read("var", r.var);
}

// "a.b.c=value;d.e=value": value is parsed as a JSON literal.
void apply_assignments(JsonValue& root, const std::string& spec) {
    std::stringstream ss(spec);
    std::string item;
    while (std::getline(ss, item, ';')) {
        if (item.find_first_not_of(" \t") == std::string::npos) continue;
        const size_t eq = item.find('=');
        if (eq == std::string::npos) {
            throw std::runtime_error("override '" + item + "' is not of the form key=value");
        }
        const std::string path = item.substr(0, eq);
        const JsonValue value = JsonValue::parse(item.substr(eq + 1));
        JsonValue* node = &root;
        size_t start = 0;
        for (;;) {
            const size_t dot = path.find('.', start);
            const std::string key = path.substr(start, dot == std::string::npos ? dot : dot - start);
            node = &node->member(key);
            if (dot == std::string::npos) break;
            start = dot + 1;
        }
        node->set_parsed(value);
    }
}

ModelConfig map_config(const JsonValue& root, const std::string& name, const Overrides& ov) {
    if (!root.is_object()) {
        throw std::runtime_error("configuration root must be a JSON object: " + name);
    }
    ModelConfig cfg;
    if (const JsonValue& n = root.get("N"); n.is_object()) {
        n.read("N_A", cfg.N_astro);
        n.read("N_E", cfg.N_exc);
        n.read("N_I", cfg.N_inh);
    }
    if (const JsonValue& s = root.get("simulation"); s.is_object()) {
        s.read("dt", cfg.dt);
        s.read("substeps", cfg.substeps);
        s.read("pre_sim_time", cfg.pre_sim_time);
        s.read("sim_time", cfg.sim_time);
        unsigned long long seed = cfg.seed;
        s.read("seed", seed);
        cfg.seed = seed;
        s.read("output_dir", cfg.output_dir);
        s.read("record_spikes", cfg.record_spikes);
        s.read("record_astro", cfg.record_astro);
        s.read("record_neuron", cfg.record_neuron);
        s.read("record_every", cfg.record_every);
        s.read("record_astro_max", cfg.record_astro_max);
        s.read("record_neuron_max", cfg.record_neuron_max);
    }
    load_astro(root.get("astrocyte"), cfg);
    load_neuron(root.get("neuron_exc"), cfg.neuron_exc);
    load_neuron(root.get("neuron_inh"), cfg.neuron_inh);
    load_input(root.get("input_astro"), cfg.input_astro);
    load_input(root.get("input_exc"), cfg.input_exc);
    load_input(root.get("input_inh"), cfg.input_inh);
    load_randomize(root.get("randomize_astro"), cfg.randomize_astro);
    load_randomize(root.get("randomize_neuron"), cfg.randomize_neuron);
    if (const JsonValue& s = root.get("synapse"); s.is_object()) {
        s.read("w_e", cfg.syn.w_e);
        s.read("w_i", cfg.syn.w_i);
        s.read("w_n2a", cfg.syn.w_n2a);
        s.read("w_a2n", cfg.syn.w_a2n);
        s.read("d_e", cfg.syn.d_e);
        s.read("d_i", cfg.syn.d_i);
        s.read("d_a2n", cfg.syn.d_a2n);
        s.read("sic_interval", cfg.syn.sic_interval);
        if (const JsonValue& stp = s.get("stp"); stp.is_object()) {
            stp.read("enabled", cfg.syn.stp_enabled);
            stp.read("U", cfg.syn.U);
            stp.read("tau_rec", cfg.syn.tau_rec);
            stp.read("tau_fac", cfg.syn.tau_fac);
        }
    }
    if (const JsonValue& a = root.get("analysis"); a.is_object()) {
        a.read("enabled", cfg.analysis.enabled);
        a.read("ca_threshold", cfg.analysis.ca_threshold);
        a.read("merge_gap", cfg.analysis.merge_gap);
        a.read("hist_window", cfg.analysis.hist_window);
        a.read("hist_shift", cfg.analysis.hist_shift);
    }
    if (const JsonValue& c = root.get("connectivity"); c.is_object()) {
        c.read("p_primary", cfg.conn.p_primary);
        c.read("p_third_if_primary", cfg.conn.p_third_if_primary);
        c.read("pool_size", cfg.conn.pool_size);
        c.read("allow_autapses", cfg.conn.allow_autapses);
        c.read("unique_third_out", cfg.conn.unique_third_out);
        std::string generator = "sequential";
        c.read("generator", generator);
        if (generator == "per_row") {
            cfg.conn.per_row = true;
        } else if (generator != "sequential") {
            throw std::runtime_error("unknown connectivity.generator '" + generator +
                                     "' (expected \"sequential\" or \"per_row\")");
        }
        std::string pool_type = "block";
        c.read("pool_type", pool_type);
        if (pool_type == "block") {
            cfg.conn.pool_type = PoolType::Block;
        } else if (pool_type == "random") {
            cfg.conn.pool_type = PoolType::Random;
        } else {
            throw std::runtime_error("unknown pool_type '" + pool_type +
                                     "' (expected \"block\" or \"random\")");
        }
    }
    // An unset noise interval is the noise_generator default, 10 * dt.
    for (asg_input_params* in : {&cfg.input_astro, &cfg.input_exc, &cfg.input_inh}) {
        if (in->noise_dt <= 0.0) in->noise_dt = 10.0 * cfg.dt;
    }
    if (cfg.substeps < 1) throw std::runtime_error("simulation.substeps must be at least 1");
    if (cfg.dt <= 0.0) throw std::runtime_error("simulation.dt must be positive");
    if (cfg.record_every < 1) throw std::runtime_error("simulation.record_every must be at least 1");
    if (cfg.syn.sic_interval < 1) throw std::runtime_error("synapse.sic_interval must be at least 1");
    if (cfg.conn.pool_size < 1) throw std::runtime_error("connectivity.pool_size must be at least 1");

    std::vector<std::string> unused;
    root.unused_keys("", unused);
    if (!unused.empty()) {
        std::string msg = "unrecognised keys in " + name + ":";
        for (const std::string& k : unused) msg += "\n  " + k;
        msg += "\nEither the name is wrong or this binary predates the option.";
        throw std::runtime_error(msg);
    }

    if (!ov.output_dir.empty()) cfg.output_dir = ov.output_dir;
    if (ov.seed >= 0) cfg.seed = static_cast<uint64_t>(ov.seed);
    if (ov.sim_time >= 0.0) cfg.sim_time = ov.sim_time;
    if (ov.pre_time >= 0.0) cfg.pre_sim_time = ov.pre_time;
    if (ov.n_astro >= 0) cfg.N_astro = static_cast<uint32_t>(ov.n_astro);
    return cfg;
}

}  // namespace

ModelConfig::ModelConfig() {
    // AstrocyteParams defaults, NEST astrocyte_lr_1994 (parameters.hpp:14-38).
    astro.c.Kd_IP3_1 = 0.13;
    astro.c.Kd_IP3_2 = 0.9434;
    astro.c.Kd_act = 0.08234;
    astro.c.Kd_inh = 1.049;
    astro.c.Km_SERCA = 0.1;
    astro.c.k_IP3R = 0.0002;
    astro.c.rate_IP3R = 0.006;
    astro.c.rate_L = 0.00011;
    astro.c.rate_SERCA = 0.0009;
    astro.c.ratio_ER_cyt = 0.185;
    astro.SIC_scale = 1.0;
    astro.SIC_th = 0.19669;
    astro.Ca_init = 0.073;
    astro.h_init = 0.793;
    astro.IP3_init = 0.16;
    neuron_exc = neuron_defaults();
    neuron_inh = neuron_defaults();
    input_astro = input_defaults();
    input_exc = input_defaults();
    input_inh = input_defaults();
}

ModelConfig load_config_text(const std::string& text, const std::string& name,
                             const Overrides& ov) {
    JsonValue root = JsonValue::parse(text);
    if (!ov.assignments.empty()) apply_assignments(root, ov.assignments);
    return map_config(root, name, ov);
}

ModelConfig load_config(const std::string& path, const Overrides& ov) {
    JsonValue root = JsonValue::parse_file(path);
    if (!ov.assignments.empty()) apply_assignments(root, ov.assignments);
    return map_config(root, path, ov);
}

}  // namespace asg_host

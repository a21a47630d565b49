#include "config.h"

#include <sstream>
#include <stdexcept>
#include <vector>

#include "config_doc.h"

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

void load_astro(ConfigDoc::Node j, ModelConfig& c) {  // parameters.cpp:13-35
    if (!j.is_object()) return;
    asg_astro_params& a = c.astro;
    j.take("Ca_tot", c.astro_Ca_tot);
    j.take("IP3_0", c.astro_IP3_0);
    j.take("Kd_IP3_1", a.c.Kd_IP3_1);
    j.take("Kd_IP3_2", a.c.Kd_IP3_2);
    j.take("Kd_act", a.c.Kd_act);
    j.take("Kd_inh", a.c.Kd_inh);
    j.take("Km_SERCA", a.c.Km_SERCA);
    j.take("SIC_scale", a.SIC_scale);
    j.take("SIC_th", a.SIC_th);
    j.take("delta_IP3", c.astro_delta_IP3);
    j.take("k_IP3R", a.c.k_IP3R);
    j.take("rate_IP3R", a.c.rate_IP3R);
    j.take("rate_L", a.c.rate_L);
    j.take("rate_SERCA", a.c.rate_SERCA);
    j.take("ratio_ER_cyt", a.c.ratio_ER_cyt);
    j.take("tau_IP3", c.astro_tau_IP3);
    j.take("Ca_init", a.Ca_init);
    j.take("h_init", a.h_init);
    j.take("IP3", a.IP3_init);        // alias, then the long name wins
    j.take("IP3_init", a.IP3_init);
}

void load_neuron(ConfigDoc::Node j, asg_neuron_params& p) {  // parameters.cpp:37-57
    if (!j.is_object()) return;
    j.take("C_m", p.C_m);
    j.take("g_L", p.g_L);
    j.take("E_L", p.E_L);
    j.take("V_th", p.V_th);
    j.take("Delta_T", p.Delta_T);
    j.take("a", p.a);
    j.take("b", p.b);
    j.take("tau_w", p.tau_w);
    j.take("V_reset", p.V_reset);
    j.take("V_peak", p.V_peak);
    j.take("t_ref", p.t_ref);
    j.take("E_ex", p.E_ex);
    j.take("E_in", p.E_in);
    j.take("tau_syn_ex", p.tau_syn_ex);
    j.take("tau_syn_in", p.tau_syn_in);
    j.take("I_e", p.I_e);
    j.take("V_m", p.V_m_init);
}

void load_input(ConfigDoc::Node j, asg_input_params& p) {  // parameters.cpp:59-69
    if (!j.is_object()) return;
    j.take("poiss_rate", p.poiss_rate);
    j.take("poiss_weight", p.poiss_weight);
    // Both names set the standard deviation; the second wins if both exist.
    j.take("gauss_noise_var", p.gauss_noise_std);
    j.take("gauss_noise_std", p.gauss_noise_std);
    j.take("noise_dt", p.noise_dt);
    bool indep = p.independent_noise != 0;
    j.take("independent_noise", indep);
    p.independent_noise = indep ? 1 : 0;
}

void load_randomize(ConfigDoc::Node j, RandomizeSpec& r) {  // parameters.cpp:71-81
    if (!j.is_object()) return;
    r.enabled = true;
    j.take("enabled", r.enabled);
    j.take("lower", r.lower);
    j.take("upper", r.upper);
    j.take("var", r.var);
}

// "a.b.c=value;d.e=value": value is parsed as a JSON literal.
void apply_assignments(ConfigDoc& doc, const std::string& spec) {
    std::stringstream ss(spec);
    std::string item;
    while (std::getline(ss, item, ';')) {
        if (item.find_first_not_of(" \t") == std::string::npos) continue;
        const size_t eq = item.find('=');
        if (eq == std::string::npos) {
            throw std::runtime_error("override '" + item + "' is not of the form key=value");
        }
        doc.assign(item.substr(0, eq), item.substr(eq + 1));
    }
}

ModelConfig map_config(const ConfigDoc& doc, const std::string& name, const Overrides& ov) {
    const ConfigDoc::Node root = doc.root();
    if (!root.is_object()) {
        throw std::runtime_error("configuration root must be a JSON object: " + name);
    }
    ModelConfig cfg;
    if (const ConfigDoc::Node n = root.at("N"); n.is_object()) {
        n.take("N_A", cfg.N_astro);
        n.take("N_E", cfg.N_exc);
        n.take("N_I", cfg.N_inh);
    }
    if (const ConfigDoc::Node s = root.at("simulation"); s.is_object()) {
        s.take("dt", cfg.dt);
        s.take("substeps", cfg.substeps);
        s.take("pre_sim_time", cfg.pre_sim_time);
        s.take("sim_time", cfg.sim_time);
        unsigned long long seed = cfg.seed;
        s.take("seed", seed);
        cfg.seed = seed;
        s.take("output_dir", cfg.output_dir);
        s.take("record_spikes", cfg.record_spikes);
        s.take("record_astro", cfg.record_astro);
        s.take("record_neuron", cfg.record_neuron);
        s.take("record_every", cfg.record_every);
        s.take("record_astro_max", cfg.record_astro_max);
        s.take("record_neuron_max", cfg.record_neuron_max);
    }
    load_astro(root.at("astrocyte"), cfg);
    load_neuron(root.at("neuron_exc"), cfg.neuron_exc);
    load_neuron(root.at("neuron_inh"), cfg.neuron_inh);
    load_input(root.at("input_astro"), cfg.input_astro);
    load_input(root.at("input_exc"), cfg.input_exc);
    load_input(root.at("input_inh"), cfg.input_inh);
    load_randomize(root.at("randomize_astro"), cfg.randomize_astro);
    load_randomize(root.at("randomize_neuron"), cfg.randomize_neuron);
    if (const ConfigDoc::Node s = root.at("synapse"); s.is_object()) {
        s.take("w_e", cfg.syn.w_e);
        s.take("w_i", cfg.syn.w_i);
        s.take("w_n2a", cfg.syn.w_n2a);
        s.take("w_a2n", cfg.syn.w_a2n);
        s.take("d_e", cfg.syn.d_e);
        s.take("d_i", cfg.syn.d_i);
        s.take("d_a2n", cfg.syn.d_a2n);
        s.take("sic_interval", cfg.syn.sic_interval);
        if (const ConfigDoc::Node stp = s.at("stp"); stp.is_object()) {
            stp.take("enabled", cfg.syn.stp_enabled);
            stp.take("U", cfg.syn.U);
            stp.take("tau_rec", cfg.syn.tau_rec);
            stp.take("tau_fac", cfg.syn.tau_fac);
        }
    }
    if (const ConfigDoc::Node a = root.at("analysis"); a.is_object()) {
        a.take("enabled", cfg.analysis.enabled);
        a.take("ca_threshold", cfg.analysis.ca_threshold);
        a.take("merge_gap", cfg.analysis.merge_gap);
        a.take("hist_window", cfg.analysis.hist_window);
        a.take("hist_shift", cfg.analysis.hist_shift);
    }
    if (const ConfigDoc::Node c = root.at("connectivity"); c.is_object()) {
        c.take("p_primary", cfg.conn.p_primary);
        c.take("p_third_if_primary", cfg.conn.p_third_if_primary);
        c.take("pool_size", cfg.conn.pool_size);
        c.take("allow_autapses", cfg.conn.allow_autapses);
        c.take("unique_third_out", cfg.conn.unique_third_out);
        std::string generator = "sequential";
        c.take("generator", generator);
        if (generator == "per_row") {
            cfg.conn.per_row = true;
        } else if (generator != "sequential") {
            throw std::runtime_error("unknown connectivity.generator '" + generator +
                                     "' (expected \"sequential\" or \"per_row\")");
        }
        std::string pool_type = "block";
        c.take("pool_type", pool_type);
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

    const std::vector<std::string> unused = doc.unread_paths();
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
    ConfigDoc doc = ConfigDoc::parse(text);
    if (!ov.assignments.empty()) apply_assignments(doc, ov.assignments);
    return map_config(doc, name, ov);
}

ModelConfig load_config(const std::string& path, const Overrides& ov) {
    ConfigDoc doc = ConfigDoc::load(path);
    if (!ov.assignments.empty()) apply_assignments(doc, ov.assignments);
    return map_config(doc, path, ov);
}

}  // namespace asg_host

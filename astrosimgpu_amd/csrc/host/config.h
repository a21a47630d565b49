// Model configuration: every parameter of one simulation, with the
// reference's defaults (include/astrosimgpu/parameters.hpp:14-198), and the
// strict JSON loader (src/parameters.cpp:85-204).
#pragma once

#include <cstdint>
#include <string>

#include "asg.h"

namespace asg_host {

struct RandomizeSpec {  // parameters.hpp:141-146
    bool enabled = false;
    double lower = 1.0;
    double upper = 1.0;
    double var = 0.0;
};

enum class PoolType { Block, Random };

struct Connectivity {  // parameters.hpp:104-122
    double p_primary = 0.2;
    double p_third_if_primary = 0.2;
    int pool_size = 1;
    PoolType pool_type = PoolType::Block;
    bool allow_autapses = false;
    bool unique_third_out = false;
    // Connection generator: false = the reference's single sequential stream
    // (bit-identical networks, O(N_pre * N_post) draws); true = one counter
    // stream per row with geometric gap sampling (the same Bernoulli
    // distribution of connections, O(connections) draws, rows in parallel).
    // Key "generator": "sequential" | "per_row"; not a reference key.
    bool per_row = false;
};

struct Synapses {  // parameters.hpp:69-99
    double w_e = 5.0, w_i = -5.0, w_n2a = 0.2, w_a2n = 1.0;
    double d_e = 1.0, d_i = 1.0, d_a2n = 1.0;
    int sic_interval = 1;
    bool stp_enabled = true;
    double U = 0.5, tau_rec = 800.0, tau_fac = 0.0;
};

struct Analysis {  // parameters.hpp:160-166
    bool enabled = true;
    double ca_threshold = -1.0;
    double merge_gap = 2000.0;
    double hist_window = 2000.0;
    double hist_shift = 400.0;
};

struct ModelConfig {
    uint32_t N_astro = 100, N_exc = 400, N_inh = 100;
    double dt = 0.1;
    int substeps = 1;
    double pre_sim_time = 0.0;
    double sim_time = 0.0;
    uint64_t seed = 1;

    asg_astro_params astro{};
    double astro_Ca_tot = 2.0, astro_IP3_0 = 0.16, astro_tau_IP3 = 7142.0,
           astro_delta_IP3 = 0.0002;
    asg_neuron_params neuron_exc{}, neuron_inh{};
    Synapses syn{};
    Connectivity conn{};
    asg_input_params input_astro{}, input_exc{}, input_inh{};
    RandomizeSpec randomize_astro{}, randomize_neuron{};
    Analysis analysis{};

    bool record_spikes = true, record_astro = true, record_neuron = false;
    int record_every = 10;
    uint32_t record_astro_max = 100, record_neuron_max = 50;
    std::string output_dir = "results";

    ModelConfig();

    // TimeGrid, types.hpp:30-43.
    int64_t pre_steps() const { return static_cast<int64_t>(pre_sim_time / dt + 0.5); }
    int64_t sim_steps() const { return static_cast<int64_t>(sim_time / dt + 0.5); }
};

// Command-line style overrides applied after the file (main.cpp:214-233), and
// "section.key=value;..." assignments applied to the JSON before mapping.
struct Overrides {
    int64_t seed = -1;
    double sim_time = -1.0;
    double pre_time = -1.0;
    int64_t n_astro = -1;
    std::string assignments;
    std::string output_dir;
};

// Throws std::runtime_error on a malformed file, an unknown key or an invalid
// value, with the reference's messages.
ModelConfig load_config(const std::string& path, const Overrides& ov);
ModelConfig load_config_text(const std::string& text, const std::string& name, const Overrides& ov);

}  // namespace asg_host

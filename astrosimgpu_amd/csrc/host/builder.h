// Network construction on the host: per-cell randomised parameters, the
// astrocyte pools, the pairwise primary connections with tripartite
// recruitment, and the four CSR sets the device step consumes
// (src/network.cpp:56-228, astrocyte.cpp:28-53, neuron.cpp:23-75).
//
// Every draw comes from one counter generator (seed, stream 0x5EED) consumed
// in the reference's order, so a given seed yields the reference's network
// bit for bit. Transcendentals here are the host C library's, as in the
// reference build.
#pragma once

#include <cstdint>
#include <vector>

#include "asg.h"
#include "config.h"

namespace asg_host {

struct NetworkStats {  // network.hpp:47-55
    uint32_t n_astro = 0, n_exc = 0, n_inh = 0;
    uint32_t n_primary_exc = 0, n_primary_inh = 0;
    uint32_t n_neuron_to_astro = 0, n_astro_to_neuron = 0;
};

struct Csr {
    std::vector<uint32_t> row_start, target;
};

// A built network: owns every array an asg_net_desc points into.
struct Model {
    ModelConfig cfg;
    NetworkStats stats;
    std::vector<double> V_reset, b;                           // per neuron
    std::vector<double> Ca_tot, IP3_0, tau_IP3, delta_IP3;    // per astrocyte
    Csr exc, inh, n2a, a2n;
    int32_t d_exc = 1, d_inh = 1, d_a2n = 1;
    asg_net_desc desc{};                                      // points into the vectors above

    void refresh_desc(int32_t precision);
};

// The reference's host-side counter generator (rng.hpp:56-143): stateful
// wrapper over the stateless splitmix draw, with Box-Muller pair caching.
class CounterRng {
public:
    CounterRng(uint64_t seed, uint64_t stream) : seed_(seed), stream_(stream) {}
    double uniform();
    double normal();
    double normal(double mean, double stddev) { return mean + stddev * normal(); }
    double normal_redraw(double mean, double stddev, double lo, double hi, int max_tries = 1000);

private:
    uint64_t seed_, stream_, counter_ = 0;
    bool has_spare_ = false;
    double spare_ = 0.0;
};

int delay_to_steps(double delay_ms, double dt);  // network.cpp:32-35

void build_model(Model& m);

}  // namespace asg_host

// The simulation step of the tripartite neuron-astrocyte network on gfx950.
//
// One engine steps a set of independent networks (one for asg_net_*, many for
// asg_batch_*) through "windows" of W consecutive steps, where W is the
// smallest synaptic delay in steps (10 in every shipped configuration). No
// signal emitted inside a window can arrive inside it, so within a window
// every cell is independent of every other cell and the per-cell state stays
// in registers for W steps. A window is five launches (six when FP64 and
// FP32 networks share a batch):
//
//   k_drive    one lane per neuron, W steps: the Poisson event counts and
//              noise current of the external drive, from counter streams.
//   k_cells    one lane per neuron / astrocyte, W steps each, one kernel per
//              precision: arrivals from the ring slots, drive, RK4, spike
//              test (a wavefront ballot writes one 64-bit spike word per 64
//              neurons and step), astrocyte update and its SIC factor.
//   k_compact  one workgroup per (network, step): prefix scan of the spike
//              words into a source-ordered spike list, one work item per
//              spike.
//   k_deliver  lanes per (neuron, step): the SIC each neuron receives from
//              that step's exchange (skipped when no astrocyte of the network
//              has a nonzero factor), and each spiking source's
//              Tsodyks-Markram update in time order and raster entries;
//              workgroups per spike: set the spike's bit in the per-target
//              64-bit mask of every synapse of its rows.
//   k_sum      one lane per (target, step): adds the marked spikes' weights
//              into the ring slot in ascending spike (= source) order.
//
// Marking + ordered sum replace the reference's sequential scatter
// (network.cpp:249-279) with an order-independent OR followed by a sum in the
// reference's floating-point order (source ascending, then CSR order),
// without atomics on doubles.
//
// Exact reformulations (each gives the reference's bits, see below):
//  * STP state is kept per source, not per synapse: every synapse of a row is
//    initialised alike (network.cpp:43-52) and updated at the same instants
//    with the same t_now (network.cpp:255-277), so all of a source's exc and
//    n2a synapses carry identical (x, u, t_last).
//  * SIC delivery is a gather at arrival time: target t adds w_a2n*factor(a)
//    once per (a, t) synapse, astrocytes ascending, which is exactly the
//    sequence of += the reference's scatter (network.cpp:281-299) applies to
//    ring_sic[slot][t]; the pending flag of a slot is a function of the step.
//  * n2a rows keep (astrocyte, multiplicity) pairs sorted by astrocyte; a
//    source's duplicates hit the same accumulator consecutively either way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "asg.h"
#include "amath.h"
#include "cell_math.h"
#include "cell_math_f32.h"
#include "raster_order.h"
#include "rng.h"

namespace asg {

// --------------------------------------------------------------- errors ----

namespace {
thread_local std::string g_last_error;

bool abort_on_error() {
    static const bool value = [] {
        const char* v = std::getenv("ASG_ABORT_ON_ERROR");
        return v != nullptr && v[0] == '1';
    }();
    return value;
}

asg_status fail(asg_status code, const std::string& what) {
    g_last_error = what;
    if (abort_on_error()) {
        std::fprintf(stderr, "asg: %s\n", what.c_str());
        std::abort();
    }
    return code;
}

struct HipError {
    hipError_t err;
    std::string what;
};

void hip_check(hipError_t err, const char* call, const char* file, int line) {
    if (err != hipSuccess) {
        char buf[512];
        std::snprintf(buf, sizeof(buf), "HIP error at %s:%d during %s: %s", file, line, call,
                      hipGetErrorString(err));
        throw HipError{err, buf};
    }
}
#define HIP_CHECK(call) ::asg::hip_check((call), #call, __FILE__, __LINE__)

template <typename F>
asg_status guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(e.err == hipErrorOutOfMemory ? ASG_ERR_NOMEM : ASG_ERR_DEVICE, e.what);
    } catch (const std::bad_alloc&) {
        return fail(ASG_ERR_NOMEM, "host allocation failed");
    }
}
}  // namespace

// --------------------------------------------------------------- layout ----

constexpr int kCellBlock = 64;     // one wavefront per block of the per-cell kernels
constexpr int kWindowCap = 16;     // longest window (bounds scratch memory)
constexpr double kFixScale = 4294967296.0;        // 2^32: fixed-point ring units
constexpr double kFixUnit = 1.0 / 4294967296.0;   // 2^-32
constexpr int64_t kSicWaves = 4096; // SIC wavefronts that fill the device (256 CUs x 16)
constexpr int64_t kGroupWaves = 16384;  // per-step wavefronts beyond which steps are grouped
constexpr int kDeliverWaves = 512;  // delivery wavefronts sharing out the spiking blocks
constexpr int kSicWorkers = 0;      // FP32 SIC wavefronts (0: one per FP32 neuron block)
constexpr int kMaxLanes = 3;        // FP32 lanes (network groups on streams of their own)
constexpr int kLanes = 2;           // lanes by default
constexpr int kMaxTeams = 1024;     // teams of the persistent FP32 launch
constexpr size_t kMaskBudget = size_t(256) << 20;  // delivery masks per network
constexpr size_t kMinChunks = 64;

struct InputK {
    double poiss_rate, poiss_weight, lambda, limit;
    double noise_std;
    int32_t independent;
    int32_t steps_per_noise;
};

struct AstroPopK {
    AstroK k;
    double SIC_th, SIC_scale;
    InputK in;
};

// Device-resident description of one network: sizes, constants, and the
// addresses of its state, connectivity and scratch arrays.
struct DevNet {
    uint32_t n_exc, n_inh, n_neur, n_astro;
    uint32_t words;        // ceil(n_neur / 64)
    uint32_t cw;           // 64-spike chunks of window spike indices the masks hold
    // Row strides (elements) of the [rows][neurons] and [rows][astrocytes]
    // arrays -- rings, drive, SIC factors, delivery masks, SIC gather lists:
    // n_neur and n_astro rounded up to 64, so that a wavefront's 64 lanes
    // read whole aligned lines.
    uint32_t ldn, lda;
    int32_t S;             // ring slots: longest delay + longest window
    int32_t Wmax;          // longest window this network's scratch holds
    int32_t d_exc, d_inh, d_n2a, d_a2n;
    int32_t sic_interval, substeps;
    int32_t win_W;         // length of the current window
    int32_t overflow;      // bit 0: raster buffer full; bit 1: a window had more
                           // spikes than the delivery masks hold (cw * 64)
    int64_t step;          // first step of the next window
    int64_t target;        // step this call stops at
    int64_t win_T0;        // first step of the current window
    int32_t win_sic_send;  // first window step of the current window that sends
                           // a SIC exchange (window offset; may be >= its length)
    int32_t win_pad;
    uint64_t seed;
    double dt, h, half_h, h_sixth;
    NeuronK pk[2];
    InputK in[2];
    AstroPopK ap;
    int32_t precision;     // ASG_FP64 or ASG_FP32
    NeuronKf pkf[2];
    AstroKf akf;
    int32_t stp_on, fac_on;
    double U, one_minus_U, tau_rec, tau_fac;
    double w_exc, w_inh, w_n2a, w_a2n;
    // neuron state
    double *V, *w, *gex, *dgex, *gin, *dgin, *Isic;
    int32_t* refr;
    const double *V_reset, *b;
    // astrocyte state and parameters
    double *Ca, *IP3, *hh;
    const double *Ca_tot, *IP3_0, *tau_IP3, *delta_IP3;
    // rings [S][targets]: exc, inh and SIC per neuron, n2a per astrocyte
    double *ring_exc, *ring_inh, *ring_sic, *ring_astro;
    double* fac;           // [Wmax][n_astro]: SIC factor of each step of the window
    // External drive of the window's steps, written by k_drive before the
    // cell pass: Poisson event counts [Wmax][n_neur] (when a population has
    // a rate) and the noise current [Wmax][n_neur] (when one has noise).
    int32_t* drive_n;
    double* drive_noise;
    float* fdrive_noise;
    // FP32 networks keep their conductances, SIC current, parameters and
    // rings here instead (their V, w, Ca, IP3 and h are the FP64 arrays)
    float *fgex, *fdgex, *fgin, *fdgin, *fIsic;
    float *fCa_tot, *fIP3_0, *finv_tau, *fdelta;
    float *fring_sic, *ffac;
    uint32_t* ffac_nz;     // [n_astro]: bit k = entry k of the astrocyte's ffac line is nonzero
    // FP32 networks' exc, inh and n2a rings in fixed point (value * 2^32):
    // delivered by 64-bit atomic adds, whose integer sums do not depend on
    // their order
    int64_t *qring_exc, *qring_inh, *qring_astro;
    double *stp_x, *stp_u, *stp_tl;
    // connectivity
    const uint32_t *exc_row, *exc_tgt, *inh_row, *inh_tgt;
    const uint32_t *n2a_row, *n2a_tgt, *n2a_mult, *n2a_raw_row;
    // SIC gather, column-major: pair q of target t at [q * n_neur + t];
    // multiplicity 0 marks padding (a target's pairs come first, astrocytes
    // ascending)
    uint32_t sic_width;
    const uint32_t *sic_src, *sic_mult;
    int64_t sic_pairs;
    // window scratch
    uint64_t* spk_bits;    // [Wmax][words]
    uint32_t* spk_woff;    // [Wmax][words]: spikes before each word in its step
    uint32_t* spk_list;    // [Wmax][n_neur]
    uint32_t* spk_n;       // [Wmax][2]: total, excitatory
    double* spk_c;         // [Wmax][n_neur][2]: exc-or-inh weight, n2a weight
    // delivery masks, one bit per (window spike, target): [cw][n_neur], [cw][n_astro]
    uint64_t *mask_n, *mask_a;
    // raster
    int64_t* r_step;
    uint32_t* r_src;
    int64_t r_cap, r_head;
    // astrocyte recording: record r = (step - rec_start) / rec_every lands in
    // ring slot r % rec_cap as [Ca x rec_n, IP3 x rec_n]
    int64_t rec_start, rec_cap;
    int32_t rec_every;
    uint32_t rec_n;
    double* rec_buf;
    // counters
    int64_t c_steps, c_spikes, c_syn, c_sic;
    uint32_t sic_live;     // bit k: some astrocyte has a nonzero SIC factor at
                           // window step k (set by k_cells, cleared at the close)
    uint32_t sic_live_p[2];  // FP32 networks: the same, one word per window
    // FP32 networks whose inhibitory neurons begin inside a 64-neuron block
    // (n_exc % 64 != 0): that block's cell pass is two wavefronts, one per
    // population. The inhibitory one's spike words ([Wmax]) are spk_extra,
    // which delivery ORs into the block's; split_state (per window parity)
    // collects bit 0 / bit 1 as each part finishes and bit 2 if either
    // spiked, so that exactly one of them lists the block.
    uint64_t* spk_extra;
    uint32_t split_state[2];
                             // parity (the close empties the other parity's)
    int32_t open;          // 1 from a window's cell pass to its close
    // Window clock: the step-periodic phases of the window that starts at
    // `step` (set_clock), kept beside it so that the cell pass divides
    // nothing: every wavefront would otherwise repeat the same 64-bit
    // divisions.
    int32_t clk_row;          // step % S
    int32_t clk_sic_arrive;   // first_hit(step, d_a2n, sic_interval)
    int32_t clk_sic_send;     // first_hit(step, 0, sic_interval)
    int32_t clk_nz_next[2];   // per neuron population: first noise-interval
                              // change after the window's first step
    int32_t clk_anz_next;     // the same for the astrocytes' noise
    int64_t clk_anz_idx;      // step / astrocyte steps_per_noise
    int64_t clk_rec_next, clk_rec_slot;  // astrocyte recording cursor
    // bit T mod (hist_mask + 1): the SIC exchange arriving at step T had a
    // nonzero factor, i.e. its ring_sic row was written; 0 means every
    // neuron's SIC is 0 (a power-of-two ring of at least S + 32 bits)
    uint32_t* sic_hist;
    uint32_t hist_mask;
};

// Bit k: the SIC exchange arriving at step T0 + k had a nonzero factor, so
// its ring row holds the sums (bits of the next 32 steps; two loads).
__device__ __forceinline__ uint32_t live_arrivals(const DevNet& n, int64_t T0) {
    const uint32_t b = static_cast<uint32_t>(T0) & n.hist_mask;
    const uint32_t w = b >> 5, words = (n.hist_mask + 1) >> 5;
    const uint64_t lo = n.sic_hist[w], hi = n.sic_hist[(w + 1) & (words - 1)];
    return static_cast<uint32_t>(((hi << 32) | lo) >> (b & 31));
}

// Recording cadence of one window: `next` is the window offset of the next
// recorded step, `slot` its ring slot (from the window clock); advance()
// moves to the one after.
struct RecCursor {
    int64_t next, slot;
    __device__ __forceinline__ explicit RecCursor(const DevNet& n)
        : next(n.clk_rec_next), slot(n.clk_rec_slot) {}
    __device__ __forceinline__ void advance(const DevNet& n) {
        next += n.rec_every;
        slot = slot + 1 == n.rec_cap ? 0 : slot + 1;
    }
};

__device__ inline void record_astro(const DevNet& n, uint32_t a, int64_t slot, double Ca,
                                    double IP3) {
    if (a < n.rec_n) {
        double* rec = n.rec_buf + static_cast<size_t>(slot) * 2 * n.rec_n;
        rec[a] = Ca;
        rec[n.rec_n + a] = IP3;
    }
}

// Bit k of `bits`: this lane's SIC factor at window step k is nonzero. One
// atomic per wavefront publishes the OR over its active lanes into the
// network's live word of the window. Returns true on the one lane (of one
// wavefront per network and window) whose atomic turned it from 0 to
// nonzero.
__device__ inline bool mark_sic_live(uint32_t* live_word, uint32_t bits, int W) {
    uint32_t any = 0;
    for (int k = 0; k < W; ++k) {
        if (__ballot((bits >> k) & 1u) != 0) {
            any |= 1u << k;
        }
    }
    const uint64_t active = __ballot(true);
    if (any != 0 &&
        (threadIdx.x & 63) == static_cast<uint32_t>(__ffsll(static_cast<long long>(active)) - 1)) {
        return atomicOr(live_word, any) == 0u;
    }
    return false;
}

// Per-window state of the FP32 post pass (k_post), one per window parity:
// the count of SIC items -- neuron blocks of networks whose SIC exchange is
// live in the window -- the cell pass listed (emptied by the post pass of
// the window before).
struct PostState {
    uint32_t sic_items;
    uint32_t pad;
};

struct CellBlock {
    int32_t net;
    int32_t kind;   // 0 neurons, 1 astrocytes
    uint32_t first;
    uint32_t lead;  // 1 for exactly one block per network
};

__device__ inline int window_len(const DevNet& n, int32_t Wg) {
    const int64_t rem = n.target - n.step;
    return rem <= 0 ? 0 : (rem < Wg ? static_cast<int>(rem) : Wg);
}

// The windows carry every step-periodic quantity as a counter from T0 instead
// of dividing by the period each step: the ring row T % S, and the window
// offset of the next step T >= origin with (T - origin) % period == 0.
__device__ __forceinline__ int ring_next(int r, int S) { return r + 1 == S ? 0 : r + 1; }

__device__ __forceinline__ int64_t first_hit(int64_t T0, int64_t origin, int64_t period) {
    const int64_t d = T0 - origin;
    if (d <= 0) {
        return -d;
    }
    const int64_t r = d % period;
    return r == 0 ? 0 : period - r;
}

// The window clock of the window that starts at step T (the DevNet clk_
// fields). Written by one lane per network, where `step` changes: the
// close of a window and the per-call launch that sets the run targets
// (after which any host-side change of the step or of the recording
// settings is taken up).
__device__ inline void set_clock(DevNet& n, int64_t T) {
    n.clk_row = static_cast<int32_t>(T % n.S);
    n.clk_sic_arrive = static_cast<int32_t>(first_hit(T, n.d_a2n, n.sic_interval));
    n.clk_sic_send = static_cast<int32_t>(first_hit(T, 0, n.sic_interval));
    for (int p = 0; p < 2; ++p) {
        const int64_t period = n.in[p].noise_std > 0.0 ? n.in[p].steps_per_noise : 1;
        n.clk_nz_next[p] = static_cast<int32_t>(first_hit(T + 1, 0, period) + 1);
    }
    const int64_t spn = n.ap.in.steps_per_noise;
    n.clk_anz_idx = spn > 0 ? T / spn : 0;
    n.clk_anz_next = spn > 0 ? static_cast<int32_t>(first_hit(T + 1, 0, spn) + 1) : INT32_MAX;
    n.clk_rec_next = INT64_MAX;
    n.clk_rec_slot = 0;
    if (n.rec_n != 0) {
        const int64_t d = T - n.rec_start;
        const int64_t q = d <= 0 ? 0 : (d + n.rec_every - 1) / n.rec_every;
        n.clk_rec_next = n.rec_start + q * n.rec_every - T;
        n.clk_rec_slot = q % n.rec_cap;
    }
}

__device__ inline void add_i64(int64_t* p, int64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// --------------------------------------------------------------- kernels ---

__device__ __forceinline__ uint64_t neuron_window(const DevNet& n, uint32_t i, int64_t T0, int W,
                                                  const uint64_t* exp_tab) {
    // Everything the loop reads is copied into registers once; the rings are
    // reached through restrict pointers so their stores cannot force reloads.
    const uint32_t N = n.n_neur;
    const size_t LN = n.ldn;  // row stride of the [rows][neurons] arrays
    const bool valid = i < N;
    const int pi = (i < n.n_exc) ? 0 : 1;
    const NeuronK p = n.pk[pi];
    const InputK in = n.in[pi];
    const int S = n.S;
    const int sic_interval = n.sic_interval;
    const int substeps = n.substeps;
    const double hs = n.h, half_h = n.half_h, h_sixth = n.h_sixth;
    const uint32_t words = n.words;
    double* __restrict__ ring_exc = n.ring_exc;
    double* __restrict__ ring_inh = n.ring_inh;
    const double* __restrict__ ring_sic = n.ring_sic;
    uint64_t* __restrict__ spk_bits = n.spk_bits;
    const int32_t* __restrict__ drive_n = n.drive_n;
    const double* __restrict__ drive_noise = n.drive_noise;

    int row = n.clk_row;  // ring row of the current step
    int64_t sic_next = n.clk_sic_arrive;  // first_hit(T0, d_a2n, sic_interval)
    const uint32_t live_in = live_arrivals(n, T0);
    const bool sic_live_n = live_in & 1u;
    double V = 0, w = 0, gex = 0, dgex = 0, gin = 0, dgin = 0, Isic = 0, Vr = 0, b = 0;
    int refr = 0;
    double ex_n = 0, ih_n = 0, sic_n = 0;  // ring values of the next step
    int ev_n = 0;                           // drive of the next step
    double noise_n = 0.0;
    const bool has_poiss = in.poiss_rate > 0.0, has_noise = in.noise_std > 0.0;
    // The drive holds one noise value per noise interval of the window:
    // value nz_j from window step nz_next - steps_per_noise on.
    const int64_t nz_period = has_noise ? in.steps_per_noise : 1;
    int64_t nz_next = n.clk_nz_next[pi];  // first_hit(T0 + 1, 0, nz_period) + 1
    uint32_t nz_j = 0;
    if (valid) {
        V = n.V[i];
        w = n.w[i];
        gex = n.gex[i];
        dgex = n.dgex[i];
        gin = n.gin[i];
        dgin = n.dgin[i];
        Isic = n.Isic[i];
        refr = n.refr[i];
        Vr = n.V_reset[i];
        b = n.b[i];
        const size_t s0 = static_cast<size_t>(row) * LN + i;
        ex_n = ring_exc[s0];
        ih_n = ring_inh[s0];
        if (sic_live_n) sic_n = ring_sic[s0];
        if (has_poiss) ev_n = drive_n[i];
        if (has_noise) noise_n = drive_noise[i];  // the window's first noise interval
    }
    const uint32_t lane = threadIdx.x & 63;
    uint64_t spiked_any = 0;  // OR of the window's spike words (uniform)
    for (int k = 0; k < W; ++k) {
        bool spiked = false;
        const int row_n = ring_next(row, S);
        const bool sic_step = k == sic_next;
        if (sic_step) {
            sic_next += sic_interval;
        }
        const bool next_live = (live_in >> (k + 1)) & 1u;
        if (valid) {
            const size_t slot = static_cast<size_t>(row) * LN + i;
            const double ex = ex_n, ih = ih_n, sic = sic_n;
            const int ev = ev_n;
            const double noise = noise_n;
            if (k + 1 < W) {
                const size_t nx = static_cast<size_t>(row_n) * LN + i;
                ex_n = ring_exc[nx];
                ih_n = ring_inh[nx];
                sic_n = next_live ? ring_sic[nx] : 0.0;
                const size_t dk1 = static_cast<size_t>(k + 1) * LN + i;
                if (has_poiss) ev_n = drive_n[dk1];
                if (has_noise && k + 1 == nz_next) {
                    // a new noise interval: its value is the next one stored
                    nz_next += nz_period;
                    ++nz_j;
                    noise_n = drive_noise[static_cast<size_t>(nz_j) * LN + i];
                }
            }
            double exc_in = 0.0, inh_in = 0.0;
            // apply_arrivals, network.cpp:301-327
            if (ex != 0.0) {
                if (ex > 0.0) {
                    exc_in += ex;
                } else if (ex < 0.0) {
                    inh_in -= ex;
                }
                ring_exc[slot] = 0.0;
            }
            if (ih != 0.0) {
                const double wgt = -ih;
                if (wgt > 0.0) {
                    exc_in += wgt;
                } else if (wgt < 0.0) {
                    inh_in -= wgt;
                }
                ring_inh[slot] = 0.0;
            }
            // A new SIC arrives exactly when an exchange was sent d_a2n
            // steps earlier; otherwise the last value is held. An exchange
            // without a nonzero factor (sic_hist 0) is 0 for every neuron
            // and its ring row is not read.
            if (sic_step) {
                Isic = sic;
            }
            // NeuronPopulation::update, neuron.cpp:143-170 (draws from k_drive)
            if (has_poiss && ev > 0) {
                exc_in += in.poiss_weight * static_cast<double>(ev);
            }
            double I_ext = Isic;
            if (has_noise) {
                I_ext += noise;
            }
            if (exc_in != 0.0) {
                dgex += exc_in * p.psc_ex;
            }
            if (inh_in != 0.0) {
                dgin += inh_in * p.psc_in;
            }
            spiked = adex_step(p, Vr, b, I_ext, hs, half_h, h_sixth, substeps, V, w, gex, dgex,
                               gin, dgin, refr, exp_tab);
        }
        const uint64_t mask = __ballot(spiked);
        if (lane == 0 && (i & ~63u) < N) {
            spk_bits[static_cast<size_t>(k) * words + (i >> 6)] = mask;
        }
        spiked_any |= mask;
        row = row_n;
    }
    if (valid) {
        n.V[i] = V;
        n.w[i] = w;
        n.gex[i] = gex;
        n.dgex[i] = dgex;
        n.gin[i] = gin;
        n.dgin[i] = dgin;
        n.Isic[i] = Isic;
        n.refr[i] = refr;
    }
    return spiked_any;
}

__device__ __forceinline__ void astro_window(const DevNet& n, uint32_t a, int64_t T0, int W) {
    const uint32_t NA = n.n_astro;
    const size_t LA = n.lda;  // row stride of the [rows][astrocytes] arrays
    if (a >= NA) {
        return;
    }
    const AstroPopK ap = n.ap;
    const int S = n.S;
    const int sic_interval = n.sic_interval;
    const int substeps = n.substeps;
    const uint64_t seed = n.seed;
    const double hs = n.h, half_h = n.half_h, h_sixth = n.h_sixth;
    double* __restrict__ ring_astro = n.ring_astro;
    double* __restrict__ fac = n.fac;
    double Ca = n.Ca[a], IP3 = n.IP3[a], hh = n.hh[a];
    const double Ca_tot = n.Ca_tot[a], IP3_0 = n.IP3_0[a], tau = n.tau_IP3[a],
                 delta = n.delta_IP3[a];
    int row = n.clk_row;  // ring row of the current step
    double v_n = ring_astro[static_cast<size_t>(row) * LA + a];
    // noise index T / steps_per_noise, redrawn at the window start and at
    // every multiple of steps_per_noise
    const int64_t spn = ap.in.steps_per_noise;
    int64_t noise_idx = n.clk_anz_idx;    // T0 / spn
    int64_t noise_next = n.clk_anz_next;  // first_hit(T0 + 1, 0, spn) + 1
    int64_t sic_next = n.clk_sic_send;    // first_hit(T0, 0, sic_interval)
    RecCursor rec(n);
    double noise = 0.0;
    uint32_t live_bits = 0;
    // Poisson stream T * kStreamStride + a: its key advances by a constant
    uint64_t pkey = rng_key(seed ^ kSaltAstroPoisson, static_cast<uint64_t>(T0) * kStreamStride + a);
    for (int k = 0; k < W; ++k, pkey += kStreamStride * kStreamMul) {
        const int row_n = ring_next(row, S);
        const size_t slot = static_cast<size_t>(row) * LA + a;
        const double v = v_n;
        if (k + 1 < W) {
            v_n = ring_astro[static_cast<size_t>(row_n) * LA + a];
        }
        double ip3_in = 0.0;
        if (v != 0.0) {
            ip3_in += v;
            ring_astro[slot] = 0.0;
        }
        // drive_astrocytes, network.cpp:340-355
        if (ap.in.poiss_rate > 0.0) {
            const int ev = rng_poisson_fresh_keyed(pkey, ap.in.lambda, ap.in.limit);
            if (ev > 0) {
                ip3_in += ap.in.poiss_weight * static_cast<double>(ev);
            }
        }
        // AstrocytePopulation::update noise, astrocyte.cpp:239-248, 310-314
        if (ap.in.noise_std > 0.0) {
            bool redraw = k == 0;
            if (k == noise_next) {
                ++noise_idx;
                noise_next += spn;
                redraw = true;
            }
            if (redraw) {
                const uint64_t nseed = seed ^ kSaltAstroNoise;
                const uint64_t stream = ap.in.independent
                                            ? static_cast<uint64_t>(noise_idx) * kStreamStride + a
                                            : static_cast<uint64_t>(noise_idx);
                noise = ap.in.noise_std * rng_normal_clamped(nseed, stream);
            }
        }
        astro_step(ap.k, Ca_tot, IP3_0, tau, delta, ip3_in, noise, hs, half_h, h_sixth, substeps,
                   Ca, IP3, hh);
        if (k == rec.next) {
            record_astro(n, a, rec.slot, Ca, IP3);
            rec.advance(n);
        }
        if (k == sic_next) {
            sic_next += sic_interval;
            const double f = sic_factor(Ca, ap.SIC_th, ap.SIC_scale);
            fac[static_cast<size_t>(k) * LA + a] = f;
            live_bits |= static_cast<uint32_t>(f != 0.0) << k;
        }
        row = row_n;
    }
    n.Ca[a] = Ca;
    n.IP3[a] = IP3;
    n.hh[a] = hh;
    mark_sic_live(const_cast<uint32_t*>(&n.sic_live), live_bits, W);
}

// ----------------------------------------------------- FP32 cell windows ---

// Global-memory access of the FP32 windows: a base address uniform across
// the wavefront (scalar registers) plus a 32-bit byte offset per lane, so
// each load or store is one instruction with no 64-bit address arithmetic.
// A network's arrays fit (plan_of refuses an FP32 network whose rings,
// drive or factor lines exceed 4 GiB).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> uniform_global(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (gptr<T>)(reinterpret_cast<T*>((static_cast<uint64_t>(hi) << 32) | lo));
}

template <typename T>
__device__ __forceinline__ gptr<T> at_byte(gptr<T> base, uint32_t byte_off) {
    return (gptr<T>)((__attribute__((address_space(1))) char*)base + byte_off);
}

__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi =
        __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// The W steps of the lanes of population pi (0 excitatory, 1 inhibitory;
// uniform across the wavefront, so that the population's constants stay in
// scalar registers) of one neuron block. Lane k ORs window step k's spike
// word into spk_word (the caller stores the words). Returns the OR of the
// window's spike words.
__device__ __forceinline__ uint64_t neuron_window_f32(const DevNet& n, uint32_t i, int64_t T0,
                                                      int W, int pi, uint64_t& spk_word) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    const uint32_t N = n.n_neur;
    const bool valid = i < N && (i < n.n_exc) == (pi == 0);
    const NeuronKf pf = n.pkf[pi];
    const NeuronK pd = n.pk[pi];
    const InputK in = n.in[pi];
    // the window's uniform quantities, in scalar registers
    const uint32_t LN4 = __builtin_amdgcn_readfirstlane(n.ldn) * 4u;  // a [.][neurons] row
    const uint32_t LN8 = 2u * LN4;
    const uint32_t S8 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(n.S)) * LN8;
    const int sic_interval = __builtin_amdgcn_readfirstlane(n.sic_interval);
    const int substeps = __builtin_amdgcn_readfirstlane(n.substeps);
    const double hs = n.h, half_h = n.half_h, h_sixth = n.h_sixth;
    gptr<int64_t> ring_exc = uniform_global(n.qring_exc);
    gptr<int64_t> ring_inh = uniform_global(n.qring_inh);
    gptr<const float> ring_sic = uniform_global(static_cast<const float*>(n.fring_sic));
    gptr<const int32_t> drive_n = uniform_global(static_cast<const int32_t*>(n.drive_n));
    gptr<const float> drive_noise = uniform_global(static_cast<const float*>(n.fdrive_noise));

    // byte offsets: of the current step's ring row, and of this lane's cell
    uint32_t row8 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(n.clk_row)) * LN8;
    const uint32_t i8 = i * 8u, i4 = i * 4u;
    int64_t sic_next = uniform_i64(n.clk_sic_arrive);  // first_hit(T0, d_a2n, sic_interval)
    const uint32_t live_in = __builtin_amdgcn_readfirstlane(live_arrivals(n, T0));
    const bool sic_live_n = live_in & 1u;
    // FP64 registers; the conductances are rounded to FP32 at the end of
    // every step, V and w stay FP64
    double V = 0, w = 0, gex = 0, dgex = 0, gin = 0, dgin = 0, Isic = 0, Vr = 0, b = 0;
    int refr = 0;
    int64_t ex_n = 0, ih_n = 0;
    float sic_n = 0;
    int ev_n = 0;
    float noise_n = 0.0f;
    const bool has_poiss = in.poiss_rate > 0.0, has_noise = in.noise_std > 0.0;
    // The drive holds one noise value per noise interval of the window:
    // value nz_j from window step nz_next - steps_per_noise on.
    const int64_t nz_period = has_noise ? in.steps_per_noise : 1;
    int64_t nz_next = n.clk_nz_next[pi];  // first_hit(T0 + 1, 0, nz_period) + 1
    uint32_t nz_off = i4;                 // byte offset of the current noise value
    if (valid) {
        V = n.V[i];
        w = n.w[i];
        gex = n.fgex[i];
        dgex = n.fdgex[i];
        gin = n.fgin[i];
        dgin = n.fdgin[i];
        Isic = n.fIsic[i];
        refr = n.refr[i];
        Vr = n.V_reset[i];
        b = n.b[i];
        ex_n = *at_byte(ring_exc, row8 + i8);
        ih_n = *at_byte(ring_inh, row8 + i8);
        if (sic_live_n) sic_n = *at_byte(ring_sic, row8 / 2u + i4);
        if (has_poiss) ev_n = *at_byte(drive_n, i4);
        if (has_noise) noise_n = *at_byte(drive_noise, i4);  // the window's first noise interval
    }
    const uint32_t lane = threadIdx.x & 63;
    uint64_t spiked_any = 0;  // OR of the window's spike words (uniform)
    // No store inside the step loop: on this GPU loads and stores complete
    // in order, so a store issued in a step (conditional, so uncounted by
    // the compiler) would make the next step's prefetched loads wait for it.
    // The ring slots that held input are cleared after the loop (bit k of
    // clr_ex / clr_ih: window step k's slot), and lane k keeps step k's spike
    // word for the caller's store after the loop. Nothing else touches this
    // window's ring rows before the post pass.
    const uint32_t row0_8 = row8;
    uint32_t clr_ex = 0, clr_ih = 0;
    for (int k = 0; k < W; ++k) {
        bool spiked = false;
        const uint32_t rown8 = row8 + LN8 == S8 ? 0u : row8 + LN8;
        const bool sic_step = k == sic_next;
        if (sic_step) {
            sic_next += sic_interval;
        }
        const bool next_live = (live_in >> (k + 1)) & 1u;
        if (valid) {
            const int64_t ex = ex_n, ih = ih_n;
            const float sic = sic_n;
            const int ev = ev_n;
            const float noise = noise_n;
            if (k + 1 < W) {
                const uint32_t nx8 = rown8 + i8;
                ex_n = *at_byte(ring_exc, nx8);
                ih_n = *at_byte(ring_inh, nx8);
                sic_n = next_live ? *at_byte(ring_sic, rown8 / 2u + i4) : 0.0f;
                if (has_poiss) ev_n = *at_byte(drive_n, static_cast<uint32_t>(k + 1) * LN4 + i4);
                if (has_noise && k + 1 == nz_next) {
                    // a new noise interval: its value is the next one stored
                    nz_next += nz_period;
                    nz_off += LN4;
                    noise_n = *at_byte(drive_noise, nz_off);
                }
            }
            double exc_in = 0.0, inh_in = 0.0;
            if (ex != 0) {
                const double v = static_cast<double>(ex) * kFixUnit;
                if (v > 0.0) {
                    exc_in += v;
                } else {
                    inh_in -= v;
                }
                clr_ex |= 1u << k;
            }
            if (ih != 0) {
                const double v = static_cast<double>(ih) * kFixUnit;
                if (v < 0.0) {
                    exc_in -= v;
                } else {
                    inh_in += v;
                }
                clr_ih |= 1u << k;
            }
            if (sic_step) {
                Isic = sic;
            }
            // Same counter streams as the FP64 step (k_drive): identical
            // Poisson events, the noise current from FP32 draws.
            if (has_poiss && ev > 0) {
                exc_in += in.poiss_weight * static_cast<double>(ev);
            }
            double I_ext = Isic;
            if (has_noise) {
                I_ext += noise;
            }
            if (exc_in != 0.0) {
                dgex += exc_in * pd.psc_ex;
            }
            if (inh_in != 0.0) {
                dgin += inh_in * pd.psc_in;
            }
            for (int sidx = 0; sidx < substeps; ++sidx) {
                const bool sp = adex_substep_reg(pd, pf, Vr, b, I_ext, hs, half_h, h_sixth, V, w,
                                                 gex, dgex, gin, dgin, refr);
                spiked = spiked || sp;
            }
            // the conductances are FP32 between steps, so results do not
            // depend on where windows (call lengths, the batch's shortest
            // delay) fall
            gex = round_f32(gex);
            dgex = round_f32(dgex);
            gin = round_f32(gin);
            dgin = round_f32(dgin);
        }
        const uint64_t mask = __ballot(spiked);
        spk_word = lane == static_cast<uint32_t>(k) ? (spk_word | mask) : spk_word;
        spiked_any |= mask;
        row8 = rown8;
    }
    for (uint32_t m = clr_ex; m != 0; m &= m - 1) {
        const uint32_t r8 = row0_8 + static_cast<uint32_t>(__ffs(m) - 1) * LN8;
        *at_byte(ring_exc, (r8 >= S8 ? r8 - S8 : r8) + i8) = 0;
    }
    for (uint32_t m = clr_ih; m != 0; m &= m - 1) {
        const uint32_t r8 = row0_8 + static_cast<uint32_t>(__ffs(m) - 1) * LN8;
        *at_byte(ring_inh, (r8 >= S8 ? r8 - S8 : r8) + i8) = 0;
    }
    if (valid) {
        n.V[i] = V;
        n.w[i] = w;
        n.fgex[i] = static_cast<float>(gex);
        n.fdgex[i] = static_cast<float>(dgex);
        n.fgin[i] = static_cast<float>(gin);
        n.fdgin[i] = static_cast<float>(dgin);
        n.fIsic[i] = static_cast<float>(Isic);
        n.refr[i] = refr;
    }
    return spiked_any;
}

// Returns true on the lane whose update made the network's SIC exchange
// live for the window (see mark_sic_live).
__device__ __forceinline__ bool astro_window_f32(const DevNet& n, uint32_t a, int64_t T0, int W,
                                                 int par) {
    // FP32 path: mul-add pairs may fuse (it has no host build to match)
    #pragma clang fp contract(fast)
    const uint32_t NA = n.n_astro;
    if (a >= NA) {
        return false;
    }
    const AstroKf kf = n.akf;
    const InputK ain = n.ap.in;
    const double SIC_th = n.ap.SIC_th, SIC_scale = n.ap.SIC_scale;
    const int sic_interval = __builtin_amdgcn_readfirstlane(n.sic_interval);
    const int substeps = __builtin_amdgcn_readfirstlane(n.substeps);
    const uint64_t seed = n.seed;
    const float hs = static_cast<float>(n.h), half_h = static_cast<float>(n.half_h),
                h_sixth = static_cast<float>(n.h_sixth);
    // uniform bases and 32-bit byte offsets (see at_byte)
    gptr<int64_t> ring_astro = uniform_global(n.qring_astro);
    gptr<float> fac = uniform_global(n.ffac);
    const uint32_t LA8 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(n.lda)) * 8u;
    const uint32_t S8 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(n.S)) * LA8;
    const uint32_t a8 = a * 8u, line = a * 64u;  // this astrocyte's ring column, factor line
    // bit k: entry k of this astrocyte's factor line is nonzero
    uint32_t* __restrict__ fac_nz = n.ffac_nz;
    const uint32_t nz_prev = fac_nz[a];
    uint32_t sent = 0;  // bits of the window's steps that send an exchange
    double Ca = n.Ca[a], IP3 = n.IP3[a], hh = n.hh[a];
    const float Ca_tot = n.fCa_tot[a], IP3_0 = n.fIP3_0[a], inv_tau = n.finv_tau[a],
                delta = n.fdelta[a];
    // byte offset of the ring row of the current step
    uint32_t row8 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(n.clk_row)) * LA8;
    int64_t v_n = *at_byte(ring_astro, row8 + a8);
    // noise index T / steps_per_noise, redrawn at the window start and at
    // every multiple of steps_per_noise
    const int64_t spn = ain.steps_per_noise;
    int64_t noise_idx = uniform_i64(n.clk_anz_idx);    // T0 / spn
    int64_t noise_next = uniform_i64(n.clk_anz_next);  // first_hit(T0 + 1, 0, spn) + 1
    int64_t sic_next = uniform_i64(n.clk_sic_send);    // first_hit(T0, 0, sic_interval)
    // ring slots that held input, cleared after the loop (see neuron_window_f32)
    const uint32_t row0_8 = row8;
    uint32_t clr = 0;
    RecCursor rec(n);
    float noise = 0.0f;
    uint32_t live_bits = 0;
    // Poisson stream T * kStreamStride + a: its key advances by a constant
    uint64_t pkey = rng_key(seed ^ kSaltAstroPoisson, static_cast<uint64_t>(T0) * kStreamStride + a);
    for (int k = 0; k < W; ++k, pkey += kStreamStride * kStreamMul) {
        const uint32_t rown8 = row8 + LA8 == S8 ? 0u : row8 + LA8;
        const int64_t v = v_n;
        if (k + 1 < W) {
            v_n = *at_byte(ring_astro, rown8 + a8);
        }
        float ip3_in = 0.0f;
        if (v != 0) {
            ip3_in += static_cast<float>(static_cast<double>(v) * kFixUnit);
            clr |= 1u << k;
        }
        if (ain.poiss_rate > 0.0) {
            // Knuth's loop only: an FP32 network with lambda > 30 here (the
            // reference's rounded normal) is refused at creation
            // (check_fp32_limits), so the normal's FP64 log and cos need no
            // registers in this kernel
            const int ev = rng_poisson_knuth_keyed(pkey, ain.limit);
            if (ev > 0) {
                ip3_in += static_cast<float>(ain.poiss_weight * static_cast<double>(ev));
            }
        }
        if (ain.noise_std > 0.0) {
            bool redraw = k == 0;
            if (k == noise_next) {
                ++noise_idx;
                noise_next += spn;
                redraw = true;
            }
            if (redraw) {
                const uint64_t stream = ain.independent
                                            ? static_cast<uint64_t>(noise_idx) * kStreamStride + a
                                            : static_cast<uint64_t>(noise_idx);
                noise = static_cast<float>(ain.noise_std) *
                        rng_normal_clamped_f32(rng_key(seed ^ kSaltAstroNoise, stream));
            }
        }
        astro_step_f32(kf, Ca_tot, IP3_0, inv_tau, delta, ip3_in, noise, hs, half_h, h_sixth,
                       substeps, Ca, IP3, hh);
        if (k == rec.next) {
            record_astro(n, a, rec.slot, Ca, IP3);
            rec.advance(n);
        }
        if (k == sic_next) {
            sic_next += sic_interval;
            const float f = sic_factor_mixed(Ca, SIC_th, SIC_scale);
            sent |= 1u << k;
            // a zero over a zero is not stored: the line is written where
            // it changes (an astrocyte below threshold writes nothing)
            if (f != 0.0f || ((nz_prev >> k) & 1u)) {
                *at_byte(fac, line + static_cast<uint32_t>(k) * 4u) = f;
            }
            live_bits |= static_cast<uint32_t>(f != 0.0f) << k;
        }
        row8 = rown8;
    }
    for (uint32_t m = clr; m != 0; m &= m - 1) {
        const uint32_t r8 = row0_8 + static_cast<uint32_t>(__ffs(m) - 1) * LA8;
        *at_byte(ring_astro, (r8 >= S8 ? r8 - S8 : r8) + a8) = 0;
    }
    n.Ca[a] = Ca;
    n.IP3[a] = IP3;
    n.hh[a] = hh;
    const uint32_t nz = (nz_prev & ~sent) | live_bits;
    if (nz != nz_prev) {
        fac_nz[a] = nz;
    }
    return mark_sic_live(const_cast<uint32_t*>(&n.sic_live_p[par]), live_bits, W);
}

// FP32 SIC of one neuron for every window step that sends an exchange:
// sum over (astrocyte, multiplicity) of m * w_a2n * factor, compensated.
// The factors are stored astrocyte-major, 16 to a 64-byte line (ffac[a * 16
// + k] = factor of window step k), so one pair costs one line for all the
// window's steps rather than one line per step. Entries of steps that send
// are current (the astrocyte pass stores every new factor except a zero over
// a zero); the others are not read.
template <int GM>
__device__ __forceinline__ void deliver_sic_f32(const DevNet& n, uint32_t t, int G,
                                                uint32_t sic_live) {
    static_assert(GM <= 16, "one line of factors per astrocyte");
    const size_t LN = n.ldn;
    const int64_t T0 = n.win_T0;
    uint32_t send = 0;  // bit j: window step j sends an exchange
    for (int j = n.win_sic_send; j < G; j += n.sic_interval) {
        send |= 1u << j;
    }
    const uint32_t live = send & sic_live;
    float acc[GM], comp[GM];
#pragma unroll
    for (int j = 0; j < GM; ++j) {
        acc[j] = 0.0f;
        comp[j] = 0.0f;
    }
    if (live != 0) {
        const uint32_t* __restrict__ src = n.sic_src;
        const uint32_t* __restrict__ mult = n.sic_mult;
        const float4* __restrict__ fac = reinterpret_cast<const float4*>(n.ffac);
        const float w_a2n = static_cast<float>(n.w_a2n);
        // the 16-byte parts of a line that hold live steps (uniform: one
        // network per wavefront)
        const uint32_t parts = (live & 0xFu ? 1u : 0u) | (live & 0xF0u ? 2u : 0u) |
                               (live & 0xF00u ? 4u : 0u) | (live & 0xF000u ? 8u : 0u);
        for (uint32_t q = 0; q < n.sic_width; ++q) {
            const uint32_t m = mult[static_cast<size_t>(q) * LN + t];
            if (m == 0) {
                break;
            }
            const uint32_t a = src[static_cast<size_t>(q) * LN + t];
            const float4* row = fac + static_cast<size_t>(a) * 4;
            float v[16];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (4 * c < GM && ((parts >> c) & 1u)) {
                    x = row[c];
                }
                v[4 * c] = x.x;
                v[4 * c + 1] = x.y;
                v[4 * c + 2] = x.z;
                v[4 * c + 3] = x.w;
            }
            const float fm = static_cast<float>(m);
#pragma unroll
            for (int j = 0; j < GM; ++j) {
                if ((live >> j) & 1u) {
                    const float fa = v[j];
                    if (fa != 0.0f) {
                        comp_add(acc[j], comp[j], fm * (w_a2n * fa));
                    }
                }
            }
        }
    }
    // A live exchange writes every neuron's slot; the others are neither
    // written nor read (sic_hist).
    const int S = n.S;
    const int row0 = static_cast<int>((T0 + n.d_a2n) % S);
#pragma unroll
    for (int j = 0; j < GM; ++j) {
        const float v = acc[j] + comp[j];
        if ((live >> j) & 1u) {
            const int row = row0 + j < S ? row0 + j : row0 + j - S;
            n.fring_sic[static_cast<size_t>(row) * LN + t] = v;
        }
    }
}

// Exclusive prefix sum over the 64 lanes of a wavefront (no barriers).
__device__ inline uint32_t wave_exclusive_scan(uint32_t v, uint32_t& total) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) {
            x += y;
        }
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// Closes the current window whose SIC exchanges were live at the window
// steps of the bits of `sic_live`: their liveness at their arrival steps (a
// step that sends no exchange has a 0 bit), counters and the next window's
// first step.
__device__ inline void close_window_with(DevNet& n, uint32_t sic_live) {
    const int W = n.win_W;
    const int64_t T0 = n.win_T0;
    // the W liveness bits go to ring bits b0 .. b0 + W - 1 (W <= 16 < 32):
    // at most two words, each read and written once
    const uint32_t b0 = static_cast<uint32_t>(T0 + n.d_a2n) & n.hist_mask;
    const uint32_t w0 = b0 >> 5, sh = b0 & 31, words = (n.hist_mask + 1) >> 5;
    const uint64_t span = ((1ULL << W) - 1ULL) << sh;
    const uint64_t live = static_cast<uint64_t>(sic_live & ((1u << W) - 1u)) << sh;
    uint32_t* hist = n.sic_hist;
    hist[w0] = (hist[w0] & ~static_cast<uint32_t>(span)) | static_cast<uint32_t>(live);
    if ((span >> 32) != 0) {
        const uint32_t w1 = (w0 + 1) & (words - 1);
        hist[w1] = (hist[w1] & ~static_cast<uint32_t>(span >> 32)) | static_cast<uint32_t>(live >> 32);
    }
    n.c_steps += W;
    n.step = T0 + W;
    set_clock(n, T0 + W);
    n.open = 0;
}

__device__ inline void close_window_state(DevNet& n) {
    close_window_with(n, n.sic_live);
    n.sic_live = 0;
}

// The close of an FP64 window of `total` spikes: its raster entries (written
// at r_head by the delivery pass) become part of the raster if they all fit.
__device__ inline void close_window(DevNet& n, uint32_t total) {
    if (n.r_head + total <= n.r_cap) {
        n.r_head += total;
    }
    close_window_state(n);
}

// Window pass 0: the external drive of every neuron for the window's steps,
// from the reference's counter streams -- Poisson event counts of the
// background input (neuron.cpp:143-151) and the Gaussian noise current held
// over noise_dt (neuron.cpp:152-166) -- written per (step, neuron) for the
// cell pass. One wavefront per neuron block.
//
// Knuth's Poisson draw (rng.hpp:107-124) multiplies uniforms until the
// product falls to exp(-lambda); at lambda ~ 0.2 most draws end after one
// uniform, but a wavefront looping per lane would run as long as its slowest
// lane. Here every (step, neuron) draws its first uniform together; the ones
// still above the limit go to a queue in LDS (product, step, lane, draws so
// far) and the wavefront works through the queue 64 entries at a time,
// compacting survivors in place, until it is empty. Every uniform is the
// same (seed, stream, counter) as in the sequential loop, so the counts are
// identical.
//
// drive_block computes the drive of the window after the last one opened
// (win_T0 + win_W is the step that window's close sets) for the 64 neurons
// of block cb, with the dynamic LDS ring drive_lds (64 * Wg products and
// keys). All 64 lanes of a one-wavefront workgroup call it together.
__device__ __forceinline__ void drive_block(DevNet& n, const CellBlock& cb, int32_t Wg,
                                            double* drive_lds) {
    double* q_prod = drive_lds;                                             // [64 * Wg]
    uint32_t* q_key = reinterpret_cast<uint32_t*>(drive_lds + 64 * Wg);     // [64 * Wg]
    const int64_t T0 = n.win_T0 + n.win_W;
    const int64_t rem = n.target - T0;
    const int W = rem <= 0 ? 0 : (rem < Wg ? static_cast<int>(rem) : Wg);
    if (W == 0) {
        return;
    }
    const uint32_t N = n.n_neur;
    const size_t LN = n.ldn;
    const uint32_t lane = threadIdx.x;
    const uint32_t i = cb.first + lane;
    const bool valid = i < N;
    const int pi = (i < n.n_exc) ? 0 : 1;
    const InputK in = n.in[pi];
    const uint64_t seed = n.seed;
    if (n.in[0].poiss_rate > 0.0 || n.in[1].poiss_rate > 0.0) {
        int32_t* __restrict__ out = n.drive_n;
        const uint64_t pseed = seed ^ kSaltNeuronPoisson;
        if (n.in[0].lambda > 30.0 || n.in[1].lambda > 30.0) {
            // the normal approximation above lambda = 30: one lane per draw
            for (int k = 0; valid && k < W; ++k) {
                out[static_cast<size_t>(k) * LN + i] =
                    rng_poisson_fresh(pseed, static_cast<uint64_t>(T0 + k) * kStreamStride + i,
                                      in.lambda, in.limit);
            }
        } else {
            const bool mine = valid && in.lambda > 0.0;
            uint32_t qn = 0;
            // stream (T0 + k) * kStreamStride + i: its key advances by a
            // constant per step
            constexpr uint64_t kStepKey = kStreamStride * kStreamMul;
            uint64_t key = rng_key(pseed, static_cast<uint64_t>(T0) * kStreamStride + i);
            for (int k = 0; k < W; ++k, key += kStepKey) {
                bool pend = false;
                double p = 0.0;
                if (mine) {
                    p = rng_uniform_keyed(key, 0);
                    pend = p > in.limit;
                }
                if (valid && !pend) {
                    out[static_cast<size_t>(k) * LN + i] = 0;
                }
                const uint64_t b = __ballot(pend);
                if (pend) {
                    const uint32_t pos =
                        qn + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0));
                    q_prod[pos] = p;
                    q_key[pos] = (static_cast<uint32_t>(k) << 6) | lane;  // draws so far: 0 extra
                }
                qn += static_cast<uint32_t>(__popcll(b));
            }
            __syncthreads();
            // The queue is a ring of 64 * Wg entries worked from its head 64
            // at a time, survivors appended at its tail, so a draw's next
            // uniform need not wait for the rest of its round. Each (step,
            // lane) has at most one entry, so the live span tail - head
            // never exceeds the ring.
            const uint32_t cap = 64u * static_cast<uint32_t>(Wg);
            // the stream key of (step kk, lane l) is base + (kk * stride + l) * phi
            const uint64_t base = rng_key(pseed, static_cast<uint64_t>(T0) * kStreamStride + cb.first);
            const double limit_e = n.in[0].limit, limit_i = n.in[1].limit;
            const uint32_t n_exc = n.n_exc;
            uint32_t head = 0, tail = qn;
            while (head < tail) {
                const uint32_t take = tail - head < 64u ? tail - head : 64u;
                uint32_t at = head + lane;
                at = at >= cap ? at - cap : at;
                bool pend = false;
                double p = 0.0;
                uint32_t key = 0;
                if (lane < take) {
                    p = q_prod[at];
                    key = q_key[at];
                    const uint32_t l = key & 63u, kk = (key >> 6) & 15u, cnt = (key >> 10) + 1;
                    const uint32_t ii = cb.first + l;
                    const double limit = ii < n_exc ? limit_e : limit_i;
                    const uint32_t soff = __umul24(kk, static_cast<uint32_t>(kStreamStride)) + l;
                    const uint64_t z = base + static_cast<uint64_t>(soff) * kStreamMul +
                                       static_cast<uint64_t>(cnt) * kCounterMul;
                    p *= static_cast<double>(rng_bits_keyed(z, 0) >> 11) * (1.0 / 9007199254740992.0);
                    if (p > limit) {
                        pend = true;
                        key = (cnt << 10) | (kk << 6) | l;
                    } else {
                        out[static_cast<size_t>(kk) * LN + ii] = static_cast<int32_t>(cnt);
                    }
                }
                head += take;
                head = head >= cap ? head - cap : head;
                tail = tail >= cap ? tail - cap : tail;
                if (tail < head) {
                    tail += cap;  // keep tail - head the live span
                }
                __syncthreads();  // this chunk is read before survivors overwrite it
                const uint64_t b = __ballot(pend);
                if (pend) {
                    uint32_t pos = tail + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0));
                    pos = pos >= cap ? pos - cap : pos;
                    q_prod[pos] = p;
                    q_key[pos] = key;
                }
                tail += static_cast<uint32_t>(__popcll(b));
                __syncthreads();
            }
        }
    }
    if (valid && in.noise_std > 0.0) {
        // One draw per noise interval (neuron.cpp:152-166), held across it:
        // value j of the window's noise drive holds its j-th interval.
        const uint64_t salt_shared = pi == 0 ? kSaltNoiseSharedExc : kSaltNoiseSharedInh;
        const uint64_t nseed = in.independent ? seed ^ kSaltNeuronNoise : seed ^ salt_shared;
        const int64_t spn = in.steps_per_noise;
        const int64_t idx = T0 / spn;
        int64_t next = first_hit(T0 + 1, 0, spn) + 1;  // offset of the next multiple of spn
        // stream idx * kNoiseStride + i (independent) or idx (shared): the
        // key advances by a constant per noise index
        const uint64_t idx_key = (in.independent ? kNoiseStride : 1ULL) * kStreamMul;
        uint64_t key = rng_key(nseed, in.independent ? static_cast<uint64_t>(idx) * kNoiseStride + i
                                                     : static_cast<uint64_t>(idx));
        if (n.precision == ASG_FP32) {
            // FP32 path: the same uniforms through the FP32 Box-Muller.
            const float sd = static_cast<float>(in.noise_std);
            n.fdrive_noise[i] = sd * rng_normal_fresh_f32(key);
            for (size_t j = 1; next < W; next += spn, ++j) {
                key += idx_key;
                n.fdrive_noise[j * LN + i] = sd * rng_normal_fresh_f32(key);
            }
        } else {
            uint64_t ctr = 0;
            n.drive_noise[i] = 0.0 + in.noise_std * rng_normal_fresh_keyed(key, ctr);
            for (size_t j = 1; next < W; next += spn, ++j) {
                key += idx_key;
                ctr = 0;
                n.drive_noise[j * LN + i] = 0.0 + in.noise_std * rng_normal_fresh_keyed(key, ctr);
            }
        }
    }
}

// Window pass 0 as a launch of its own: the drive of the next window of
// every network whose neuron blocks are [nblocks, nblocks + n_items), one
// wavefront per block; block 0 first empties the delivery list count
// work_n (when given) for the window's cell pass.
__global__ __launch_bounds__(64) void k_drive(DevNet* nets, const CellBlock* nblocks,
                                              int32_t n_items, int32_t Wg, int32_t* work_n) {
    extern __shared__ double drive_lds[];
    if (blockIdx.x == 0 && threadIdx.x == 0 && work_n != nullptr) {
        *work_n = 0;
    }
    const int item = static_cast<int>(blockIdx.x);
    if (item >= n_items) {
        return;
    }
    const CellBlock cb = nblocks[item];
    drive_block(nets[cb.net], cb, Wg, drive_lds);
}

// Window pass 1: every cell of every network advances W steps. One kernel per
// precision, so neither carries the other's registers; neurons and
// astrocytes share the launch, so the short astrocyte blocks overlap the
// neuron blocks.
// WPB wavefronts (64-cell blocks) per workgroup, a launch-schedule knob.
// A neuron block that spiked in the window appends itself to its
// precision's delivery list `work`, counted in *work_n. An FP32 network
// whose SIC exchange becomes live in the window lists its neuron blocks
// (nbr[2 * slot + 1] of them from index nbr[2 * slot] of the FP32 neuron
// blocks) as the post pass's SIC items in sic_item. FP32 windows alternate
// between two sets of these lists and live words by parity `par` (the host
// passes the set of the window's parity), so that the post pass can empty
// the other set while it reads this one.
// The cell pass of one 64-cell block, run by one whole wavefront (the
// block is uniform across its lanes).
template <int P>
__device__ __forceinline__ void cells_item(DevNet* nets, const CellBlock& cb, int32_t Wg,
                                           CellBlock* work, int32_t* work_n, int32_t par,
                                           PostState* post, int32_t* sic_item,
                                           const int32_t* nbr, const uint64_t* exp_tab) {
    const uint32_t lane = threadIdx.x & 63;
    DevNet& n = nets[cb.net];
    const int W = window_len(n, Wg);
    const int64_t T0 = n.step;
    if (cb.lead && lane == 0) {
        n.win_T0 = T0;
        n.win_W = W;
        n.win_sic_send = n.clk_sic_send;
        n.open = 1;
    }
    if (W == 0) {
        return;
    }
    const uint32_t i = cb.first + lane;
    if constexpr (P == ASG_FP32) {
        if (cb.kind != 1) {
            // one population per wavefront (its constants in scalar
            // registers): kind 0 the population of the block's first neuron,
            // kind 2 the inhibitory part of the block where they begin
            const uint32_t n_exc = __builtin_amdgcn_readfirstlane(n.n_exc);
            const int pi = cb.kind == 2 ? 1 : (cb.first < n_exc ? 0 : 1);
            uint64_t spk_word = 0;
            const uint64_t any = neuron_window_f32(n, i, T0, W, pi, spk_word);
            const bool split = cb.first < n_exc && cb.first + 63 >= n_exc && n_exc < n.n_neur;
            if (lane < static_cast<uint32_t>(W) && cb.first < n.n_neur) {
                if (cb.kind == 2) {
                    n.spk_extra[lane] = spk_word;
                } else {
                    n.spk_bits[static_cast<size_t>(lane) * n.words + (cb.first >> 6)] = spk_word;
                }
            }
            bool list = any != 0;
            if (split) {
                // the part that finishes second lists the block if either
                // part spiked; the atomic's old value is all it reads
                uint32_t old = 0;
                if (lane == 0) {
                    old = atomicOr(&n.split_state[par],
                                   (cb.kind == 2 ? 2u : 1u) | (any != 0 ? 4u : 0u));
                }
                old = __builtin_amdgcn_readfirstlane(old);
                list = (old & (cb.kind == 2 ? 1u : 2u)) != 0 && ((old & 4u) != 0 || any != 0);
            }
            if (list && lane == 0) {
                work[atomicAdd(work_n, 1)] = CellBlock{cb.net, 0, cb.first, cb.lead};
            }
        } else {
            // the wavefront whose update made the exchange live lists the
            // network's blocks, its lanes together
            const uint64_t made = __ballot(astro_window_f32(n, i, T0, W, par));
            if (made != 0) {
                const int32_t nb = nbr[2 * cb.net + 1], first = nbr[2 * cb.net];
                uint32_t at = 0;
                if (lane == static_cast<uint32_t>(__ffsll(static_cast<long long>(made)) - 1)) {
                    at = atomicAdd(&post->sic_items, static_cast<uint32_t>(nb));
                }
                at = __builtin_amdgcn_readlane(at, __ffsll(static_cast<long long>(made)) - 1);
                for (int32_t q = static_cast<int32_t>(lane); q < nb; q += 64) {
                    sic_item[at + q] = first + q;
                }
            }
        }
    } else {
        if (cb.kind == 0) {
            if (neuron_window(n, i, T0, W, exp_tab) != 0 && lane == 0) {
                work[atomicAdd(work_n, 1)] = cb;
            }
        } else {
            astro_window(n, i, T0, W);
        }
    }
}

// Window pass 1: every cell of every network advances W steps. One kernel per
// precision, so neither carries the other's registers; neurons and
// astrocytes share the launch, so the short astrocyte blocks overlap the
// neuron blocks.
// WPB wavefronts (64-cell blocks) per workgroup, a launch-schedule knob.
// A neuron block that spiked in the window appends itself to its
// precision's delivery list `work`, counted in *work_n. An FP32 network
// whose SIC exchange becomes live in the window lists its neuron blocks
// (nbr[2 * slot + 1] of them from index nbr[2 * slot] of the FP32 neuron
// blocks) as the post pass's SIC items in sic_item. FP32 windows alternate
// between two sets of these lists and live words by parity `par` (the host
// passes the set of the window's parity), so that the post pass can empty
// the other set while it reads this one.
// The FP32 instantiation is held to five wavefronts per SIMD (96 VGPRs,
// which its neuron and astrocyte paths fit without spilling).
template <int P, int WPB>
__global__ __launch_bounds__(kCellBlock * WPB)
__attribute__((amdgpu_waves_per_eu(P == ASG_FP32 ? 5 : 1))) void k_cells(DevNet* nets, const CellBlock* blocks,
                                                            int32_t count, int32_t Wg,
                                                            CellBlock* work, int32_t* work_n,
                                                            int32_t par, PostState* post,
                                                            int32_t* sic_item, const int32_t* nbr) {
    // The wavefront's block index is uniform across its lanes: said so
    // explicitly, so the block and network descriptors stay in scalar
    // registers.
    const int32_t b = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(blockIdx.x) * WPB +
                                                     static_cast<int32_t>(threadIdx.x >> 6));
    if constexpr (P == ASG_FP64) {
        // exp's table in LDS: its lookups sit on the RK4 stages' dependency
        // chain, and an LDS read returns sooner than the global table's
        __shared__ uint64_t exp_tab[256];
        for (uint32_t j = threadIdx.x; j < 256; j += blockDim.x) {
            exp_tab[j] = ASG_EXP_TAB[j];
        }
        __syncthreads();
        if (b >= count) {
            return;
        }
        cells_item<P>(nets, blocks[b], Wg, work, work_n, par, post, sic_item, nbr, exp_tab);
    } else {
        if (b >= count) {
            return;
        }
        cells_item<P>(nets, blocks[b], Wg, work, work_n, par, post, sic_item, nbr, nullptr);
    }
}


// Window pass 2: one wavefront per (network, step) turns that step's spike
// words into the source-ordered spike list (the reference sorts each step's
// spikes by source, neuron.cpp:243-248) and each word's spike offset.
__global__ __launch_bounds__(64) void k_compact(DevNet* nets, const int32_t* active) {
    DevNet& n = nets[active[blockIdx.x]];
    const int k = static_cast<int>(blockIdx.y);
    if (k >= n.win_W) {
        return;
    }
    const uint32_t lane = threadIdx.x;
    const uint32_t words = n.words;
    const uint32_t per = (words + 63) / 64;
    const uint32_t w_lo = min(words, lane * per), w_hi = min(words, w_lo + per);
    const uint64_t* bits = n.spk_bits + static_cast<size_t>(k) * words;
    uint32_t cnt = 0, cntE = 0;
    for (uint32_t q = w_lo; q < w_hi; ++q) {
        const uint64_t m = bits[q];
        cnt += __popcll(m);
        const uint32_t base = q * 64;
        if (base + 64 <= n.n_exc) {
            cntE += __popcll(m);
        } else if (base < n.n_exc) {
            cntE += __popcll(m & ((1ULL << (n.n_exc - base)) - 1ULL));
        }
    }
    uint32_t total, totalE;
    uint32_t off = wave_exclusive_scan(cnt, total);
    wave_exclusive_scan(cntE, totalE);
    uint32_t* list = n.spk_list + static_cast<size_t>(k) * n.n_neur;
    uint32_t* woff = n.spk_woff + static_cast<size_t>(k) * words;
    for (uint32_t q = w_lo; q < w_hi; ++q) {
        woff[q] = off;
        uint64_t m = bits[q];
        while (m) {
            const uint32_t bit = static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
            list[off++] = q * 64 + bit;
            m &= m - 1;
        }
    }
    if (lane == 0) {
        n.spk_n[2 * k] = total;
        n.spk_n[2 * k + 1] = totalE;
        if (total > 0) {
            add_i64(&n.c_spikes, total);
        }
    }
}

// The SIC neuron t receives from the exchange sent at window step k
// (network.cpp:281-299): w_a2n * factor(a) added once per (a, t) synapse,
// astrocytes ascending, written to ring_sic at the arrival slot.
template <int GM>
__device__ void deliver_sic(const DevNet& n, uint32_t t, int g0, int G) {
    const size_t LN = n.ldn;
    const int64_t T0 = n.win_T0;
    uint32_t send = 0;  // bit j: window step g0 + j sends an exchange
    for (int64_t j = first_hit(T0 + g0, 0, n.sic_interval); j < G; j += n.sic_interval) {
        send |= 1u << j;
    }
    // Steps at which no astrocyte of the network has a nonzero factor have
    // an empty sum.
    const uint32_t live = send & (n.sic_live >> g0);
    double acc[GM];
#pragma unroll
    for (int j = 0; j < GM; ++j) {
        acc[j] = 0.0;
    }
    if (live != 0) {
        const uint32_t* __restrict__ src = n.sic_src;
        const uint32_t* __restrict__ mult = n.sic_mult;
        const double* __restrict__ fac = n.fac + static_cast<size_t>(g0) * n.lda;
        const size_t LA = n.lda;
        const double w_a2n = n.w_a2n;
        for (uint32_t q = 0; q < n.sic_width; ++q) {
            const uint32_t m = mult[static_cast<size_t>(q) * LN + t];
            if (m == 0) {
                break;
            }
            const uint32_t a = src[static_cast<size_t>(q) * LN + t];
            double v[GM];
            uint32_t nz = 0;
#pragma unroll
            for (int j = 0; j < GM; ++j) {
                v[j] = 0.0;
                if ((live >> j) & 1u) {
                    const double fa = fac[j * LA + a];
                    v[j] = w_a2n * fa;
                    nz |= static_cast<uint32_t>(fa != 0.0) << j;
                }
            }
            // the m additions of one (astrocyte, target) pair; the chains of
            // the group's steps are independent and interleave
            if (nz != 0) {
                for (uint32_t r = 0; r < m; ++r) {
#pragma unroll
                    for (int j = 0; j < GM; ++j) {
                        acc[j] = ((nz >> j) & 1u) ? acc[j] + v[j] : acc[j];
                    }
                }
            }
        }
    }
    // A live exchange writes every neuron's slot; the others are neither
    // written nor read (sic_hist).
    const int S = n.S;
    const int row0 = static_cast<int>((T0 + g0 + n.d_a2n) % S);
#pragma unroll
    for (int j = 0; j < GM; ++j) {
        if ((live >> j) & 1u) {
            const int row = row0 + j < S ? row0 + j : row0 + j - S;
            n.ring_sic[static_cast<size_t>(row) * LN + t] = acc[j];
        }
    }
}

// Set bit (j mod 64) of chunk j / 64 in the mask of every target of one row,
// the whole wavefront together, with 8 target loads in flight per lane.
__device__ inline void mark_row(const uint32_t* __restrict__ tgt, uint32_t q0, uint32_t q1,
                                uint64_t* base, uint64_t bit, uint32_t lane) {
    constexpr int U = 8;
    for (uint32_t q = q0 + lane; q < q1; q += 64 * U) {
        uint32_t tt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t qq = q + static_cast<uint32_t>(u) * 64;
            tt[u] = qq < q1 ? tgt[qq] : UINT32_MAX;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (tt[u] != UINT32_MAX) {
                atomicOr(reinterpret_cast<unsigned long long*>(base + tt[u]), bit);
            }
        }
    }
}

// Lane k's value of v; k must be uniform across the wavefront.
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int k) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), k));
}
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int k) {
    return static_cast<uint64_t>(lane_u32(static_cast<uint32_t>(v), k)) |
           (static_cast<uint64_t>(lane_u32(static_cast<uint32_t>(v >> 32), k)) << 32);
}

// What one wavefront's 64 neurons spiked at window steps [k0, k0 + nk),
// nk <= 64, loaded in one round: lane kk holds step k0 + kk's spike word of
// the block, the spikes of the step before that word, and the step's
// total and excitatory spike counts.
struct StepWords {
    uint64_t word;
    uint32_t off, tot, nE;
    __device__ __forceinline__ StepWords(const DevNet& n, uint32_t word_i, uint32_t lane, int k0,
                                         int nk) {
        word = 0;
        off = tot = nE = 0;
        if (lane < static_cast<uint32_t>(nk)) {
            const size_t q = static_cast<size_t>(k0 + static_cast<int>(lane)) * n.words + word_i;
            word = n.spk_bits[q];
            off = n.spk_woff[q];
            tot = n.spk_n[2 * (k0 + static_cast<int>(lane))];
            nE = n.spk_n[2 * (k0 + static_cast<int>(lane)) + 1];
        }
    }
};

// The spikes one wavefront's 64 neurons emitted at window steps [k0, k1):
// each spike's bit -- its window index, the spikes of the earlier steps plus
// its index in its step's list -- is set in the window mask of every target
// of its rows. OR is order-independent, so wavefronts may mark in any order.
// The window's words and the spiking neurons' row bounds are loaded once.
__device__ void mark_steps(const DevNet& n, uint32_t first, uint32_t lane, int k0, int k1) {
    const uint32_t NE = n.n_exc;
    const StepWords sw(n, first >> 6, lane, 0, n.win_W);
    uint32_t window_total = 0;
    const uint32_t before_k = wave_exclusive_scan(sw.tot, window_total);
    uint64_t any = 0;
    for (int kk = k0; kk < k1; ++kk) {
        any |= lane_u64(sw.word, kk);
    }
    if (any == 0) {
        return;
    }
    const uint32_t s_lane = first + lane;
    uint32_t r0 = 0, r1 = 0, a0 = 0, a1 = 0;
    if ((any >> lane) & 1ULL) {
        if (s_lane < NE) {
            r0 = n.exc_row[s_lane];
            r1 = n.exc_row[s_lane + 1];
            a0 = n.n2a_row[s_lane];
            a1 = n.n2a_row[s_lane + 1];
        } else {
            r0 = n.inh_row[s_lane - NE];
            r1 = n.inh_row[s_lane - NE + 1];
        }
    }
    const uint32_t cap = 64 * n.cw;
    for (int k = k0; k < k1; ++k) {
        const uint64_t word = lane_u64(sw.word, k);
        if (word == 0) {
            continue;
        }
        const uint32_t base = lane_u32(before_k, k) + lane_u32(sw.off, k);
        uint64_t m = word;
        while (m) {
            const int L = __ffsll(static_cast<long long>(m)) - 1;
            m &= m - 1;
            const uint32_t jw = base + static_cast<uint32_t>(__popcll(word & ((1ULL << L) - 1ULL)));
            if (jw >= cap) {
                break;  // beyond the delivery capacity; deliver_sources flags it
            }
            const uint64_t bit = 1ULL << (jw & 63);
            const size_t chunk = jw >> 6;
            const uint32_t q0 = lane_u32(r0, L), q1 = lane_u32(r1, L);
            mark_row(first + static_cast<uint32_t>(L) < NE ? n.exc_tgt : n.inh_tgt, q0, q1,
                     n.mask_n + chunk * n.ldn, bit, lane);
            if (first + static_cast<uint32_t>(L) < NE) {
                mark_row(n.n2a_tgt, lane_u32(a0, L), lane_u32(a1, L), n.mask_a + chunk * n.lda,
                         bit, lane);
            }
        }
    }
}

// For the 64 neurons of one wavefront, over the whole window in time order:
// each spike's Tsodyks-Markram update (network.cpp:230-247), the weights its
// rows deliver and its raster entry. The window's spike words come in one
// round; a spiking neuron's plasticity state and row lengths are loaded
// once and its state stored once.
__device__ void deliver_sources(DevNet& n, uint32_t first, uint32_t lane) {
    const uint32_t N = n.n_neur;
    const uint32_t t = first + lane;
    const bool valid = t < N;
    const int W = n.win_W;
    const int64_t T0 = n.win_T0;
    const int64_t r_head = n.r_head;
    const StepWords sw(n, first >> 6, lane, 0, W);
    uint32_t window_total = 0;
    const uint32_t before_k = wave_exclusive_scan(sw.tot, window_total);  // spikes of earlier steps
    const bool room = r_head + static_cast<int64_t>(window_total) <= n.r_cap;
    uint32_t mine = 0;  // bit k: this lane's neuron spiked at window step k
    for (int k2 = 0; k2 < W; ++k2) {
        mine |= static_cast<uint32_t>((lane_u64(sw.word, k2) >> lane) & 1ULL) << k2;
    }
    const bool exc = t < n.n_exc;
    const bool stp = n.stp_on != 0;
    int64_t row_syn = 0;
    double x = 0.0, u = 0.0, tl = 0.0;
    if (valid && mine != 0) {
        if (exc) {
            row_syn = (n.exc_row[t + 1] - n.exc_row[t]) + (n.n2a_raw_row[t + 1] - n.n2a_raw_row[t]);
        } else {
            row_syn = n.inh_row[t - n.n_exc + 1] - n.inh_row[t - n.n_exc];
        }
        if (stp) {
            x = n.stp_x[t];
            u = n.stp_u[t];
            tl = n.stp_tl[t];
        }
    }
    const uint64_t below = (1ULL << lane) - 1ULL;
    int64_t syn = 0;
    for (int k2 = 0; k2 < W; ++k2) {
        const uint64_t word = lane_u64(sw.word, k2);
        if (word == 0) {
            continue;
        }
        const uint32_t off = lane_u32(sw.off, k2);
        const uint32_t before = lane_u32(before_k, k2);
        if ((mine >> k2) & 1u) {
            const uint32_t j = off + static_cast<uint32_t>(__popcll(word & below));
            double c0 = exc ? n.w_exc : n.w_inh;
            double c1 = n.w_n2a;
            const double t_now = static_cast<double>(T0 + k2) * n.dt;
            if (stp) {
                const double dts = t_now - tl;
                const double x_decay = asg_exp(-dts / n.tau_rec);
                const double u_decay = n.fac_on ? asg_exp(-dts / n.tau_fac) : 0.0;
                x = 1.0 + (x - x * u - 1.0) * x_decay;
                u = n.U + u * n.one_minus_U * u_decay;
                tl = t_now;
                c0 = c0 * x * u;
                c1 = c1 * x * u;
            }
            double* cv = n.spk_c + (static_cast<size_t>(k2) * N + j) * 2;
            cv[0] = c0;
            cv[1] = c1;
            syn += row_syn;
            if (room) {
                n.r_step[r_head + before + j] = T0 + k2;
                n.r_src[r_head + before + j] = t;
            }
        }
    }
    if (valid && mine != 0 && stp) {
        n.stp_x[t] = x;
        n.stp_u[t] = u;
        n.stp_tl[t] = tl;
    }
    if (syn > 0) {
        add_i64(&n.c_syn, syn);
    }
    // every delivering wavefront of the network agrees on these
    if (!room && window_total > 0 && lane == 0) {
        atomicOr(&n.overflow, 1);
    }
    if (window_total > 64 * n.cw && lane == 0) {
        atomicOr(&n.overflow, 2);  // spikes beyond the delivery masks' capacity
    }
}

// Adds val (fixed point) into the ring slot of every target of one row, the
// whole wavefront together, 4 target loads in flight per lane.
__device__ inline void add_row(const uint32_t* __restrict__ tgt, uint32_t q0, uint32_t q1,
                               int64_t* base, int64_t val, uint32_t lane) {
    constexpr int U = 4;
    for (uint32_t q = q0 + lane; q < q1; q += 64 * U) {
        uint32_t tt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t qq = q + static_cast<uint32_t>(u) * 64;
            tt[u] = qq < q1 ? tgt[qq] : UINT32_MAX;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (tt[u] != UINT32_MAX) {
                atomicAdd(reinterpret_cast<unsigned long long*>(base + tt[u]),
                          static_cast<unsigned long long>(val));
            }
        }
    }
}

// An excitatory spike's two rows with one round of target loads: w_e into
// the exc ring slot of every exc-row target, multiplicity * w_a into the n2a
// ring slot of every (deduplicated) n2a-row target.
__device__ inline void add_rows_exc(const DevNet& n, uint32_t e0, uint32_t e1, int64_t* base_e,
                                    int64_t w_e, uint32_t a0, uint32_t a1, int64_t* base_a,
                                    int64_t w_a, uint32_t lane) {
    constexpr int UE = 4, UA = 2;
    const uint32_t* __restrict__ et = n.exc_tgt;
    const uint32_t* __restrict__ at = n.n2a_tgt;
    const uint32_t* __restrict__ am = n.n2a_mult;
    for (uint32_t qe = e0 + lane, qa = a0 + lane; qe < e1 || qa < a1; qe += 64 * UE, qa += 64 * UA) {
        uint32_t te[UE], ta[UA], ma[UA];
#pragma unroll
        for (int u = 0; u < UE; ++u) {
            const uint32_t q = qe + static_cast<uint32_t>(u) * 64;
            te[u] = q < e1 ? et[q] : UINT32_MAX;
        }
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const uint32_t q = qa + static_cast<uint32_t>(u) * 64;
            ta[u] = q < a1 ? at[q] : UINT32_MAX;
            ma[u] = q < a1 ? am[q] : 0u;
        }
#pragma unroll
        for (int u = 0; u < UE; ++u) {
            if (te[u] != UINT32_MAX) {
                atomicAdd(reinterpret_cast<unsigned long long*>(base_e + te[u]),
                          static_cast<unsigned long long>(w_e));
            }
        }
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            if (ta[u] != UINT32_MAX) {
                atomicAdd(reinterpret_cast<unsigned long long*>(base_a + ta[u]),
                          static_cast<unsigned long long>(w_a * static_cast<int64_t>(ma[u])));
            }
        }
    }
}

// Pass 3 of an FP32 network for the 64 neurons of one wavefront, over the
// whole window in time order. At each step the spiking lanes do their
// Tsodyks-Markram update (network.cpp:230-247) and raster entry, then the
// wavefront adds each spike's weights, in fixed point, into the arrival slot
// of every target of its rows: w_exc * x * u into the exc ring, -w_inh * x * u
// into the inh ring (the FP64 sums' sign convention), multiplicity * w_n2a *
// x * u into the n2a ring. The integer sums are the same in any order, so
// spikes need neither masks nor a summing pass. Raster entries are appended
// at r_head, reserved by one atomic per (wavefront, step) with spikes; their
// order within a window is restored when the raster is drained (sorted by
// step, then source).
__device__ __forceinline__ void deliver_fixed(DevNet& n, uint32_t first, uint32_t lane) {
    const uint32_t N = n.n_neur, NE = n.n_exc;
    const uint32_t t = first + lane;
    const bool valid = t < N;
    const int W = n.win_W;
    const int64_t T0 = n.win_T0;
    const int S = n.S;
    const size_t LN = n.ldn, LA = n.lda;
    // lane k < W holds the block's spike word of window step k (with the
    // inhibitory part's, in the block where the inhibitory neurons begin)
    const bool split = first < NE && first + 63 >= NE && NE < N;
    uint64_t l_word = 0;
    if (static_cast<int>(lane) < W) {
        l_word = n.spk_bits[static_cast<size_t>(lane) * n.words + (first >> 6)];
        if (split) l_word |= n.spk_extra[lane];
    }
    uint32_t mine = 0;  // bit k: this lane's neuron spiked at window step k
    for (int k2 = 0; k2 < W; ++k2) {
        mine |= static_cast<uint32_t>((lane_u64(l_word, k2) >> lane) & 1ULL) << k2;
    }
    const bool exc = t < NE;
    const bool stp = n.stp_on != 0;
    uint32_t r0 = 0, r1 = 0, a0 = 0, a1 = 0;
    double x = 0.0, u = 0.0, tl = 0.0;
    int64_t row_syn = 0;
    if (valid && mine != 0) {
        if (exc) {
            r0 = n.exc_row[t];
            r1 = n.exc_row[t + 1];
            a0 = n.n2a_row[t];
            a1 = n.n2a_row[t + 1];
            row_syn = (r1 - r0) + (n.n2a_raw_row[t + 1] - n.n2a_raw_row[t]);
        } else {
            r0 = n.inh_row[t - NE];
            r1 = n.inh_row[t - NE + 1];
            row_syn = r1 - r0;
        }
        if (stp) {
            x = n.stp_x[t];
            u = n.stp_u[t];
            tl = n.stp_tl[t];
        }
    }
    const uint64_t below = (1ULL << lane) - 1ULL;
    int rowE = static_cast<int>((T0 + n.d_exc) % S);
    int rowI = static_cast<int>((T0 + n.d_inh) % S);
    int rowA = static_cast<int>((T0 + n.d_n2a) % S);
    int64_t syn = 0;
    for (int k2 = 0; k2 < W; ++k2) {
        const uint64_t word = lane_u64(l_word, k2);
        if (word != 0) {
            // this step's raster entries: one reservation for the wavefront
            const uint32_t cnt = static_cast<uint32_t>(__popcll(word));
            int64_t base = 0;
            if (lane == 0) {
                base = static_cast<int64_t>(atomicAdd(reinterpret_cast<unsigned long long*>(&n.r_head),
                                                      static_cast<unsigned long long>(cnt)));
                add_i64(&n.c_spikes, cnt);
                if (base + cnt > n.r_cap) {
                    atomicOr(&n.overflow, 1);  // raster buffer full
                }
            }
            base = static_cast<int64_t>(lane_u64(static_cast<uint64_t>(base), 0));
            int64_t q0 = 0, q1 = 0;  // this lane's spike weights, fixed point
            if ((mine >> k2) & 1u) {
                double c0 = exc ? n.w_exc : n.w_inh;
                double c1 = n.w_n2a;
                const double t_now = static_cast<double>(T0 + k2) * n.dt;
                if (stp) {
                    const double dts = t_now - tl;
                    const double x_decay = asg_exp(-dts / n.tau_rec);
                    const double u_decay = n.fac_on ? asg_exp(-dts / n.tau_fac) : 0.0;
                    x = 1.0 + (x - x * u - 1.0) * x_decay;
                    u = n.U + u * n.one_minus_U * u_decay;
                    tl = t_now;
                    c0 = c0 * x * u;
                    c1 = c1 * x * u;
                }
                q0 = llrint(c0 * kFixScale);
                q1 = llrint(c1 * kFixScale);
                syn += row_syn;
                const int64_t at = base + __popcll(word & below);
                if (at < n.r_cap) {
                    n.r_step[at] = T0 + k2;
                    n.r_src[at] = t;
                }
            }
            uint64_t m = word;
            while (m) {
                const int L = __ffsll(static_cast<long long>(m)) - 1;
                m &= m - 1;
                const int64_t w0 = static_cast<int64_t>(lane_u64(static_cast<uint64_t>(q0), L));
                if (first + static_cast<uint32_t>(L) < NE) {
                    add_rows_exc(n, lane_u32(r0, L), lane_u32(r1, L),
                                 n.qring_exc + static_cast<size_t>(rowE) * LN, w0, lane_u32(a0, L),
                                 lane_u32(a1, L), n.qring_astro + static_cast<size_t>(rowA) * LA,
                                 static_cast<int64_t>(lane_u64(static_cast<uint64_t>(q1), L)), lane);
                } else {
                    // the inh ring holds the negated weights (as the FP64 sums)
                    add_row(n.inh_tgt, lane_u32(r0, L), lane_u32(r1, L),
                            n.qring_inh + static_cast<size_t>(rowI) * LN, -w0, lane);
                }
            }
        }
        rowE = ring_next(rowE, S);
        rowI = ring_next(rowI, S);
        rowA = ring_next(rowA, S);
    }
    if (valid && mine != 0 && stp) {
        n.stp_x[t] = x;
        n.stp_u[t] = u;
        n.stp_tl[t] = tl;
    }
    if (syn > 0) {
        add_i64(&n.c_syn, syn);
    }
}

// Window pass 3 of the FP64 networks: D wavefronts deliver the blocks the
// cell pass listed as having spiked (marking in groups of SG steps, the
// first group also doing the sources' plasticity and raster entries), then
// ceil(Wg / G) wavefronts per block sum the SIC of its neurons. G trades
// re-reading the astrocyte lists (G = Wg reads them once per window) against
// parallelism for small networks (G = 1).
template <int G>
__global__ __launch_bounds__(256) void k_deliver(DevNet* nets, const CellBlock* nblocks,
                                                 int32_t n_items, int32_t Wg, int32_t D,
                                                 int32_t SG, const CellBlock* work,
                                                 const int32_t* work_n) {
    const int e = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) * 4 +
                                                 static_cast<int>(threadIdx.x >> 6));
    if (e >= n_items) {
        return;
    }
    const uint32_t lane = threadIdx.x & 63;
    // The first D items are delivery wavefronts that share out the list of
    // blocks that spiked in the window (first, so that their chains of
    // dependent loads start with the launch): a block is ceil(Wg / SG)
    // items, one per group of SG steps to mark, the first also delivering
    // the block's sources.
    if (e < D) {
        const int groups = (Wg + SG - 1) / SG;
        const int items = *work_n * groups;
        for (int j = e; j < items; j += D) {
            const CellBlock cb = work[j / groups];
            DevNet& n = nets[cb.net];
            const int k0 = (j % groups) * SG;
            const int k1 = min(k0 + SG, n.win_W);
            if (k1 > k0) {
                mark_steps(n, cb.first, lane, k0, k1);
            }
            if (k0 == 0) {
                deliver_sources(n, cb.first, lane);
            }
        }
        return;
    }
    // the others: the SIC wavefronts, ceil(Wg / G) per block
    const int per_block = (Wg + G - 1) / G;
    const CellBlock cb = nblocks[(e - D) / per_block];
    DevNet& n = nets[cb.net];
    const int g0 = ((e - D) % per_block) * G;
    const int len = min(G, n.win_W - g0);
    const uint32_t t = cb.first + lane;
    if (len > 0 && t < n.n_neur) {
        deliver_sic<G>(n, t, g0, len);
    }
}

// Window pass 3 of the FP32 networks, with this window's close and the next
// window's drive, in one launch of one-wavefront workgroups:
//  * workgroups [0, C): the close, one lane per network: the liveness bits
//    of the window's SIC exchanges at their arrival steps, counters, the
//    next window's first step; the other parity's live word and lists are
//    emptied for the next window (nothing here reads them);
//  * [C, C + D): the delivery wavefronts, sharing out the list of neuron
//    blocks that spiked (deliver_fixed);
//  * [C + D, C + D + n_nblocks): the drive of the next window, one
//    wavefront per neuron block -- it depends on nothing this window
//    computes, and fills the device while the delivery chains wait on
//    memory; its wavefronts are the launch's longest, so they are
//    dispatched right after the delivery ones;
//  * [C + D + n_nblocks, C + D + n_nblocks + E): SIC wavefronts, sharing
//    out the neuron blocks of the networks the cell pass listed as having a
//    live exchange (a network whose astrocytes all have a zero SIC factor
//    has nothing to sum).
// The close writes only what no other workgroup of the launch reads (the
// live word it reads is this parity's, which nothing writes now), so the
// workgroups need no ordering among themselves.
// Item e of the post pass (the layout above), run by one whole wavefront.
__device__ __forceinline__ void post_item(int e, DevNet* nets, const CellBlock* nblocks,
                                          int32_t n_nblocks, int32_t Wg, int32_t C, int32_t D,
                                          int32_t E, const CellBlock* work,
                                          const int32_t* work_n, const PostState* post,
                                          const int32_t* sic_item, const int32_t* active,
                                          int32_t n_active, int32_t par, int32_t* other_work_n,
                                          PostState* other_post, double* drive_lds) {
    const uint32_t lane = threadIdx.x & 63;
    if (e < C) {
        const int j = e * 64 + static_cast<int>(lane);
        if (j < n_active) {
            DevNet& n = nets[active[j]];
            if (n.open) {
                close_window_with(n, n.sic_live_p[par]);
            }
            n.sic_live_p[par ^ 1] = 0;
            n.split_state[par ^ 1] = 0;
        }
        if (j == 0) {
            *other_work_n = 0;
            other_post->sic_items = 0;
        }
    } else if (e < C + D) {
        // each of the lane's neuron blocks is listed at most once a window
        const int items = min(*work_n, n_nblocks);
        for (int j = e - C; j < items; j += D) {
            const CellBlock cb = work[j];
            deliver_fixed(nets[cb.net], cb.first, lane);
        }
    } else if (e < C + D + n_nblocks) {
        const CellBlock cb = nblocks[e - C - D];
        drive_block(nets[cb.net], cb, Wg, drive_lds);
    } else if (e < C + D + n_nblocks + E) {
        const uint32_t items = min(post->sic_items, static_cast<uint32_t>(n_nblocks));
        for (uint32_t j = static_cast<uint32_t>(e - C - D - n_nblocks); j < items;
             j += static_cast<uint32_t>(E)) {
            const CellBlock cb = nblocks[sic_item[j]];
            DevNet& n = nets[cb.net];
            const uint32_t t = cb.first + lane;
            if (t < n.n_neur) {
                deliver_sic_f32<kWindowCap>(n, t, n.win_W, n.sic_live_p[par]);
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_post(DevNet* nets, const CellBlock* nblocks,
                                             int32_t n_nblocks, int32_t Wg, int32_t C, int32_t D,
                                             int32_t E, const CellBlock* work,
                                             const int32_t* work_n, const PostState* post,
                                             const int32_t* sic_item, const int32_t* active,
                                             int32_t n_active, int32_t par,
                                             int32_t* other_work_n, PostState* other_post) {
    extern __shared__ double drive_lds[];
    post_item(static_cast<int>(blockIdx.x), nets, nblocks, n_nblocks, Wg, C, D, E, work, work_n,
              post, sic_item, active, n_active, par, other_work_n, other_post, drive_lds);
}

// ------------------------------------------------ persistent FP32 launch ---
//
// One cooperative launch runs all of a call's windows of the FP32 networks.
// The networks are split into teams (grouped as the lanes of the per-kernel
// launches are); team t owns `wgs` one-wavefront workgroups, which run the
// call's first drive and then, window by window, the team's cell pass and
// its post pass, with a barrier among the team's workgroups after each.
// Networks exchange nothing, so no team ever waits for another: one team's
// post pass (latency-bound delivery chains) runs beside another's cell pass
// (bound by instruction issue).
struct Team {
    int32_t cell_first, cells;      // its blocks in the block table
    int32_t nblock_first, nblocks;  // its neuron blocks in the neuron-block table
    int32_t active_first, nets;     // its networks in the active list
    int32_t wgs;                    // its workgroups
    int32_t pad;
};

// Longest wait at a team barrier before the launch gives up: 2 s of the
// 100 MHz real-time counter (ASG_TEAM_WAIT_US overrides it, for tests of
// the failure path). One pass of a team takes milliseconds at most.
constexpr uint64_t kTeamWaitTicks = 200000000ull;

// Barrier among a team's workgroups, in the pattern of a grid-wide barrier:
// every arrival adds one to the team's counter (which only grows during a
// launch) and the barrier is passed once it reaches `target`; agent-scope
// release before the arrival and acquire after it make every member's
// stores before the barrier visible to every member after it. A wait longer
// than wait_ticks sets *failed and *host_failed; a member that finds
// *failed set stops waiting too. Returns false then, and the workgroup
// leaves the kernel, so every wavefront exits.
__device__ __forceinline__ bool team_barrier(uint32_t* counter, uint32_t target, int32_t* failed,
                                             int32_t* host_failed, uint64_t wait_ticks) {
    __syncthreads();
    int32_t ok = 1;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = 0;
                break;
            }
            if (__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks) {
                __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(host_failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(ok) != 0;
}

// The persistent launch: workgroup g is member wg_team[g].y of team
// wg_team[g].x. The team's lists are those a lane with the same blocks
// would use (offsets from nb0, the FP64 networks' neuron block count), and
// its windows alternate them by parity from par0. bar holds a counter per
// team, 64 bytes apart, zero at the launch.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_persist(DevNet* nets, const CellBlock* blocks,
                                                const CellBlock* nblocks, const Team* teams,
                                                const int2* wg_team, const int32_t* active,
                                                const int32_t* nbr, CellBlock* work,
                                                int32_t* work_n, PostState* post,
                                                int32_t* sic_item, int32_t nb0, int32_t Wg,
                                                int32_t windows, int32_t par0, int32_t dwaves,
                                                int32_t swaves, uint32_t* bar, int32_t* failed,
                                                int32_t* host_failed, uint64_t wait_ticks) {
    extern __shared__ double drive_lds[];
    const int t = __builtin_amdgcn_readfirstlane(wg_team[blockIdx.x].x);
    const int r = __builtin_amdgcn_readfirstlane(wg_team[blockIdx.x].y);
    const Team tm = teams[t];
    const int T = tm.wgs;
    uint32_t* counter = bar + 16 * t;
    uint32_t target = 0;
    for (int j = r; j < tm.nblocks; j += T) {  // the call's first drive
        const CellBlock cb = nblocks[tm.nblock_first + j];
        drive_block(nets[cb.net], cb, Wg, drive_lds);
    }
    target += static_cast<uint32_t>(T);
    if (!team_barrier(counter, target, failed, host_failed, wait_ticks)) {
        return;
    }
    const int rel = tm.nblock_first - nb0;
    CellBlock* work_t = work + nb0 + 2 * rel;
    int32_t* sic_t = sic_item + 2 * rel;
    int32_t* work_n_t = work_n + 1 + 2 * t;
    PostState* post_t = post + 2 * t;
    const CellBlock* nblocks_t = nblocks + tm.nblock_first;
    const int C = (tm.nets + 63) / 64;
    const int D = min(tm.nblocks, dwaves);
    const int E = swaves > 0 ? min(tm.nblocks, swaves) : tm.nblocks;
    const int items = C + D + E + tm.nblocks;
    for (int w = 0; w < windows; ++w) {
        const int p = (par0 + w) & 1;
        for (int b = r; b < tm.cells; b += T) {
            cells_item<ASG_FP32>(nets, blocks[tm.cell_first + b], Wg, work_t + p * tm.nblocks,
                                 work_n_t + p, p, post_t + p, sic_t + p * tm.nblocks, nbr, nullptr);
        }
        target += static_cast<uint32_t>(T);
        if (!team_barrier(counter, target, failed, host_failed, wait_ticks)) {
            return;
        }
        for (int e = r; e < items; e += T) {
            post_item(e, nets, nblocks_t, tm.nblocks, Wg, C, D, E, work_t + p * tm.nblocks,
                      work_n_t + p, post_t + p, sic_t + p * tm.nblocks, active + tm.active_first,
                      tm.nets, p, work_n_t + (p ^ 1), post_t + (p ^ 1), drive_lds);
        }
        target += static_cast<uint32_t>(T);
        if (!team_barrier(counter, target, failed, host_failed, wait_ticks)) {
            return;
        }
    }
}

// Zeroes the persistent launch's barrier counters and failure word.
__global__ __launch_bounds__(256) void k_persist_prep(uint32_t* bar, int32_t n, int32_t* failed) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < n) {
        bar[j] = 0;
    }
    if (j == 0) {
        *failed = 0;
    }
}

// Multiplicity of astrocyte a in source s's deduplicated n2a row.
__device__ inline uint32_t n2a_multiplicity(const DevNet& n, uint32_t s, uint32_t a) {
    uint32_t lo = n.n2a_row[s], hi = n.n2a_row[s + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (n.n2a_tgt[mid] < a) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return (lo < n.n2a_row[s + 1] && n.n2a_tgt[lo] == a) ? n.n2a_mult[lo] : 0u;
}

// The window's per-step spike counts as uniform arrays: before[k] spikes
// precede step k in window order, nE[k] of step k's are excitatory. Loaded
// by lanes 0..W-1 and broadcast, so any lane can look up any step.
struct WindowCounts {
    uint32_t before[kWindowCap + 1], nE[kWindowCap];
    __device__ __forceinline__ WindowCounts(uint32_t l_before, uint32_t l_nE, uint32_t total) {
#pragma unroll
        for (int k = 0; k < kWindowCap; ++k) {
            before[k] = lane_u32(l_before, k);
            nE[k] = lane_u32(l_nE, k);
        }
        before[kWindowCap] = total;
    }
    // The window step of window spike index jw (< total): the largest k < W
    // with before[k] <= jw.
    __device__ __forceinline__ int step_of(uint32_t jw, int W) const {
        int k = 0;
#pragma unroll
        for (int kk = 1; kk < kWindowCap; ++kk) {
            k += (kk < W && jw >= before[kk]) ? 1 : 0;
        }
        return k;
    }
    __device__ __forceinline__ uint32_t before_of(int k) const {
        uint32_t v = 0;
#pragma unroll
        for (int kk = 0; kk < kWindowCap; ++kk) {
            v = k == kk ? before[kk] : v;
        }
        return v;
    }
    __device__ __forceinline__ uint32_t nE_of(int k) const {
        uint32_t v = 0;
#pragma unroll
        for (int kk = 0; kk < kWindowCap; ++kk) {
            v = k == kk ? nE[kk] : v;
        }
        return v;
    }
};

// The ordered sums of one target over window steps [k0, k1). Its window
// mask has a bit per spike that reached it, at the spike's window index; the
// group's spikes are the indices [before[k0], before[k1]). In ascending bit
// order the spikes of one step come together, excitatory before inhibitory,
// each run in ascending spike (source) order. Each run is added into the
// arrival slot of its step and kind, starting from the slot's value, and
// stored when the run ends. The bits read are cleared: a mask word that
// lies wholly in this group's range with a store, a word shared with the
// neighbouring group's range with an atomic AND.
__device__ __forceinline__ void sum_window(const DevNet& n, int kind, uint32_t i, int64_t T0, int W,
                                           int k0, int k1, const WindowCounts& wc) {
    const size_t N = n.n_neur;
    const int S = n.S;
    const uint32_t lo = wc.before_of(k0);
    const uint32_t hi = min(k1 == W ? wc.before[kWindowCap] : wc.before_of(k1), 64 * n.cw);
    const bool neuron = kind == 0;
    const size_t targets = neuron ? N : static_cast<size_t>(n.n_astro);
    const size_t ld = neuron ? static_cast<size_t>(n.ldn) : static_cast<size_t>(n.lda);
    if (i >= targets) {
        return;
    }
    uint64_t* __restrict__ mask = (neuron ? n.mask_n : n.mask_a) + i;
    const double* __restrict__ spk_c = n.spk_c;
    const int rowE0 = static_cast<int>((T0 + n.d_exc) % S);
    const int rowI0 = static_cast<int>((T0 + n.d_inh) % S);
    const int rowA0 = static_cast<int>((T0 + n.d_n2a) % S);
    double acc = 0.0;
    double* run_ring = nullptr;  // ring and slot of the open run
    size_t run_slot = 0;
    int run_key = -1;       // 2 * step + (inhibitory ? 1 : 0)
    for (uint32_t c = lo >> 6; lo < hi && c <= (hi - 1) >> 6; ++c) {
        uint64_t m = mask[c * ld];
        const uint32_t b0 = max(lo, c * 64) - c * 64, b1 = min(hi, c * 64 + 64) - c * 64;
        m &= (b1 == 64 ? ~0ULL : (1ULL << b1) - 1ULL) & ~((1ULL << b0) - 1ULL);
        if (m == 0) {
            continue;
        }
        const bool owned = (k0 == 0 || lo <= c * 64) && (k1 == W || hi >= c * 64 + 64);
        if (owned) {
            mask[c * ld] = 0;
        } else {
            atomicAnd(reinterpret_cast<unsigned long long*>(mask + c * ld), ~m);
        }
        while (m) {
            const uint32_t jw = c * 64 + static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
            m &= m - 1;
            const int k = wc.step_of(jw, W);
            const uint32_t j = jw - wc.before_of(k);
            const size_t cv = (static_cast<size_t>(k) * N + j) * 2;
            if (neuron) {
                const bool inh = j >= wc.nE_of(k);
                const int key = 2 * k + (inh ? 1 : 0);
                if (key != run_key) {
                    if (run_ring != nullptr) {
                        run_ring[run_slot] = acc;
                    }
                    run_key = key;
                    const int row0 = inh ? rowI0 : rowE0;
                    const int row = row0 + k < S ? row0 + k : row0 + k - S;
                    run_ring = inh ? n.ring_inh : n.ring_exc;
                    run_slot = static_cast<size_t>(row) * ld + i;
                    acc = run_ring[run_slot];
                }
                acc += inh ? -spk_c[cv] : spk_c[cv];
            } else {
                if (k != run_key) {
                    if (run_ring != nullptr) {
                        run_ring[run_slot] = acc;
                    }
                    run_key = k;
                    const int row = rowA0 + k < S ? rowA0 + k : rowA0 + k - S;
                    run_ring = n.ring_astro;
                    run_slot = static_cast<size_t>(row) * ld + i;
                    acc = run_ring[run_slot];
                }
                // spike j's source counts its n2a multiplicity m times: m
                // repeated additions (network.cpp:249-279 adds once per
                // synapse)
                const uint32_t mult = n2a_multiplicity(n, n.spk_list[static_cast<size_t>(k) * N + j], i);
                const double w = spk_c[cv + 1];
                for (uint32_t r = mult; r > 0; --r) {
                    acc += w;
                }
            }
        }
    }
    if (run_ring != nullptr) {
        run_ring[run_slot] = acc;
    }
}

// Window pass 4, one lane per target and a wavefront per group of SG
// window steps, four wavefronts per workgroup: the window's ordered sums into
// the rings (network.cpp:249-279 scatters in ascending source order per
// step). The lead wavefront then closes the window.
__global__ __launch_bounds__(256) void k_sum(DevNet* nets, const CellBlock* blocks,
                                             int32_t n_items, int32_t Wg, int32_t SG) {
    const int e = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) * 4 +
                                                 static_cast<int>(threadIdx.x >> 6));
    if (e >= n_items) {
        return;
    }
    const int per_block = (Wg + SG - 1) / SG;
    const CellBlock cb = blocks[e / per_block];
    const int k0 = (e % per_block) * SG;
    const uint32_t lane = threadIdx.x & 63;
    DevNet& n = nets[cb.net];
    const int W = n.win_W;
    const int64_t T0 = n.win_T0;
    uint32_t l_tot = 0, l_nE = 0;  // lane k < W: spike counts of window step k
    if (static_cast<int>(lane) < W) {
        l_tot = n.spk_n[2 * lane];
        l_nE = n.spk_n[2 * lane + 1];
    }
    uint32_t total = 0;
    const uint32_t l_before = wave_exclusive_scan(l_tot, total);
    if (cb.lead && k0 == 0 && lane == 0) {
        close_window(n, total);
    }
    const int k1 = min(k0 + SG, W);
    if (total == 0 || k1 <= k0) {
        return;
    }
    const WindowCounts wc(l_before, l_nE, total);
    sum_window(n, cb.kind, cb.first + lane, T0, W, k0, k1, wc);
}

// Element-wise evaluation of the step's FP64 functions (diagnostics).
// Per-call bookkeeping of a batch, one launch each instead of one copy per
// network: the new run targets of the active networks, and the raster heads
// and overflow bits read back together.
__global__ __launch_bounds__(256) void k_set_targets(DevNet* nets, const int32_t* active,
                                                     int32_t n_active, const int64_t* targets) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < n_active) {
        DevNet& n = nets[active[j]];
        n.target = targets[j];
        set_clock(n, n.step);
        n.split_state[0] = 0;
        n.split_state[1] = 0;
    }
}

__global__ __launch_bounds__(256) void k_raster_heads(const DevNet* nets, const int32_t* active,
                                                      int32_t n_active, int64_t* heads,
                                                      int32_t* flags) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < n_active) {
        // an FP32 network's reservations may run past a full buffer
        heads[j] = min(nets[active[j]].r_head, nets[active[j]].r_cap);
        flags[j] = nets[active[j]].overflow;
    }
}

// Packs the raster entries of active network blockIdx.y (heads[y] of them)
// at offsets[y] of one staging buffer and empties that network's raster:
// as (step, source) pairs, or, with out_key, as keys (step - bases[y]) *
// 2^32 + source, which sort like the pairs (the host checks that every step
// offset fits 32 bits).
__global__ __launch_bounds__(256) void k_raster_gather(DevNet* nets, const int32_t* active,
                                                       const int64_t* heads,
                                                       const int64_t* offsets, int64_t* out_step,
                                                       uint32_t* out_src, const int64_t* bases,
                                                       uint64_t* out_key) {
    DevNet& n = nets[active[blockIdx.y]];
    const int64_t cnt = heads[blockIdx.y];
    const int64_t off = offsets[blockIdx.y];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    if (out_key != nullptr) {
        const int64_t base = bases[blockIdx.y];
        for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < cnt;
             e += stride) {
            out_key[off + e] = (static_cast<uint64_t>(n.r_step[e] - base) << 32) | n.r_src[e];
        }
    } else {
        for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < cnt;
             e += stride) {
            out_step[off + e] = n.r_step[e];
            out_src[off + e] = n.r_src[e];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && cnt > 0) {
        n.r_head = 0;  // every kernel that appends has finished (same stream)
    }
}

__global__ void k_math(int32_t fn, const double* x, double* y, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double v = x[i];
        if (fn == ASG_MATH_DIV) {
            // pairs (a, b): a / b by div_rn, b passed through (n is even)
            y[i] = (i & 1) ? v : div_rn(v, x[i + 1], 1.0 / x[i + 1]);
            continue;
        }
        double r;
        switch (fn) {
            case ASG_MATH_EXP: r = asg_exp(v); break;
            case ASG_MATH_LOG: r = asg_log(v); break;
            case ASG_MATH_SIN: r = asg_sin(v); break;
            case ASG_MATH_COS: r = asg_cos(v); break;
            case ASG_MATH_EXP2_F32: r = spike_exp2(v); break;
            default: r = asg_sqrt(v); break;
        }
        y[i] = r;
    }
}

// ------------------------------------------------------- host: one net -----

namespace {

int steps_per_noise(double noise_dt, double dt) {
    const long long r = std::llround(noise_dt / dt);
    return std::max(1, static_cast<int>(r));
}

NeuronK make_neuron_k(const asg_neuron_params& s, double h) {
    NeuronK p{};
    p.C_m = s.C_m;
    p.neg_g_L = -s.g_L;
    p.E_L = s.E_L;
    p.V_th = s.V_th;
    p.Delta_T = s.Delta_T;
    p.a = s.a;
    p.tau_w = s.tau_w;
    p.V_peak = s.V_peak;
    p.E_ex = s.E_ex;
    p.E_in = s.E_in;
    p.I_e = s.I_e;
    p.inv_C_m = 1.0 / s.C_m;
    p.inv_Delta_T = 1.0 / s.Delta_T;
    p.inv_tau_w = 1.0 / s.tau_w;
    p.gL_DeltaT = s.g_L * s.Delta_T;
    constexpr double kE = 2.7182818284590452353602874713527;
    p.psc_ex = kE / s.tau_syn_ex;
    p.psc_in = kE / s.tau_syn_in;
    p.ex_decay = asg_exp(-h / s.tau_syn_ex);
    p.in_decay = asg_exp(-h / s.tau_syn_in);
    p.refr_steps = static_cast<int>(s.t_ref / h + 0.5);
    return p;
}

InputK make_input_k(const asg_input_params& in, double dt) {
    InputK k{};
    k.poiss_rate = in.poiss_rate;
    k.poiss_weight = in.poiss_weight;
    k.lambda = in.poiss_rate * dt * 1e-3;
    k.limit = asg_exp(-k.lambda);
    k.noise_std = in.gauss_noise_std;
    k.independent = in.independent_noise != 0;
    k.steps_per_noise = steps_per_noise(in.noise_dt, dt);
    return k;
}

NeuronKf make_neuron_kf(const asg_neuron_params& s) {
    constexpr double kLog2e = 1.4426950408889634074;
    NeuronKf p{};
    p.inv_C_m = 1.0 / s.C_m;
    p.inv_tau_w = 1.0 / s.tau_w;
    p.exp_a = kLog2e / s.Delta_T;
    p.exp_b = -s.V_th * kLog2e / s.Delta_T;
    p.gL_EL_Ie = s.g_L * s.E_L + s.I_e;
    p.a_EL = s.a * s.E_L;
    p.g_L = s.g_L;
    return p;
}

std::string check_csr(const char* name, const asg_csr& c, uint32_t rows, uint32_t n_targets) {
    if (c.n_rows != rows) {
        return std::string(name) + ": n_rows does not match the population size";
    }
    if (c.row_start == nullptr) {
        return std::string(name) + ": row_start is NULL";
    }
    if (c.row_start[0] != 0) {
        return std::string(name) + ": row_start[0] must be 0";
    }
    for (uint32_t r = 0; r < rows; ++r) {
        if (c.row_start[r + 1] < c.row_start[r]) {
            return std::string(name) + ": row_start is not monotone";
        }
    }
    const uint32_t nnz = c.row_start[rows];
    if (nnz > 0 && c.target == nullptr) {
        return std::string(name) + ": target is NULL";
    }
    for (uint32_t q = 0; q < nnz; ++q) {
        if (c.target[q] >= n_targets) {
            return std::string(name) + ": target index out of range";
        }
    }
    if (c.delay_steps < 1) {
        return std::string(name) + ": delay_steps must be >= 1";
    }
    return "";
}

std::string validate(const asg_net_desc& d) {
    const uint32_t N = d.n_exc + d.n_inh;
    if (N + d.n_astro == 0) return "the network has no cells";
    if (d.substeps < 1) return "substeps must be at least 1";
    if (!(d.dt > 0.0)) return "dt must be positive";
    if (d.sic_interval < 1) return "sic_interval must be at least 1";
    if (d.precision != ASG_FP64 && d.precision != ASG_FP32) return "unknown precision";
    if (N > 0 && (d.V_reset == nullptr || d.b == nullptr)) return "V_reset / b arrays are NULL";
    if (d.n_astro > 0 && (d.Ca_tot == nullptr || d.IP3_0 == nullptr || d.tau_IP3 == nullptr ||
                          d.delta_IP3 == nullptr)) {
        return "astrocyte parameter arrays are NULL";
    }
    std::string e;
    if (!(e = check_csr("exc", d.exc, d.n_exc, N)).empty()) return e;
    if (!(e = check_csr("inh", d.inh, d.n_inh, N)).empty()) return e;
    if (!(e = check_csr("n2a", d.n2a, d.n_exc, d.n_astro)).empty()) return e;
    if (!(e = check_csr("a2n", d.a2n, d.n_astro, N)).empty()) return e;
    return "";
}

inline size_t round_up64(size_t n) { return (n + 63) / 64 * 64; }

// Host-side derived connectivity: deduplicated n2a rows and the SIC gather
// lists (a2n transposed to targets, astrocytes ascending, with multiplicity).
struct DerivedConn {
    std::vector<uint32_t> n2a_row, n2a_tgt, n2a_mult;
    uint32_t sic_width = 0;
    int64_t sic_pairs = 0;
    std::vector<uint32_t> sic_src, sic_mult;  // [width][round_up64(N)]
};

DerivedConn derive(const asg_net_desc& d) {
    DerivedConn out;
    const uint32_t N = d.n_exc + d.n_inh;
    out.n2a_row.assign(d.n_exc + 1, 0);
    std::vector<uint32_t> row;
    for (uint32_t s = 0; s < d.n_exc; ++s) {
        row.assign(d.n2a.target + d.n2a.row_start[s], d.n2a.target + d.n2a.row_start[s + 1]);
        std::sort(row.begin(), row.end());
        for (size_t q = 0; q < row.size();) {
            size_t e = q;
            while (e < row.size() && row[e] == row[q]) ++e;
            out.n2a_tgt.push_back(row[q]);
            out.n2a_mult.push_back(static_cast<uint32_t>(e - q));
            q = e;
        }
        out.n2a_row[s + 1] = static_cast<uint32_t>(out.n2a_tgt.size());
    }
    // Count (target, astro) pairs; astrocytes are visited ascending, so
    // each target's list comes out sorted by astrocyte.
    std::vector<uint32_t> count(N + 1, 0);
    std::vector<uint32_t> last(N, UINT32_MAX);
    for (uint32_t a = 0; a < d.n_astro; ++a) {
        for (uint32_t q = d.a2n.row_start[a]; q < d.a2n.row_start[a + 1]; ++q) {
            const uint32_t t = d.a2n.target[q];
            if (last[t] != a) {
                last[t] = a;
                ++count[t + 1];
            }
        }
    }
    // Column-major (ELL) layout: width = most pairs of any target.
    uint32_t width = 0;
    for (uint32_t t = 0; t < N; ++t) width = std::max(width, count[t + 1]);
    out.sic_width = width;
    out.sic_pairs = 0;
    const size_t ld = round_up64(N);
    out.sic_src.assign(static_cast<size_t>(width) * ld, 0);
    out.sic_mult.assign(static_cast<size_t>(width) * ld, 0);
    std::vector<uint32_t> fill(N, 0);
    std::fill(last.begin(), last.end(), UINT32_MAX);
    for (uint32_t a = 0; a < d.n_astro; ++a) {
        for (uint32_t q = d.a2n.row_start[a]; q < d.a2n.row_start[a + 1]; ++q) {
            const uint32_t t = d.a2n.target[q];
            if (last[t] != a) {
                last[t] = a;
                const size_t at = static_cast<size_t>(fill[t]) * ld + t;
                out.sic_src[at] = a;
                out.sic_mult[at] = 1;
                ++fill[t];
                ++out.sic_pairs;
            } else {
                ++out.sic_mult[static_cast<size_t>(fill[t] - 1) * ld + t];
            }
        }
    }
    return out;
}

int window_of(const asg_net_desc& d) {
    int w = std::min(std::min(d.exc.delay_steps, d.inh.delay_steps),
                     std::min(d.n2a.delay_steps, d.a2n.delay_steps));
    return std::max(1, std::min(w, kWindowCap));
}

// Byte plan of one network's device arrays, so the footprint query and the
// allocation agree.
struct Plan {
    size_t N, NA, words, S, Wmax, cw, r_cap, hist_bits, ldn, lda;
    size_t n_exc_syn, n_inh_syn, n2a_dedup, sic_pairs;
    size_t bytes;
};

Plan plan_of(const asg_net_desc& d, const DerivedConn* dc) {
    Plan p{};
    p.N = d.n_exc + d.n_inh;
    p.NA = d.n_astro;
    p.words = (p.N + 63) / 64;
    p.ldn = round_up64(p.N);
    p.lda = round_up64(p.NA);
    p.Wmax = static_cast<size_t>(window_of(d));
    const int dmax = std::max(std::max(d.exc.delay_steps, d.inh.delay_steps),
                              std::max(d.n2a.delay_steps, d.a2n.delay_steps));
    p.S = static_cast<size_t>(dmax) + p.Wmax;
    p.hist_bits = 64;
    while (p.hist_bits < p.S + 32) p.hist_bits *= 2;
    // Delivery masks hold 64 window spikes per chunk. Every neuron spiking
    // at every step of a window needs ceil(Wmax * N / 64) chunks; that is
    // kept while the masks fit in kMaskBudget, otherwise the chunk count is
    // cut to what fits (at least kMinChunks, i.e. 4096 spikes per window), and
    // a window that exceeds it is reported as ASG_ERR_OVERFLOW rather than
    // delivered in part.
    p.cw = (p.Wmax * p.N + 63) / 64;
    {
        const size_t per_chunk = (round_up64(p.N) + round_up64(p.NA)) * 8;
        if (p.cw * per_chunk > kMaskBudget && per_chunk > 0) {
            p.cw = std::min(p.cw, std::max<size_t>(kMinChunks, kMaskBudget / per_chunk));
        }
    }
    p.r_cap = std::max<size_t>(size_t(1) << 20, 16 * p.N);
    p.n_exc_syn = d.exc.row_start[d.n_exc];
    p.n_inh_syn = d.inh.row_start[d.n_inh];
    const size_t n2a_raw = d.n2a.row_start[d.n_exc];
    const size_t a2n_raw = d.a2n.row_start[d.n_astro];
    p.n2a_dedup = dc ? dc->n2a_tgt.size() : n2a_raw;
    p.sic_pairs = dc ? dc->sic_src.size() : a2n_raw;  // ELL entries
    const size_t fb = d.precision == ASG_FP32 ? 4 : 8;  // bytes of a state value
    size_t b = 0;
    // neuron state (V and w FP64 in both precisions), V_reset/b, STP
    b += p.N * (2 * 8 + 5 * fb + 4 + 2 * 8 + 3 * 8);
    b += p.NA * (3 * 8 + 4 * fb);                   // astrocyte state (FP64) and parameters
    const bool f32 = d.precision == ASG_FP32;
    // rings: FP64 exc/inh/SIC and n2a; FP32 fixed-point exc/inh and n2a, FP32 SIC
    b += f32 ? p.S * p.ldn * (8 + 8 + 4) + p.S * p.lda * 8 : p.S * p.ldn * 3 * 8 + p.S * p.lda * 8;
    // SIC factors: FP64 [Wmax][astrocytes]; FP32 one 16-float line per astrocyte
    b += (f32 ? 16 * p.lda * 4 + p.lda * 4 : p.Wmax * p.lda * 8) + p.hist_bits / 8;
    b += (d.n_exc + 1 + p.n_exc_syn + d.n_inh + 1 + p.n_inh_syn) * 4;
    b += (2 * (d.n_exc + 1) + 2 * p.n2a_dedup) * 4;
    b += 2 * p.sic_pairs * 4;
    b += p.Wmax * (p.words * 12 + p.N * 4 + 8);
    if (!f32) {
        b += p.Wmax * p.N * 16;               // spike weights
        b += p.cw * (p.ldn + p.lda) * 8;      // window masks
    }
    b += p.r_cap * 12;
    const bool poiss = d.input_exc.poiss_rate > 0.0 || d.input_inh.poiss_rate > 0.0;
    const bool noise = d.input_exc.gauss_noise_std > 0.0 || d.input_inh.gauss_noise_std > 0.0;
    b += (poiss ? p.Wmax * p.ldn * 4 : 0) + (noise ? p.Wmax * p.ldn * fb : 0);
    p.bytes = b;
    return p;
}

// Limits of the FP32 path. Its windows address a network's rings, drive
// and factor lines by 32-bit byte offsets from their bases, so each such
// array must stay below 4 GiB (about 13 million neurons or astrocytes at the
// shipped delays). Its cell kernel draws the astrocytes' Poisson input by
// Knuth's loop only, so their expected events per step must not exceed 30
// (above that the reference switches to a rounded normal: a Poisson rate
// over 300 kHz at dt = 0.1 ms, which the FP64 path runs).
std::string check_fp32_limits(const asg_net_desc& d) {
    if (d.precision != ASG_FP32) return "";
    if (d.input_astro.poiss_rate * d.dt * 1e-3 > 30.0) {
        return "FP32 path: astrocyte Poisson input above 30 events per step is not supported "
               "(run this network in FP64)";
    }
    const Plan p = plan_of(d, nullptr);
    const size_t limit = size_t(1) << 32;
    if (p.S * p.ldn * 8 >= limit || p.S * p.lda * 8 >= limit || p.lda * 64 >= limit ||
        p.Wmax * p.ldn * 4 >= limit) {
        return "FP32 network too large: a ring, drive or SIC-factor array would exceed 4 GiB";
    }
    return "";
}

// One network's device allocation and its host mirror.
// Device blocks that networks' arrays are carved from. A removed network's
// blocks wait here for the next added network, so that admission in a
// serving loop calls no hipMalloc (slow, tens of microseconds to
// milliseconds) or hipFree (which may wait for the device) in steady state.
// At most kPoolCap bytes are kept; a failed hipMalloc first returns the
// cached blocks to the device and tries again.
struct BlockPool {
    static constexpr size_t kGranule = size_t(2) << 20;
    static constexpr size_t kPoolCap = size_t(16) << 30;
    std::multimap<size_t, void*> free_blocks;  // size -> block
    size_t cached = 0;

    static size_t rounded(size_t need) {
        return (std::max<size_t>(need, 1) + kGranule - 1) / kGranule * kGranule;
    }

    // A block of at least rounded(need) bytes: a cached one of at most twice
    // that, or a new one. Returns the block and its size.
    std::pair<void*, size_t> acquire(size_t need) {
        const size_t size = rounded(need);
        auto it = free_blocks.lower_bound(size);
        if (it != free_blocks.end() && it->first <= 2 * size) {
            const std::pair<void*, size_t> b{it->second, it->first};
            cached -= it->first;
            free_blocks.erase(it);
            return b;
        }
        void* p = nullptr;
        if (hipMalloc(&p, size) != hipSuccess) {
            (void)hipGetLastError();
            flush();
            HIP_CHECK(hipMalloc(&p, size));
        }
        return {p, size};
    }
    void give_back(void* p, size_t size) {
        free_blocks.emplace(size, p);
        cached += size;
        while (cached > kPoolCap && !free_blocks.empty()) {
            auto big = std::prev(free_blocks.end());
            cached -= big->first;
            (void)hipFree(big->second);
            free_blocks.erase(big);
        }
    }
    void flush() {
        for (auto& b : free_blocks) (void)hipFree(b.second);
        free_blocks.clear();
        cached = 0;
    }
};

struct NetSlot {
    bool used = false;
    DevNet dn{};                  // host copy (device addresses)
    // The slot's arrays are carved, 256-byte aligned, from blocks of the
    // pool (one block sized by the network's planned footprint, as a rule),
    // each zeroed once when taken.
    BlockPool* pool = nullptr;
    std::vector<std::pair<void*, size_t>> blocks;
    char* carve = nullptr;        // next free byte of the current block
    size_t carve_left = 0;
    size_t block_hint = 0;        // size asked for the first block
    hipStream_t carve_stream = nullptr;
    size_t bytes = 0;
    int W = 1;
    int64_t step = 0;
    int64_t drained_to = 0;       // the raster holds entries of steps [drained_to, step)
    int64_t total = 0;            // run length; INT64_MAX for asg_net
    void* rec_alloc = nullptr;    // recording ring (owned separately: it can be replaced)
    int64_t rec_read = 0;         // records handed to the host so far
    std::vector<int64_t> q_step;  // drained raster not yet handed out
    std::vector<uint32_t> q_src;
    size_t q_pos = 0;
    int32_t overflow = 0;         // overflow bits read at the last drain
};

// `count` elements of T from the slot's arena, zero-filled.
template <typename T>
T* dev_alloc(NetSlot& s, size_t count) {
    const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
    if (bytes > s.carve_left) {
        const size_t want = std::max(bytes, s.block_hint);
        const std::pair<void*, size_t> b = s.pool->acquire(want);
        s.block_hint = 0;
        s.blocks.push_back(b);
        s.bytes += BlockPool::rounded(want);  // what the slot asked for (see Engine::add)
        HIP_CHECK(hipMemsetAsync(b.first, 0, b.second, s.carve_stream));
        s.carve = static_cast<char*>(b.first);
        s.carve_left = b.second;
    }
    T* p = reinterpret_cast<T*>(s.carve);
    s.carve += bytes;
    s.carve_left -= bytes;
    return p;
}

template <typename T>
T* dev_copy(NetSlot& s, const T* host, size_t count, hipStream_t st) {
    T* p = dev_alloc<T>(s, count);
    if (count > 0) {
        HIP_CHECK(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
    return p;
}

__global__ __launch_bounds__(256) void k_fill_f64(double* p, double value, int64_t count) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (j < count) {
        p[j] = value;
    }
}

double* dev_fill(NetSlot& s, double value, size_t count, hipStream_t st) {
    double* p = dev_alloc<double>(s, count);
    if (count > 0) {
        hipLaunchKernelGGL(k_fill_f64, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                           0, st, p, value, static_cast<int64_t>(count));
    }
    return p;
}

// The arena's blocks are zeroed when taken.
template <typename T>
T* dev_zero(NetSlot& s, size_t count, hipStream_t) {
    return dev_alloc<T>(s, count);
}

void release(NetSlot& s) {
    for (const auto& b : s.blocks) {
        s.pool->give_back(b.first, b.second);
    }
    s.blocks.clear();
    s.carve = nullptr;
    s.carve_left = 0;
    if (s.rec_alloc != nullptr) {
        (void)hipFree(s.rec_alloc);
        s.rec_alloc = nullptr;
    }
    s.rec_read = 0;
    s.used = false;
    s.q_step.clear();
    s.q_src.clear();
    s.q_pos = 0;
    s.overflow = 0;
}

// The host-side part of admitting a network: the derived connectivity and
// the plan. No device call, so any host thread can make it.
struct Prepared {
    DerivedConn dc;
    Plan pl;
};

Prepared prepare(const asg_net_desc& d) {
    Prepared p;
    p.dc = derive(d);
    p.pl = plan_of(d, &p.dc);
    return p;
}

// `prep` is prepare(d) made earlier (or null: made here).
void build_slot(NetSlot& s, const asg_net_desc& d, int64_t total, hipStream_t st,
                const Prepared* prep = nullptr) {
    Prepared local;
    if (prep == nullptr) {
        local = prepare(d);
        prep = &local;
    }
    const DerivedConn& dc = prep->dc;
    const Plan& pl = prep->pl;
    s.block_hint = pl.bytes;
    s.carve_stream = st;
    s.bytes = 0;
    DevNet& n = s.dn;
    n = DevNet{};
    n.n_exc = d.n_exc;
    n.n_inh = d.n_inh;
    n.n_neur = static_cast<uint32_t>(pl.N);
    n.n_astro = d.n_astro;
    n.words = static_cast<uint32_t>(pl.words);
    n.cw = static_cast<uint32_t>(pl.cw);
    n.ldn = static_cast<uint32_t>(pl.ldn);
    n.lda = static_cast<uint32_t>(pl.lda);
    n.S = static_cast<int32_t>(pl.S);
    n.Wmax = static_cast<int32_t>(pl.Wmax);
    n.d_exc = d.exc.delay_steps;
    n.d_inh = d.inh.delay_steps;
    n.d_n2a = d.n2a.delay_steps;
    n.d_a2n = d.a2n.delay_steps;
    n.sic_interval = d.sic_interval;
    n.substeps = d.substeps;
    n.seed = d.seed;
    n.dt = d.dt;
    n.h = d.dt / static_cast<double>(d.substeps);
    n.half_h = 0.5 * n.h;
    n.h_sixth = n.h / 6.0;
    n.pk[0] = make_neuron_k(d.neuron[0], n.h);
    n.pk[1] = make_neuron_k(d.neuron[1], n.h);
    n.in[0] = make_input_k(d.input_exc, d.dt);
    n.in[1] = make_input_k(d.input_inh, d.dt);
    n.ap.k = make_astro_k(d.astro.c);
    n.ap.SIC_th = d.astro.SIC_th;
    n.ap.SIC_scale = d.astro.SIC_scale;
    n.ap.in = make_input_k(d.input_astro, d.dt);
    n.stp_on = d.stp.enabled != 0;
    n.fac_on = !(d.stp.tau_fac < 1e-10);
    n.U = d.stp.U;
    n.one_minus_U = 1.0 - d.stp.U;
    n.tau_rec = d.stp.tau_rec;
    n.tau_fac = d.stp.tau_fac;
    n.w_exc = d.exc.weight;
    n.w_inh = d.inh.weight;
    n.w_n2a = d.n2a.weight;
    n.w_a2n = d.a2n.weight;

    n.precision = d.precision;
    n.pkf[0] = make_neuron_kf(d.neuron[0]);
    n.pkf[1] = make_neuron_kf(d.neuron[1]);
    n.akf = make_astro_kf(d.astro.c);
    const size_t N = pl.N, NA = pl.NA;
    std::vector<double> V0(N);
    for (uint32_t i = 0; i < N; ++i) {
        V0[i] = d.neuron[i < d.n_exc ? 0 : 1].V_m_init;
    }
    n.refr = dev_zero<int32_t>(s, N, st);
    n.V_reset = dev_copy(s, d.V_reset, N, st);
    n.b = dev_copy(s, d.b, N, st);
    if (d.precision == ASG_FP32) {
        auto to_f = [](const double* src, size_t cnt) {
            std::vector<float> v(cnt);
            for (size_t q = 0; q < cnt; ++q) v[q] = static_cast<float>(src[q]);
            return v;
        };
        n.V = dev_copy(s, V0.data(), N, st);
        n.w = dev_zero<double>(s, N, st);
        n.fgex = dev_zero<float>(s, N, st);
        n.fdgex = dev_zero<float>(s, N, st);
        n.fgin = dev_zero<float>(s, N, st);
        n.fdgin = dev_zero<float>(s, N, st);
        n.fIsic = dev_zero<float>(s, N, st);
        n.Ca = dev_fill(s, d.astro.Ca_init, NA, st);
        n.IP3 = dev_fill(s, d.astro.IP3_init, NA, st);
        n.hh = dev_fill(s, d.astro.h_init, NA, st);
        const std::vector<float> Ca_tot = to_f(d.Ca_tot, NA), IP3_0 = to_f(d.IP3_0, NA),
                                 delta = to_f(d.delta_IP3, NA);
        std::vector<float> inv_tau(NA);
        for (size_t q = 0; q < NA; ++q) inv_tau[q] = static_cast<float>(1.0 / d.tau_IP3[q]);
        n.fCa_tot = dev_copy(s, Ca_tot.data(), NA, st);
        n.fIP3_0 = dev_copy(s, IP3_0.data(), NA, st);
        n.finv_tau = dev_copy(s, inv_tau.data(), NA, st);
        n.fdelta = dev_copy(s, delta.data(), NA, st);
        n.qring_exc = dev_zero<int64_t>(s, pl.S * pl.ldn, st);
        n.qring_inh = dev_zero<int64_t>(s, pl.S * pl.ldn, st);
        n.fring_sic = dev_zero<float>(s, pl.S * pl.ldn, st);
        n.qring_astro = dev_zero<int64_t>(s, pl.S * pl.lda, st);
        n.ffac = dev_zero<float>(s, 16 * pl.lda, st);
        n.ffac_nz = dev_zero<uint32_t>(s, pl.lda, st);
        HIP_CHECK(hipStreamSynchronize(st));
    } else {
        n.V = dev_copy(s, V0.data(), N, st);
        n.w = dev_zero<double>(s, N, st);
        n.gex = dev_zero<double>(s, N, st);
        n.dgex = dev_zero<double>(s, N, st);
        n.gin = dev_zero<double>(s, N, st);
        n.dgin = dev_zero<double>(s, N, st);
        n.Isic = dev_zero<double>(s, N, st);
        n.Ca = dev_fill(s, d.astro.Ca_init, NA, st);
        n.IP3 = dev_fill(s, d.astro.IP3_init, NA, st);
        n.hh = dev_fill(s, d.astro.h_init, NA, st);
        n.Ca_tot = dev_copy(s, d.Ca_tot, NA, st);
        n.IP3_0 = dev_copy(s, d.IP3_0, NA, st);
        n.tau_IP3 = dev_copy(s, d.tau_IP3, NA, st);
        n.delta_IP3 = dev_copy(s, d.delta_IP3, NA, st);
        n.ring_exc = dev_zero<double>(s, pl.S * pl.ldn, st);
        n.ring_inh = dev_zero<double>(s, pl.S * pl.ldn, st);
        n.ring_sic = dev_zero<double>(s, pl.S * pl.ldn, st);
        n.ring_astro = dev_zero<double>(s, pl.S * pl.lda, st);
        n.fac = dev_zero<double>(s, pl.Wmax * pl.lda, st);
    }
    n.stp_x = dev_fill(s, 1.0, N, st);
    n.stp_u = dev_fill(s, d.stp.U, N, st);
    n.stp_tl = dev_fill(s, -1e12, N, st);
    n.sic_hist = dev_zero<uint32_t>(s, pl.hist_bits / 32, st);
    n.hist_mask = static_cast<uint32_t>(pl.hist_bits - 1);
    n.exc_row = dev_copy(s, d.exc.row_start, d.n_exc + 1, st);
    n.exc_tgt = dev_copy(s, d.exc.target, pl.n_exc_syn, st);
    n.inh_row = dev_copy(s, d.inh.row_start, d.n_inh + 1, st);
    n.inh_tgt = dev_copy(s, d.inh.target, pl.n_inh_syn, st);
    n.n2a_row = dev_copy(s, dc.n2a_row.data(), dc.n2a_row.size(), st);
    n.n2a_tgt = dev_copy(s, dc.n2a_tgt.data(), dc.n2a_tgt.size(), st);
    n.n2a_mult = dev_copy(s, dc.n2a_mult.data(), dc.n2a_mult.size(), st);
    n.n2a_raw_row = dev_copy(s, d.n2a.row_start, d.n_exc + 1, st);
    n.sic_width = dc.sic_width;
    n.sic_pairs = dc.sic_pairs;
    n.sic_src = dev_copy(s, dc.sic_src.data(), dc.sic_src.size(), st);
    n.sic_mult = dev_copy(s, dc.sic_mult.data(), dc.sic_mult.size(), st);
    n.spk_bits = dev_zero<uint64_t>(s, pl.Wmax * pl.words, st);
    n.spk_extra = dev_zero<uint64_t>(s, pl.Wmax, st);
    n.spk_woff = dev_zero<uint32_t>(s, pl.Wmax * pl.words, st);
    n.spk_list = dev_zero<uint32_t>(s, pl.Wmax * N, st);
    n.spk_n = dev_zero<uint32_t>(s, pl.Wmax * 2, st);
    if (d.precision != ASG_FP32) {
        // FP64 delivery: spike weights and window masks for the ordered sums
        n.spk_c = dev_zero<double>(s, pl.Wmax * N * 2, st);
        n.mask_n = dev_zero<uint64_t>(s, pl.cw * pl.ldn, st);
        n.mask_a = dev_zero<uint64_t>(s, pl.cw * pl.lda, st);
    }
    if (n.in[0].poiss_rate > 0.0 || n.in[1].poiss_rate > 0.0) {
        n.drive_n = dev_zero<int32_t>(s, pl.Wmax * pl.ldn, st);
    }
    if (n.in[0].noise_std > 0.0 || n.in[1].noise_std > 0.0) {
        if (d.precision == ASG_FP32) {
            n.fdrive_noise = dev_zero<float>(s, pl.Wmax * pl.ldn, st);
        } else {
            n.drive_noise = dev_zero<double>(s, pl.Wmax * pl.ldn, st);
        }
    }
    n.r_cap = static_cast<int64_t>(pl.r_cap);
    n.r_step = dev_alloc<int64_t>(s, pl.r_cap);
    n.r_src = dev_alloc<uint32_t>(s, pl.r_cap);
    HIP_CHECK(hipStreamSynchronize(st));
    s.W = static_cast<int>(pl.Wmax);
    s.step = 0;
    s.drained_to = 0;
    s.total = total;
    s.used = true;
}

}  // namespace

// --------------------------------------------------------------- engine ----

// A set of networks stepped together. asg_net wraps an engine with one slot.
struct Engine {
    hipStream_t stream = nullptr;
    // New networks' buffers are filled on a stream of their own, so that an
    // add need not wait for a step call in flight.
    hipStream_t up_stream = nullptr;
    int device = 0;
    int max_nets = 0;
    int64_t budget = 0;
    BlockPool pool;                    // the slots' device blocks
    std::vector<NetSlot> slots;
    DevNet* d_nets = nullptr;
    CellBlock* d_blocks = nullptr;
    CellBlock* d_nblocks = nullptr;   // neuron blocks only
    int32_t* d_active = nullptr;
    // FP32 lanes: the FP32 networks are split into `lanes` groups of about
    // equal cell count, each with its own lists and stream, so that one
    // group's cell pass (bound by instruction issue) overlaps another's post
    // pass (bound by memory latency). The groups share no network, so
    // nothing orders them. The FP64 networks and lane 0 use `stream`.
    hipStream_t lane_stream[kMaxLanes] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[kMaxLanes] = {};
    int sched_lanes = kLanes;  // requested
    int lanes = 1;             // of the current tables
    // FP64 groups: the FP64 networks split like the FP32 lanes (as many as
    // the lane count, each on lane stream g), so that one group's cell pass
    // overlaps another's latency-bound compaction, delivery, sums and
    // drive. Group g's blocks, neuron blocks and networks are [c64[g],
    // c64[g + 1]), [nb64[g], nb64[g + 1]) and [a64[g], a64[g + 1]) of range 0.
    int lanes64 = 1;
    std::vector<int> c64, nb64, a64;
    // The window's delivery lists: neuron blocks that spiked, appended by
    // k_cells: the FP64 networks' from entry 0, counted in d_work_n[0]
    // (zeroed by the FP64 networks' k_drive); each FP32 lane's in two lists
    // by window parity p (work32), counted in d_work_n[1 + 2 lane + p]
    // (zeroed by the lane's post pass of the window before).
    CellBlock* d_work = nullptr;
    int32_t* d_work_n = nullptr;
    // The FP32 post pass's per-window state (2 per lane, by parity) and its
    // SIC items, the neuron blocks of the networks with a live SIC exchange
    // (per lane and parity, a list of the lane's neuron blocks' count), and
    // per slot the first of its neuron blocks in its lane's range of
    // d_nblocks and their count.
    PostState* d_post = nullptr;
    int32_t* d_sic_item = nullptr;
    size_t sic_item_cap = 0;
    int32_t* d_nbr = nullptr;
    std::vector<int32_t> h_nbr;
    size_t blocks_cap = 0, nblocks_cap = 0, active_cap = 0, work_cap = 0;
    size_t work_n_cap = 0, post_cap = 0;
    int n_blocks = 0, n_nblocks = 0, n_active = 0, Wg = 1;
    std::vector<int32_t> h_active;     // host copy of d_active
    bool dirty = true;
    // Batch-wide bookkeeping buffers: targets in, raster heads and overflow
    // bits out (pinned host + device, max_nets each), and the raster staging
    // buffer that one drain packs every network's new entries into.
    int64_t* h_targets = nullptr;
    int64_t* d_targets = nullptr;
    int64_t* h_heads = nullptr;
    int64_t* d_heads = nullptr;
    int32_t* h_flags = nullptr;
    int32_t* d_flags = nullptr;
    int64_t* h_offsets = nullptr;      // max_nets + 1: segment bounds of the staging buffer
    int64_t* d_offsets = nullptr;
    int64_t* h_bases = nullptr;        // max_nets: step of key offset 0 per network
    int64_t* d_bases = nullptr;
    // The drain's sort: packed keys of every entry, sorted per network on
    // the device (rocPRIM's segmented radix sort), and its scratch.
    uint64_t* d_stage_key = nullptr;
    uint64_t* d_stage_key2 = nullptr;
    uint64_t* h_stage_key = nullptr;
    void* d_sort_temp = nullptr;
    size_t sort_temp_cap = 0;
    int64_t* d_stage_step = nullptr;
    uint32_t* d_stage_src = nullptr;
    int64_t* h_stage_step = nullptr;
    uint32_t* h_stage_src = nullptr;
    size_t stage_cap = 0;       // of the key arrays
    size_t stage_pair_cap = 0;  // of the (step, source) arrays
    bool raster_dirty = false;         // steps ran since the last drain
    // scratch of order_raster
    std::vector<uint64_t> sort_keys, sort_tmp;
    // Launch schedule (results do not depend on it): cell-kernel wavefronts
    // per workgroup (1, 2, 4) and SIC group (1-16 steps, 0 = by batch size).
    int sched_cell_waves = 1;
    int sched_sic_group = 0;
    int sched_step_group = 0;
    int sched_deliver_waves = kDeliverWaves;
    int sched_sic_workers = kSicWorkers;
    float tuned_ms_per_window = 0.0f;   // the autotuner's best median
    // Window sequences replayed from captured graphs (when use_graphs and
    // not profiling); cleared whenever the tables change.
    bool use_graphs = false;
    // One graph per stream (per lane), replayed each on its own stream
    // between the fork and the join, so that the lanes overlap as eager
    // launches do.
    struct GraphEntry {
        int64_t windows;
        int Wg, G, wpb, SG, parity;  // parity of the first window
        hipGraphExec_t exec[kMaxLanes];
    };
    // Windows launched so far: FP32 windows alternate their lists by the
    // parity of this count, across calls (a graph replays the parities it
    // was captured with, so graphs are kept per starting parity).
    int64_t win_seq = 0;
    std::vector<GraphEntry> graphs;
    bool pending = false;  // async work in flight
    // Per-kernel event timing (off unless asked for).
    bool profiling = false;
    int prof_every = 1;               // time one window in prof_every
    int64_t prof_seq = 0;             // windows launched while profiling
    std::vector<hipEvent_t> events;   // 6 per timed window of the last step() call
    int64_t timed_windows = 0, pending_windows = 0;
    // [0] cells, [1] compaction, [2] delivery pass, [3] ordered sums, [4] drive
    double kernel_ms[5] = {0, 0, 0, 0, 0};
    // d_blocks holds ranges [cell_range[r], cell_range[r + 1]): r = 0 the
    // FP64 networks' blocks, r = 1 + l those of FP32 lane l (l < lanes).
    std::vector<int> cell_range = std::vector<int>(kMaxLanes + 2, 0);
    // d_nblocks and d_active likewise: neuron blocks (active slots) of the
    // FP64 networks, then of each FP32 lane.
    std::vector<int> nblock_range = std::vector<int>(kMaxLanes + 2, 0);
    std::vector<int> active_range = std::vector<int>(kMaxLanes + 2, 0);
    // The persistent launch (sched_teams > 0): the FP32 networks' groups of
    // the tables are then teams (see k_persist) rather than stream lanes,
    // with their descriptors, each workgroup's (team, member), a barrier
    // counter per team, and the failure words (device; pinned host, which a
    // timed-out barrier sets and wait() reads).
    int sched_teams = 0;
    bool persistent = false;   // the current tables are team tables
    int max_resident = 0;      // workgroups of k_persist the device holds at once
    Team* d_teams = nullptr;
    int2* d_wg_team = nullptr;
    uint32_t* d_bar = nullptr;
    int32_t* d_pfail = nullptr;
    int32_t* h_pfail = nullptr;
    int32_t* h_pfail_dev = nullptr;  // h_pfail as the device addresses it
    size_t teams_cap = 0, wg_cap = 0, bar_cap = 0;
    int n_teams = 0, persist_grid = 0;
    uint64_t team_wait_ticks = kTeamWaitTicks;

    explicit Engine(int max_n, int64_t budget_bytes) : max_nets(max_n), budget(budget_bytes) {
        HIP_CHECK(hipGetDevice(&device));
        // Streams share the device's few hardware queues in the order they
        // are created. A batch creates its upload and lane streams here; a
        // single-network engine creates them only when first needed (it
        // normally uses its main stream alone), so that it leaves the other
        // queues to other engines and several of them run side by side.
        HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        if (max_n > 1) {
            HIP_CHECK(hipStreamCreateWithFlags(&up_stream, hipStreamNonBlocking));
        }
        slots.resize(static_cast<size_t>(max_n));
        for (NetSlot& s : slots) s.pool = &pool;
        const size_t m = static_cast<size_t>(max_n);
        HIP_CHECK(hipMalloc(&d_nets, sizeof(DevNet) * m));
        HIP_CHECK(hipHostMalloc(&h_targets, sizeof(int64_t) * m, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_heads, sizeof(int64_t) * m, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_flags, sizeof(int32_t) * m, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_offsets, sizeof(int64_t) * (m + 1), hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_bases, sizeof(int64_t) * m, hipHostMallocDefault));
        HIP_CHECK(hipMalloc(&d_targets, sizeof(int64_t) * m));
        HIP_CHECK(hipMalloc(&d_heads, sizeof(int64_t) * m));
        HIP_CHECK(hipMalloc(&d_flags, sizeof(int32_t) * m));
        HIP_CHECK(hipMalloc(&d_offsets, sizeof(int64_t) * (m + 1)));
        HIP_CHECK(hipMalloc(&d_bases, sizeof(int64_t) * m));
        // The delivery list counts and the post pass's state exist for every
        // batch: k_post empties them even when no network has neurons.
        work_n_cap = 1 + 2 * kMaxLanes;
        post_cap = 2 * kMaxLanes;
        HIP_CHECK(hipMalloc(&d_work_n, work_n_cap * sizeof(int32_t)));
        HIP_CHECK(hipMemsetAsync(d_work_n, 0, work_n_cap * sizeof(int32_t), stream));
        HIP_CHECK(hipMalloc(&d_post, post_cap * sizeof(PostState)));
        HIP_CHECK(hipMemsetAsync(d_post, 0, post_cap * sizeof(PostState), stream));
        HIP_CHECK(hipMalloc(&d_pfail, sizeof(int32_t)));
        HIP_CHECK(hipMemsetAsync(d_pfail, 0, sizeof(int32_t), stream));
        HIP_CHECK(hipHostMalloc(&h_pfail, sizeof(int32_t), hipHostMallocDefault));
        *h_pfail = 0;
        HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&h_pfail_dev), h_pfail, 0));
        if (const char* us = std::getenv("ASG_TEAM_WAIT_US")) {
            team_wait_ticks = std::strtoull(us, nullptr, 10) * 100;  // 100 MHz counter
        }
        lane_stream[0] = stream;
        if (max_n > 1) {
            ensure_lane_streams(kMaxLanes);
        }
        HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        for (int l = 0; l < kMaxLanes; ++l) {
            HIP_CHECK(hipEventCreateWithFlags(&ev_join[l], hipEventDisableTiming));
        }
        HIP_CHECK(hipMalloc(&d_nbr, 2 * sizeof(int32_t) * m));
        h_nbr.assign(2 * m, 0);
        HIP_CHECK(hipMemcpyAsync(d_nbr, h_nbr.data(), 2 * sizeof(int32_t) * m,
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    ~Engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        clear_graphs();
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (NetSlot& s : slots) release(s);
        pool.flush();
        (void)hipFree(d_nets);
        (void)hipFree(d_blocks);
        (void)hipFree(d_nblocks);
        (void)hipFree(d_active);
        (void)hipFree(d_work);
        (void)hipFree(d_work_n);
        (void)hipFree(d_targets);
        (void)hipFree(d_heads);
        (void)hipFree(d_flags);
        (void)hipFree(d_offsets);
        (void)hipFree(d_bases);
        (void)hipFree(d_stage_key);
        (void)hipFree(d_stage_key2);
        (void)hipFree(d_sort_temp);
        (void)hipFree(d_post);
        (void)hipFree(d_sic_item);
        (void)hipFree(d_nbr);
        (void)hipFree(d_teams);
        (void)hipFree(d_wg_team);
        (void)hipFree(d_bar);
        (void)hipFree(d_pfail);
        (void)hipHostFree(h_pfail);
        (void)hipFree(d_stage_step);
        (void)hipFree(d_stage_src);
        (void)hipHostFree(h_targets);
        (void)hipHostFree(h_heads);
        (void)hipHostFree(h_flags);
        (void)hipHostFree(h_offsets);
        (void)hipHostFree(h_bases);
        (void)hipHostFree(h_stage_key);
        (void)hipHostFree(h_stage_step);
        (void)hipHostFree(h_stage_src);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (int l = 0; l < kMaxLanes; ++l) {
            if (ev_join[l]) (void)hipEventDestroy(ev_join[l]);
        }
        for (int l = 1; l < kMaxLanes; ++l) {
            if (lane_stream[l]) (void)hipStreamDestroy(lane_stream[l]);
        }
        if (up_stream) (void)hipStreamDestroy(up_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int64_t bytes_in_use() const {
        int64_t b = 0;
        for (const NetSlot& s : slots) {
            if (s.used) b += static_cast<int64_t>(s.bytes);
        }
        return b;
    }

    void wait() {
        if (pending) {
            HIP_CHECK(hipStreamSynchronize(stream));
            pending = false;
            if (*h_pfail != 0) {
                *h_pfail = 0;
                throw HipError{hipErrorLaunchTimeOut,
                               "persistent launch: a team barrier waited longer than 2 s and the "
                               "launch stopped; the networks' state is that of an unfinished window"};
            }
            if (pending_windows > 0) {
                // event intervals: drive, cells, compaction, delivery, sums
                constexpr int kSlot[5] = {4, 0, 1, 2, 3};
                for (int64_t w = 0; w < pending_windows; ++w) {
                    for (int k = 0; k < 5; ++k) {
                        float ms = 0.0f;
                        HIP_CHECK(hipEventElapsedTime(&ms, events[6 * w + k], events[6 * w + k + 1]));
                        kernel_ms[kSlot[k]] += ms;
                    }
                }
                timed_windows += pending_windows;
                pending_windows = 0;
            }
        }
    }

    void ensure_events(int64_t windows) {
        const size_t need = static_cast<size_t>(windows) * 6;
        while (events.size() < need) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            events.push_back(e);
        }
    }

    void upload(int slot) {
        HIP_CHECK(hipMemcpyAsync(d_nets + slot, &slots[slot].dn, sizeof(DevNet),
                                 hipMemcpyHostToDevice, stream));
    }

    // Pull the device copy of a slot's descriptor (its step, raster head and
    // counters move on the device).
    void download(int slot) {
        HIP_CHECK(hipMemcpyAsync(&slots[slot].dn, d_nets + slot, sizeof(DevNet),
                                 hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    template <typename T>
    static void grow(T*& p, size_t& cap, size_t need) {
        if (need <= cap) return;
        (void)hipFree(p);
        p = nullptr;
        cap = std::max(need, cap * 2);
        HIP_CHECK(hipMalloc(&p, cap * sizeof(T)));
    }

    // FP32 lane of every slot: the FP32 networks by cell blocks, largest
    // first, each to the lane with the fewest blocks so far. One lane while
    // profiling, so that the per-kernel events see every FP32 kernel.
    std::vector<int> assign_lanes() {
        std::vector<int> lane_of(static_cast<size_t>(max_nets), 0);
        std::vector<std::pair<int64_t, int>> w;
        for (int sidx = 0; sidx < max_nets; ++sidx) {
            const NetSlot& s = slots[sidx];
            if (!s.used || s.dn.precision != ASG_FP32) continue;
            const int64_t nb = (s.dn.n_neur + kCellBlock - 1) / kCellBlock +
                               (s.dn.n_astro + kCellBlock - 1) / kCellBlock;
            w.emplace_back(-nb, sidx);
        }
        const int n32 = static_cast<int>(w.size());
        persistent = !profiling && sched_teams > 0 && n32 > 0;
        lanes = profiling    ? 1
                : persistent ? std::max(1, std::min(sched_teams, n32))
                             : std::max(1, std::min(sched_lanes, n32));
        // the FP64 networks' groups, balanced the same way
        std::vector<std::pair<int64_t, int>> w64;
        for (int sidx = 0; sidx < max_nets; ++sidx) {
            const NetSlot& s = slots[sidx];
            if (!s.used || s.dn.precision == ASG_FP32) continue;
            const int64_t nb = (s.dn.n_neur + kCellBlock - 1) / kCellBlock +
                               (s.dn.n_astro + kCellBlock - 1) / kCellBlock;
            w64.emplace_back(-nb, sidx);
        }
        const int n64 = static_cast<int>(w64.size());
        lanes64 = profiling || persistent ? 1 : std::max(1, std::min(sched_lanes, n64));
        auto spread = [&](std::vector<std::pair<int64_t, int>>& v, int k) {
            std::sort(v.begin(), v.end());
            std::vector<int64_t> load(static_cast<size_t>(k), 0);
            for (const auto& e : v) {
                int best = 0;
                for (int l = 1; l < k; ++l) {
                    if (load[l] < load[best]) best = l;
                }
                lane_of[static_cast<size_t>(e.second)] = best;
                load[best] -= e.first;
            }
        };
        spread(w, lanes);
        spread(w64, lanes64);
        return lane_of;
    }

    // Descriptors and workgroups of the teams (the FP32 groups 1 + t of the
    // tables) for the persistent launch. The grid is what the device holds
    // at once (the occupancy API; a cooperative launch needs every
    // workgroup resident), shared among the teams by their blocks per
    // window (cell blocks plus neuron blocks, whose post items -- drive,
    // delivery, SIC -- scale with them), at least one each and no more than
    // a team has items in its larger pass. Workgroups are dispatched to the
    // eight XCDs round-robin by index, so a team takes its members from the
    // indices of one XCD while it can: its lists, rings and counters then
    // stay in one L2.
    void build_teams() {
        int on = 0, cus = 0;
        HIP_CHECK(hipDeviceGetAttribute(&on, hipDeviceAttributeCooperativeLaunch, device));
        if (!on) {
            throw HipError{hipErrorNotSupported, "the device does not support cooperative launches"};
        }
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        int per_cu = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(k_persist), 64, drive_lds_bytes()));
        max_resident = per_cu * cus;
        n_teams = lanes;
        if (n_teams > max_resident) {
            throw HipError{hipErrorInvalidValue, "more persistent teams than resident workgroups"};
        }
        std::vector<Team> tm(static_cast<size_t>(n_teams));
        std::vector<int64_t> weight(tm.size()), useful(tm.size());
        int64_t wsum = 0, usum = 0;
        for (int t = 0; t < n_teams; ++t) {
            Team& x = tm[static_cast<size_t>(t)];
            x.cell_first = cell_range[1 + t];
            x.cells = cell_range[2 + t] - cell_range[1 + t];
            x.nblock_first = nblock_range[1 + t];
            x.nblocks = nblock_range[2 + t] - nblock_range[1 + t];
            x.active_first = active_range[1 + t];
            x.nets = active_range[2 + t] - active_range[1 + t];
            x.pad = 0;
            const int post_items = (x.nets + 63) / 64 + 3 * x.nblocks;
            weight[static_cast<size_t>(t)] = x.cells + x.nblocks;
            useful[static_cast<size_t>(t)] = std::max(1, std::max(x.cells, post_items));
            wsum += weight[static_cast<size_t>(t)];
            usum += useful[static_cast<size_t>(t)];
        }
        const int64_t G = std::min<int64_t>(max_resident, usum);
        int64_t given = 0;
        for (size_t t = 0; t < tm.size(); ++t) {
            const int64_t share = wsum > 0 ? G * weight[t] / wsum : 1;
            tm[t].wgs = static_cast<int32_t>(std::max<int64_t>(1, std::min(share, useful[t])));
            given += tm[t].wgs;
        }
        // hand out what rounding left, largest weight first, while any team
        // can use more; take back any excess from the largest teams
        std::vector<int> order(tm.size());
        for (size_t t = 0; t < tm.size(); ++t) order[t] = static_cast<int>(t);
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return weight[static_cast<size_t>(a)] > weight[static_cast<size_t>(b)];
        });
        for (bool more = true; given < G && more;) {
            more = false;
            for (int t : order) {
                if (given >= G) break;
                if (tm[static_cast<size_t>(t)].wgs < useful[static_cast<size_t>(t)]) {
                    ++tm[static_cast<size_t>(t)].wgs;
                    ++given;
                    more = true;
                }
            }
        }
        while (given > max_resident) {
            for (int t : order) {
                if (given <= max_resident) break;
                if (tm[static_cast<size_t>(t)].wgs > 1) {
                    --tm[static_cast<size_t>(t)].wgs;
                    --given;
                }
            }
        }
        persist_grid = static_cast<int>(given);
        // members: indices x, x + 8, ... belong to XCD x
        constexpr int kXcds = 8;
        std::vector<std::vector<int>> free_of(kXcds);
        for (int g = persist_grid - 1; g >= 0; --g) free_of[static_cast<size_t>(g % kXcds)].push_back(g);
        std::vector<int2> wg(static_cast<size_t>(persist_grid));
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return tm[static_cast<size_t>(a)].wgs > tm[static_cast<size_t>(b)].wgs;
        });
        for (int t : order) {
            int rank = 0;
            while (rank < tm[static_cast<size_t>(t)].wgs) {
                size_t x = 0;
                for (size_t y = 1; y < free_of.size(); ++y) {
                    if (free_of[y].size() > free_of[x].size()) x = y;
                }
                while (rank < tm[static_cast<size_t>(t)].wgs && !free_of[x].empty()) {
                    wg[static_cast<size_t>(free_of[x].back())] = int2{t, rank++};
                    free_of[x].pop_back();
                }
            }
        }
        grow(d_teams, teams_cap, tm.size());
        grow(d_wg_team, wg_cap, wg.size());
        grow(d_bar, bar_cap, 16 * tm.size());
        HIP_CHECK(hipMemcpyAsync(d_teams, tm.data(), tm.size() * sizeof(Team),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d_wg_team, wg.data(), wg.size() * sizeof(int2),
                                 hipMemcpyHostToDevice, stream));
    }

    // Windows per persistent launch at most: each team counter then stays
    // below (1 + 2 * kPersistWindows) * max_resident < 2^32.
    static constexpr int64_t kPersistWindows = int64_t(1) << 16;

    // The FP32 networks' `windows` windows from window sequence number seq0
    // as persistent launches on `stream`.
    void launch_persistent(int64_t windows, int64_t seq0) {
        const int32_t nb0 = nblock_range[1];
        const int32_t dw = sched_deliver_waves, sw = sched_sic_workers, wg = Wg;
        const int32_t bars = 16 * n_teams;
        for (int64_t done = 0; done < windows; done += kPersistWindows) {
            int32_t win = static_cast<int32_t>(std::min(kPersistWindows, windows - done));
            int32_t par0 = static_cast<int32_t>((seq0 + done) & 1);
            hipLaunchKernelGGL(k_persist_prep, dim3((bars + 255) / 256), dim3(256), 0, stream,
                               d_bar, bars, d_pfail);
            void* args[] = {&d_nets, &d_blocks, &d_nblocks, &d_teams, &d_wg_team, &d_active,
                            &d_nbr, &d_work, &d_work_n, &d_post, &d_sic_item,
                            const_cast<int32_t*>(&nb0), const_cast<int32_t*>(&wg), &win, &par0,
                            const_cast<int32_t*>(&dw), const_cast<int32_t*>(&sw), &d_bar,
                            &d_pfail, &h_pfail_dev, &team_wait_ticks};
            HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_persist),
                                                 dim3(persist_grid), dim3(64), args,
                                                 static_cast<unsigned>(drive_lds_bytes()), stream));
        }
    }

    void rebuild_tables() {
        std::vector<CellBlock> blocks, nblocks;
        std::vector<int32_t> active;
        Wg = kWindowCap;
        std::fill(h_nbr.begin(), h_nbr.end(), 0);
        const std::vector<int> lane_of = assign_lanes();
        cell_range.assign(static_cast<size_t>(lanes) + 2, 0);
        nblock_range.assign(static_cast<size_t>(lanes) + 2, 0);
        active_range.assign(static_cast<size_t>(lanes) + 2, 0);
        // Blocks in contiguous ranges, one per cell launch: the FP64
        // networks' blocks, then each FP32 lane's. A range lists all its
        // astrocyte blocks before its neuron blocks: an astrocyte wavefront
        // runs longer than a neuron one (a counter-stream draw every step),
        // and a launch of more wavefronts than the device holds at once
        // finishes sooner with the longest dispatched first (on the
        // 256-network sweep, FP32 42.8 against 45.0 us per window, FP64
        // 113.3 against 114.2). Each network's first neuron block (its
        // first astrocyte block if it has no neurons) is its lead block.
        // The FP64 range is its groups' blocks one group after another.
        c64.assign(static_cast<size_t>(lanes64) + 1, 0);
        nb64.assign(static_cast<size_t>(lanes64) + 1, 0);
        a64.assign(static_cast<size_t>(lanes64) + 1, 0);
        for (int range = 0; range < 1 + lanes; ++range) {
          const int groups = range == 0 ? lanes64 : 1;
          for (int grp = 0; grp < groups; ++grp) {
            std::vector<CellBlock> neuron_part;  // after the astrocyte blocks
            for (int sidx = 0; sidx < max_nets; ++sidx) {
                const NetSlot& s = slots[sidx];
                if (!s.used || (s.dn.precision == ASG_FP32) != (range > 0)) continue;
                if (lane_of[static_cast<size_t>(sidx)] != (range > 0 ? range - 1 : grp)) continue;
                active.push_back(sidx);
                Wg = std::min(Wg, s.W);
                if (range > 0) {
                    h_nbr[2 * sidx] = static_cast<int32_t>(nblocks.size()) - nblock_range[range];
                    h_nbr[2 * sidx + 1] =
                        static_cast<int32_t>((s.dn.n_neur + kCellBlock - 1) / kCellBlock);
                }
                uint32_t lead = 1;
                for (uint32_t f = 0; f < s.dn.n_neur; f += kCellBlock) {
                    neuron_part.push_back(CellBlock{sidx, 0, f, lead});
                    nblocks.push_back(CellBlock{sidx, 0, f, lead});
                    lead = 0;
                    // FP32: the block where the inhibitory neurons begin
                    // gets a second cell wavefront for them (kind 2)
                    const uint32_t ne = s.dn.n_exc;
                    if (range > 0 && f < ne && f + kCellBlock > ne && ne < s.dn.n_neur) {
                        neuron_part.push_back(CellBlock{sidx, 2, f, 0});
                    }
                }
                for (uint32_t f = 0; f < s.dn.n_astro; f += kCellBlock) {
                    blocks.push_back(CellBlock{sidx, 1, f, lead});
                    lead = 0;
                }
            }
            blocks.insert(blocks.end(), neuron_part.begin(), neuron_part.end());
            if (range == 0) {
                c64[grp + 1] = static_cast<int>(blocks.size());
                nb64[grp + 1] = static_cast<int>(nblocks.size());
                a64[grp + 1] = static_cast<int>(active.size());
            }
          }
            cell_range[range + 1] = static_cast<int>(blocks.size());
            nblock_range[range + 1] = static_cast<int>(nblocks.size());
            active_range[range + 1] = static_cast<int>(active.size());
        }
        n_blocks = static_cast<int>(blocks.size());
        n_nblocks = static_cast<int>(nblocks.size());
        n_active = static_cast<int>(active.size());
        h_active = active;
        if (n_blocks > 0) {
            grow(d_blocks, blocks_cap, blocks.size());
            HIP_CHECK(hipMemcpyAsync(d_blocks, blocks.data(), blocks.size() * sizeof(CellBlock),
                                     hipMemcpyHostToDevice, stream));
        }
        if (n_nblocks > 0) {
            grow(d_nblocks, nblocks_cap, nblocks.size());
            HIP_CHECK(hipMemcpyAsync(d_nblocks, nblocks.data(), nblocks.size() * sizeof(CellBlock),
                                     hipMemcpyHostToDevice, stream));
            // each neuron block is listed at most once per window (FP32: in
            // its lane's lists of the window's parity)
            grow(d_work, work_cap, nblocks.size() + (nblocks.size() - nblock_range[1]));
            grow(d_sic_item, sic_item_cap, 2 * (nblocks.size() - nblock_range[1]));
        }
        if (n_active > 0) {
            grow(d_active, active_cap, active.size());
            HIP_CHECK(hipMemcpyAsync(d_active, active.data(), active.size() * sizeof(int32_t),
                                     hipMemcpyHostToDevice, stream));
        }
        HIP_CHECK(hipMemcpyAsync(d_nbr, h_nbr.data(), h_nbr.size() * sizeof(int32_t),
                                 hipMemcpyHostToDevice, stream));
        // Every list starts empty under the new tables. A lane's post pass
        // empties the lists of the parity after its own, so after a call the
        // lists of its last window's parity still hold that window's counts;
        // that is harmless while the lane runs on (its next window has the
        // other parity), not for a lane that stops and later runs again on
        // other networks: its stale entries would name blocks of networks
        // that may be gone, and its appends would run past its lists.
        grow(d_work_n, work_n_cap, 1 + 2 * static_cast<size_t>(lanes) + (lanes64 - 1));
        grow(d_post, post_cap, 2 * static_cast<size_t>(lanes));
        if (persistent) build_teams();
        HIP_CHECK(hipMemsetAsync(d_work_n, 0, work_n_cap * sizeof(int32_t), stream));
        HIP_CHECK(hipMemsetAsync(d_post, 0, post_cap * sizeof(PostState), stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        dirty = false;
        clear_graphs();
    }

    int add(const asg_net_desc& d, int64_t total, const Prepared* prep = nullptr) {
        int free_slot = -1;
        for (int i = 0; i < max_nets; ++i) {
            if (!slots[i].used) {
                free_slot = i;
                break;
            }
        }
        if (free_slot < 0) return -1;
        const DerivedConn* none = nullptr;
        const Plan pl = prep != nullptr ? prep->pl : plan_of(d, none);
        // A slot counts the rounded sizes it asks the pool for (a reused
        // block can be larger; the pool keeps what it caches bounded).
        const int64_t need = static_cast<int64_t>(BlockPool::rounded(pl.bytes));
        if (budget > 0 && bytes_in_use() + need > budget) return -2;
        // No wait for a step call in flight: the free slot is in none of the
        // tables it runs on (a removed slot's last call was drained by the
        // remove), so its device descriptor and buffers are not read until
        // the next table rebuild, by which time the uploads have finished.
        NetSlot& s = slots[free_slot];
        // uploads go on the upload stream, so that they overlap a step call
        // in flight; a single-network engine without one uploads on its
        // main stream while idle and creates one otherwise
        if (pending && up_stream == nullptr) {
            HIP_CHECK(hipStreamCreateWithFlags(&up_stream, hipStreamNonBlocking));
        }
        hipStream_t up = up_stream != nullptr ? up_stream : stream;
        try {
            build_slot(s, d, total, up, prep);
            HIP_CHECK(hipMemcpyAsync(d_nets + free_slot, &s.dn, sizeof(DevNet),
                                     hipMemcpyHostToDevice, up));
            HIP_CHECK(hipStreamSynchronize(up));
        } catch (...) {
            (void)hipStreamSynchronize(up);
            release(s);
            throw;
        }
        dirty = true;
        return free_slot;
    }

    void remove(int slot) {
        // Drain first: the active list the drain walks still holds this slot
        // until the tables are rebuilt, and its buffers are about to go.
        drain_all();
        release(slots[slot]);
        dirty = true;
    }

    // k_deliver of the FP64 networks for SIC group size G (its accumulators
    // are registers).
    template <int G>
    void launch_deliver_g(int32_t Wg_, int MG, int grp) {
        const int count = nb64[grp + 1] - nb64[grp];
        if (count == 0) return;
        // delivery wavefronts (before the SIC ones) over the listed blocks
        const int groups = (Wg_ + MG - 1) / MG;
        const int D = std::min(count * groups, sched_deliver_waves);
        const int items = count * ((Wg_ + G - 1) / G) + D;
        hipLaunchKernelGGL(k_deliver<G>, dim3((items + 3) / 4), dim3(256), 0, lane_stream[grp],
                           d_nets, d_nblocks + nb64[grp], items, Wg_, D, MG, d_work + nb64[grp],
                           work_n64(grp));
    }

    // Pass 3 of FP64 group grp: marking in groups of MG steps and the SIC
    // sums of every block.
    void launch_deliver64(int G, int32_t Wg_, int MG, int grp) {
        switch (G) {
            case 1: launch_deliver_g<1>(Wg_, MG, grp); break;
            case 2: launch_deliver_g<2>(Wg_, MG, grp); break;
            case 4: launch_deliver_g<4>(Wg_, MG, grp); break;
            case 8: launch_deliver_g<8>(Wg_, MG, grp); break;
            default: launch_deliver_g<16>(Wg_, MG, grp); break;
        }
    }

    // FP64 group grp's delivery list count (its list is d_work + nb64[grp]).
    int32_t* work_n64(int grp) const {
        return grp == 0 ? d_work_n : d_work_n + 1 + 2 * lanes + (grp - 1);
    }

    // FP64 group grp's window launches (on lane stream grp): the drive
    // (whose first block empties the group's list count), the cell pass,
    // the compaction of its spike words and its ordered sums.
    void launch_drive64(int grp) {
        const int count = nb64[grp + 1] - nb64[grp];
        if (count == 0) return;
        hipLaunchKernelGGL(k_drive, dim3(count), dim3(64), drive_lds_bytes(), lane_stream[grp],
                           d_nets, d_nblocks + nb64[grp], count, Wg, work_n64(grp));
    }
    void launch_cells64(int grp, int32_t Wg_, int wpb, int p) {
        // post / sic_item are not used by the FP64 cell pass
        launch_cells_k<ASG_FP64>(d_blocks + c64[grp], c64[grp + 1] - c64[grp], Wg_, wpb, p,
                                 d_work + nb64[grp], work_n64(grp), d_post, d_sic_item,
                                 lane_stream[grp]);
    }
    void launch_compact64(int grp) {
        const int n = a64[grp + 1] - a64[grp];
        if (n == 0) return;
        hipLaunchKernelGGL(k_compact, dim3(n, Wg), dim3(64), 0, lane_stream[grp], d_nets,
                           d_active + a64[grp]);
    }
    void launch_sum64(int grp, int SG) {
        const int items = (c64[grp + 1] - c64[grp]) * ((Wg + SG - 1) / SG);
        if (items == 0) return;
        hipLaunchKernelGGL(k_sum, dim3((items + 3) / 4), dim3(256), 0, lane_stream[grp], d_nets,
                           d_blocks + c64[grp], items, Wg, SG);
    }

    size_t drive_lds_bytes() const {
        return static_cast<size_t>(64 * Wg) * (sizeof(double) + sizeof(uint32_t));
    }

    // The stream of range `range` (0 the FP64 networks, 1 + l FP32 lane l).
    hipStream_t range_stream(int range) const {
        return range == 0 ? stream : lane_stream[range - 1];
    }

    // The drive of the call's first window of FP32 lane `range` - 1 as a
    // launch of its own (later windows are driven by the post pass).
    void launch_drive(int range) {
        const int count = nblock_range[range + 1] - nblock_range[range];
        if (count == 0) return;
        hipLaunchKernelGGL(k_drive, dim3(count), dim3(64), drive_lds_bytes(), range_stream(range),
                           d_nets, d_nblocks + nblock_range[range], count, Wg, nullptr);
    }

    // FP32 lane l's delivery list, its count, the post pass's state and the
    // SIC items of window parity p.
    int lane_nblocks(int l) const { return nblock_range[2 + l] - nblock_range[1 + l]; }
    CellBlock* work32(int l, int p) const {
        return d_work + nblock_range[1] + 2 * (nblock_range[1 + l] - nblock_range[1]) +
               p * lane_nblocks(l);
    }
    int32_t* work_n32(int l, int p) const { return d_work_n + 1 + 2 * l + p; }
    PostState* post32(int l, int p) const { return d_post + 2 * l + p; }
    int32_t* sic_items32(int l, int p) const {
        return d_sic_item + 2 * (nblock_range[1 + l] - nblock_range[1]) + p * lane_nblocks(l);
    }

    // Pass 3 of FP32 lane l with the window's close and the next window's
    // drive (window parity p).
    void launch_post(int l, int p) {
        const int n32 = active_range[2 + l] - active_range[1 + l];
        if (n32 == 0) return;
        const int count = lane_nblocks(l);
        const int C = (n32 + 63) / 64;
        const int D = std::min(count, sched_deliver_waves);
        // SIC wavefronts: by default one per neuron block, so that every
        // live block has one however many networks are live (those beyond
        // the window's items exit at once)
        const int E = sched_sic_workers > 0 ? std::min(count, sched_sic_workers) : count;
        hipLaunchKernelGGL(k_post, dim3(C + D + E + count), dim3(64), drive_lds_bytes(),
                           lane_stream[l], d_nets, d_nblocks + nblock_range[1 + l], count, Wg, C,
                           D, E, work32(l, p), work_n32(l, p), post32(l, p), sic_items32(l, p),
                           d_active + active_range[1 + l], n32, p, work_n32(l, p ^ 1),
                           post32(l, p ^ 1));
    }

    // Window steps per marking / summing wavefront: as requested, or (0) the
    // whole window once per-step wavefronts would number beyond kGroupWaves.
    int step_group_for(int requested) const {
        if (requested > 0) return std::min(requested, Wg);
        return static_cast<int64_t>(n_blocks) * Wg > kGroupWaves ? Wg : 1;
    }

    // One cell-pass launch: `count` blocks from `blocks` (window parity p),
    // appending the blocks that spiked to work / *work_n, on stream st.
    template <int P>
    void launch_cells_k(const CellBlock* blocks, int count, int32_t Wg_, int wpb, int p,
                        CellBlock* work, int32_t* work_n, PostState* post, int32_t* sic_item,
                        hipStream_t st) {
        if (count == 0) return;
        const dim3 grid((count + wpb - 1) / wpb);
        switch (wpb) {
            case 4:
                hipLaunchKernelGGL((k_cells<P, 4>), grid, dim3(4 * kCellBlock), 0, st, d_nets,
                                   blocks, count, Wg_, work, work_n, p, post, sic_item, d_nbr);
                break;
            case 2:
                hipLaunchKernelGGL((k_cells<P, 2>), grid, dim3(2 * kCellBlock), 0, st, d_nets,
                                   blocks, count, Wg_, work, work_n, p, post, sic_item, d_nbr);
                break;
            default:
                hipLaunchKernelGGL((k_cells<P, 1>), grid, dim3(kCellBlock), 0, st, d_nets,
                                   blocks, count, Wg_, work, work_n, p, post, sic_item, d_nbr);
                break;
        }
    }

    // The cell pass of FP32 lane l (window parity p).
    void launch_cells32(int l, int32_t Wg_, int wpb, int p) {
        launch_cells_k<ASG_FP32>(d_blocks + cell_range[1 + l], cell_range[2 + l] - cell_range[1 + l],
                                 Wg_, wpb, p, work32(l, p), work_n32(l, p), post32(l, p),
                                 sic_items32(l, p), lane_stream[l]);
    }

    // The SIC group the schedule asks for, or (0) the largest power of two
    // up to the window that still leaves kSicWaves wavefronts.
    int sic_group_for(int requested) const {
        int top = 1;
        while (top < Wg) top *= 2;
        if (requested > 0) return std::min(requested, top);
        int g = top;
        while (g > 1 && static_cast<int64_t>(n_nblocks) * ((Wg + g - 1) / g) < kSicWaves) {
            g /= 2;
        }
        return g;
    }

    // Lanes beyond the first wait for what `stream` holds so far (fork), and
    // `stream` for what they hold (join).
    // Streams of the FP32 lanes: none beyond `stream` under persistent
    // launches, whose groups are teams within one launch.
    int stream_lanes() const { return persistent ? 1 : std::max(lanes, lanes64); }
    // Creates the lane streams below `count` that do not exist yet.
    void ensure_lane_streams(int count) {
        for (int l = 1; l < count; ++l) {
            if (lane_stream[l] == nullptr) {
                HIP_CHECK(hipStreamCreateWithFlags(&lane_stream[l], hipStreamNonBlocking));
            }
        }
    }
    void fork_lanes() {
        if (stream_lanes() > 1) {
            HIP_CHECK(hipEventRecord(ev_fork, stream));
            for (int l = 1; l < stream_lanes(); ++l) {
                HIP_CHECK(hipStreamWaitEvent(lane_stream[l], ev_fork, 0));
            }
        }
    }
    void join_lanes() {
        for (int l = 1; l < stream_lanes(); ++l) {
            HIP_CHECK(hipEventRecord(ev_join[l], lane_stream[l]));
            HIP_CHECK(hipStreamWaitEvent(stream, ev_join[l], 0));
        }
    }

    // The launches of `windows` windows, the first of sequence number seq0
    // (eager, or recorded into a graph when the streams are capturing):
    // every stream's (only < 0), or only those of lane stream `only` (FP32
    // lane `only` and FP64 group `only`). With `marks`, one window in prof_every
    // is bracketed by events on `stream` (6 per timed window). Returns the
    // windows timed.
    //
    // FP64 groups: drive, cells, compaction, delivery (marking + SIC),
    // ordered sums (whose lead wavefront closes the window). FP32 lanes:
    // cells, then k_post (delivery, SIC of the live networks, the next
    // window's drive, close); the first window's drive is launched before
    // the loop. Each lane's launches go to its own stream.
    int64_t launch_windows(int64_t windows, int G, int wpb, int SG, bool marks, int64_t seq0,
                           int only = -1, bool fp32 = true) {
        int64_t timed = 0;
        if (windows <= 0) {
            return 0;
        }
        auto on64 = [&](int g) { return g < lanes64 && (only < 0 || only == g); };
        auto on = [&](int l) { return fp32 && (only < 0 || only == l); };
        for (int l = 0; l < lanes; ++l) {
            if (on(l)) launch_drive(1 + l);
        }
        for (int64_t wdx = 0; wdx < windows; ++wdx) {
            const int p = static_cast<int>((seq0 + wdx) & 1);  // the window's parity
            const bool timed_here = marks && (prof_seq++ % prof_every) == 0;
            auto mark = [&](int k) {
                if (timed_here) HIP_CHECK(hipEventRecord(events[6 * timed + k], stream));
            };
            mark(0);
            for (int g = 0; g < lanes64; ++g) {
                if (on64(g)) launch_drive64(g);
            }
            mark(1);
            for (int g = 0; g < lanes64; ++g) {
                if (on64(g)) launch_cells64(g, Wg, wpb, p);
            }
            for (int l = 0; l < lanes; ++l) {
                if (on(l)) launch_cells32(l, Wg, wpb, p);
            }
            mark(2);
            for (int g = 0; g < lanes64; ++g) {
                if (on64(g)) launch_compact64(g);
            }
            mark(3);
            for (int g = 0; g < lanes64; ++g) {
                if (on64(g)) launch_deliver64(G, Wg, SG, g);
            }
            for (int l = 0; l < lanes; ++l) {
                if (on(l)) launch_post(l, p);
            }
            mark(4);
            for (int g = 0; g < lanes64; ++g) {
                if (on64(g)) launch_sum64(g, SG);
            }
            mark(5);
            if (timed_here) ++timed;
        }
        return timed;
    }

    // Graphs of `windows` windows from the current window sequence number,
    // for the current tables and schedule, captured on first use.
    const GraphEntry& graph_for(int64_t windows, int G, int wpb, int SG) {
        const int parity = static_cast<int>(win_seq & 1);
        for (const GraphEntry& g : graphs) {
            if (g.windows == windows && g.G == G && g.wpb == wpb && g.SG == SG && g.Wg == Wg &&
                g.parity == parity) {
                return g;
            }
        }
        GraphEntry ge{windows, Wg, G, wpb, SG, parity, {}};
        try {
            for (int l = 0; l < stream_lanes(); ++l) {
                hipGraph_t graph = nullptr;
                HIP_CHECK(hipStreamBeginCapture(lane_stream[l], hipStreamCaptureModeThreadLocal));
                launch_windows(windows, G, wpb, SG, false, win_seq, l);
                HIP_CHECK(hipStreamEndCapture(lane_stream[l], &graph));
                const hipError_t e = hipGraphInstantiate(&ge.exec[l], graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_CHECK(e);
            }
        } catch (...) {
            for (hipGraphExec_t x : ge.exec) {
                if (x) (void)hipGraphExecDestroy(x);
            }
            throw;
        }
        graphs.push_back(ge);
        return graphs.back();
    }

    void clear_graphs() {
        for (GraphEntry& g : graphs) {
            for (hipGraphExec_t x : g.exec) {
                if (x) (void)hipGraphExecDestroy(x);
            }
        }
        graphs.clear();
    }

    // Advance every used slot by up to n steps (to its total).
    void step(int64_t n, bool async) {
        wait();
        if (n <= 0) return;
        if (dirty) rebuild_tables();
        if (n_active == 0) return;
        ensure_lane_streams(stream_lanes());
        int64_t windows = 0;
        // The run target is the only descriptor field the host changes
        // between calls: all of them go down in one copy and one launch.
        for (int j = 0; j < n_active; ++j) {
            NetSlot& s = slots[h_active[j]];
            const int64_t target = std::min(s.total, s.step + n);
            s.dn.target = target;
            h_targets[j] = target;
            const int64_t rem = target - s.step;
            windows = std::max(windows, (rem + Wg - 1) / Wg);
            s.step = target;
        }
        HIP_CHECK(hipMemcpyAsync(d_targets, h_targets, sizeof(int64_t) * n_active,
                                 hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_set_targets, dim3((n_active + 255) / 256), dim3(256), 0, stream,
                           d_nets, d_active, n_active, d_targets);
        raster_dirty = true;
        const int G = sic_group_for(sched_sic_group);
        const int SG = step_group_for(sched_step_group);
        if (persistent) {
            // the FP64 networks of a mixed batch by their kernels, then the
            // FP32 teams in one launch
            launch_windows(windows, G, sched_cell_waves, SG, false, win_seq, -1, false);
            launch_persistent(windows, win_seq);
        } else if (profiling) {
            // per-kernel events need eager launches
            ensure_events(windows);
            fork_lanes();
            pending_windows = launch_windows(windows, G, sched_cell_waves, SG, true, win_seq);
            join_lanes();
        } else if (use_graphs && windows > 0) {
            const GraphEntry& g = graph_for(windows, G, sched_cell_waves, SG);
            fork_lanes();
            for (int l = 0; l < stream_lanes(); ++l) {
                HIP_CHECK(hipGraphLaunch(g.exec[l], lane_stream[l]));
            }
            join_lanes();
        } else {
            fork_lanes();
            launch_windows(windows, G, sched_cell_waves, SG, false, win_seq);
            join_lanes();
        }
        win_seq += windows;
        HIP_CHECK(hipGetLastError());
        pending = true;
        if (!async) wait();
    }

    // Autotuner: one knob at a time -- cell wavefronts per workgroup; then,
    // with FP64 networks, their SIC group and step group; then, with FP32
    // networks, the delivery wavefront count -- each candidate run
    // as step calls of `wpt` windows, `rounds` times round-robin so slow
    // drifts in activity hit all alike, keeping the smallest median. The
    // schedule changes launch geometry only, never results, so tuning
    // advances the simulation like any other call.
    struct Sched {
        int wpb, g, sg, dw, sw, lanes, teams;
    };
    Sched current_schedule() const {
        return Sched{sched_cell_waves, sched_sic_group, sched_step_group, sched_deliver_waves,
                     sched_sic_workers, sched_lanes, sched_teams};
    }
    void apply_schedule(const Sched& c) {
        sched_cell_waves = c.wpb;
        sched_sic_group = c.g;
        sched_step_group = c.sg;
        sched_deliver_waves = c.dw;
        sched_sic_workers = c.sw;
        set_lanes(c.lanes);
        set_teams(c.teams);
    }
    // A new lane or team count takes effect with the next table rebuild.
    void set_lanes(int l) {
        if (l != sched_lanes) {
            sched_lanes = l;
            dirty = true;
        }
    }
    void set_teams(int t) {
        if (t != sched_teams) {
            sched_teams = t;
            dirty = true;
        }
    }

    // Times every candidate; applies and returns the fastest (median ms per
    // call of wpt windows).
    float tune_among(const std::vector<Sched>& cand, int wpt, int rounds, hipEvent_t e0,
                     hipEvent_t e1) {
        std::vector<std::vector<float>> times(cand.size());
        for (int r = 0; r < rounds; ++r) {
            for (size_t c = 0; c < cand.size(); ++c) {
                apply_schedule(cand[c]);
                if (dirty) {
                    // new tables (and graphs): built by an untimed call
                    step(static_cast<int64_t>(wpt) * Wg, true);
                }
                HIP_CHECK(hipEventRecord(e0, stream));
                step(static_cast<int64_t>(wpt) * Wg, true);
                HIP_CHECK(hipEventRecord(e1, stream));
                wait();
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                times[c].push_back(ms);
            }
        }
        size_t best = 0;
        float best_t = 0.0f;
        for (size_t c = 0; c < cand.size(); ++c) {
            std::vector<float> v = times[c];
            std::sort(v.begin(), v.end());
            const float med = v[v.size() / 2];
            if (c == 0 || med < best_t) {
                best = c;
                best_t = med;
            }
        }
        apply_schedule(cand[best]);
        return best_t;
    }

    void autotune(int wpt, int rounds) {
        wait();
        if (dirty) rebuild_tables();
        if (n_active == 0 || wpt <= 0 || rounds <= 0) return;
        int top = 1;
        while (top < Wg) top *= 2;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        const bool prof = profiling;
        profiling = false;
        float best_t = 0.0f;
        try {
            std::vector<Sched> cand;
            const int n32 = active_range[1 + lanes] - active_range[1];
            if (n32 > 0) {
                // how the FP32 windows are launched: per-kernel launches on
                // 1..kMaxLanes streams, or one persistent launch of teams
                for (int l = 1; l <= std::min(kMaxLanes, n32); ++l) {
                    Sched c = current_schedule();
                    c.lanes = l;
                    c.teams = 0;
                    cand.push_back(c);
                }
                for (int t : {4, 16, 64, 256}) {
                    if (t > n32 && t != 4) continue;
                    Sched c = current_schedule();
                    c.teams = std::min(t, n32);
                    cand.push_back(c);
                }
                best_t = tune_among(cand, wpt, rounds, e0, e1);
                cand.clear();
            }
            for (int wpb : {1, 2, 4}) {
                Sched c = current_schedule();
                c.wpb = wpb;
                cand.push_back(c);
            }
            best_t = tune_among(cand, wpt, rounds, e0, e1);
            if (active_range[1] > 0) {  // FP64 networks
                cand.clear();
                for (int g = 1; g <= top; g *= 2) {
                    for (int sg : {1, Wg}) {
                        Sched c = current_schedule();
                        c.g = g;
                        c.sg = sg;
                        cand.push_back(c);
                    }
                }
                best_t = tune_among(cand, wpt, rounds, e0, e1);
            }
            if (active_range[1 + lanes] > active_range[1]) {  // FP32 networks
                cand.clear();
                // Not the SIC wavefront count: the SIC load grows with the
                // activity (networks whose astrocytes pass the threshold),
                // so a count tuned early in a run can starve it later; the
                // default gives every neuron block a wavefront.
                for (int dw : {512, 1024, 2048}) {
                    Sched c = current_schedule();
                    c.dw = dw;
                    cand.push_back(c);
                }
                best_t = tune_among(cand, wpt, rounds, e0, e1);
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            profiling = prof;
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        profiling = prof;
        tuned_ms_per_window = best_t / static_cast<float>(wpt);
    }

    // Move the device rasters of all networks into their host queues: one
    // launch + copy for the heads and overflow bits, one launch packing every
    // network's new entries into the staging buffer, one copy back.
    // Grows a device array and its pinned host copy to `count` elements.
    template <typename T>
    static void grow_pair(T*& dev, T*& host, size_t count) {
        (void)hipFree(dev);
        (void)hipHostFree(host);
        dev = nullptr;
        host = nullptr;
        HIP_CHECK(hipMalloc(&dev, count * sizeof(T)));
        HIP_CHECK(hipHostMalloc(&host, count * sizeof(T), hipHostMallocDefault));
    }

    // Moves every active network's new raster entries into its host queue,
    // in the reference's order (step, then source). The entries are packed
    // as keys (step offset from the step the network was last drained to,
    // then source) and sorted per network on the device, so the host only
    // unpacks them; FP32 delivery appends a window's entries in the order
    // its wavefronts reserve them, and this is where the order is restored.
    // Should a network have run 2^32 steps since its last drain, the pairs
    // come over as they are and the host sorts them instead.
    void drain_all() {
        wait();
        if (!raster_dirty || n_active == 0) return;
        hipLaunchKernelGGL(k_raster_heads, dim3((n_active + 255) / 256), dim3(256), 0, stream,
                           d_nets, d_active, n_active, d_heads, d_flags);
        HIP_CHECK(hipMemcpyAsync(h_heads, d_heads, sizeof(int64_t) * n_active,
                                 hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(h_flags, d_flags, sizeof(int32_t) * n_active,
                                 hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        int64_t total = 0, most = 0, span = 1;
        for (int j = 0; j < n_active; ++j) {
            NetSlot& s = slots[h_active[j]];
            h_offsets[j] = total;
            total += h_heads[j];
            most = std::max(most, h_heads[j]);
            s.overflow = h_flags[j];
            h_bases[j] = s.drained_to;
            span = std::max(span, s.step - s.drained_to);
            s.drained_to = s.step;
        }
        h_offsets[n_active] = total;
        raster_dirty = false;
        if (total == 0) return;
        const bool keyed = span < (int64_t(1) << 32) && total < (int64_t(1) << 32);
        if (keyed && static_cast<size_t>(total) > stage_cap) {
            const size_t cap = std::max(static_cast<size_t>(total), 2 * stage_cap);
            stage_cap = 0;
            grow_pair(d_stage_key, h_stage_key, cap);
            (void)hipFree(d_stage_key2);
            d_stage_key2 = nullptr;
            HIP_CHECK(hipMalloc(&d_stage_key2, cap * sizeof(uint64_t)));
            stage_cap = cap;
        }
        if (!keyed && static_cast<size_t>(total) > stage_pair_cap) {
            const size_t cap = std::max(static_cast<size_t>(total), 2 * stage_pair_cap);
            stage_pair_cap = 0;
            grow_pair(d_stage_step, h_stage_step, cap);
            grow_pair(d_stage_src, h_stage_src, cap);
            stage_pair_cap = cap;
        }
        HIP_CHECK(hipMemcpyAsync(d_offsets, h_offsets, sizeof(int64_t) * (n_active + 1),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d_bases, h_bases, sizeof(int64_t) * n_active,
                                 hipMemcpyHostToDevice, stream));
        const int64_t gx = std::min<int64_t>(64, (most + 255) / 256);
        hipLaunchKernelGGL(k_raster_gather, dim3(static_cast<uint32_t>(gx), n_active), dim3(256),
                           0, stream, d_nets, d_active, d_heads, d_offsets, d_stage_step,
                           d_stage_src, d_bases, keyed ? d_stage_key : nullptr);
        HIP_CHECK(hipGetLastError());
        if (keyed) {
            int bits = 1;
            while (bits < 32 && (span >> bits) != 0) ++bits;
            const unsigned end_bit = 32u + static_cast<unsigned>(bits);
            const unsigned items = static_cast<unsigned>(total);
            const unsigned segs = static_cast<unsigned>(n_active);
            size_t need = 0;
            HIP_CHECK(rocprim::segmented_radix_sort_keys(nullptr, need, d_stage_key, d_stage_key2,
                                                         items, segs, d_offsets, d_offsets + 1,
                                                         0u, end_bit, stream));
            if (need > sort_temp_cap) {
                (void)hipFree(d_sort_temp);
                d_sort_temp = nullptr;
                sort_temp_cap = 0;
                HIP_CHECK(hipMalloc(&d_sort_temp, need));
                sort_temp_cap = need;
            }
            HIP_CHECK(rocprim::segmented_radix_sort_keys(d_sort_temp, need, d_stage_key,
                                                         d_stage_key2, items, segs, d_offsets,
                                                         d_offsets + 1, 0u, end_bit, stream));
            HIP_CHECK(hipMemcpyAsync(h_stage_key, d_stage_key2, sizeof(uint64_t) * total,
                                     hipMemcpyDeviceToHost, stream));
        } else {
            HIP_CHECK(hipMemcpyAsync(h_stage_step, d_stage_step, sizeof(int64_t) * total,
                                     hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipMemcpyAsync(h_stage_src, d_stage_src, sizeof(uint32_t) * total,
                                     hipMemcpyDeviceToHost, stream));
        }
        HIP_CHECK(hipStreamSynchronize(stream));
        for (int j = 0; j < n_active; ++j) {
            const int64_t nr = h_heads[j];
            if (nr == 0) continue;
            NetSlot& s = slots[h_active[j]];
            const size_t at = s.q_step.size();
            s.q_step.resize(at + static_cast<size_t>(nr));
            s.q_src.resize(at + static_cast<size_t>(nr));
            int64_t* qs = s.q_step.data() + at;
            uint32_t* qr = s.q_src.data() + at;
            if (keyed) {
                const uint64_t* k = h_stage_key + h_offsets[j];
                const int64_t base = h_bases[j];
                for (int64_t e = 0; e < nr; ++e) {
                    qs[e] = base + static_cast<int64_t>(k[e] >> 32);
                    qr[e] = static_cast<uint32_t>(k[e]);
                }
            } else {
                std::memcpy(qs, h_stage_step + h_offsets[j], sizeof(int64_t) * nr);
                std::memcpy(qr, h_stage_src + h_offsets[j], sizeof(uint32_t) * nr);
                order_raster(qs, qr, static_cast<size_t>(nr), sort_keys, sort_tmp);
            }
        }
    }

    // 0, or the overflow bits of a slot (see DevNet::overflow).
    int overflow_bits(int slot) {
        drain_all();
        return slots[slot].overflow;
    }

    // OR of the overflow bits of every network.
    int overflow_bits_any() {
        drain_all();
        int bits = 0;
        for (const NetSlot& s : slots) {
            if (s.used) bits |= s.overflow;
        }
        return bits;
    }

    int64_t spikes_out(int slot, int64_t* step, uint32_t* src, int64_t cap) {
        drain_all();
        return take_queued(slot, step, src, cap);
    }

    // Every slot's queued pairs, back to back in slot order (counts per slot).
    int64_t spikes_out_all(int64_t* counts, int64_t* step, uint32_t* src, int64_t cap) {
        drain_all();
        int64_t done = 0;
        for (int s = 0; s < max_nets; ++s) {
            counts[s] = 0;
            if (slots[s].used && done < cap) {
                counts[s] = take_queued(s, step + done, src + done, cap - done);
                done += counts[s];
            }
        }
        return done;
    }

    // Up to cap of a slot's host-queued pairs, oldest first.
    int64_t take_queued(int slot, int64_t* step, uint32_t* src, int64_t cap) {
        NetSlot& s = slots[slot];
        const int64_t avail = static_cast<int64_t>(s.q_step.size() - s.q_pos);
        const int64_t m = std::min(avail, cap);
        if (m > 0) {
            std::memcpy(step, s.q_step.data() + s.q_pos, m * sizeof(int64_t));
            std::memcpy(src, s.q_src.data() + s.q_pos, m * sizeof(uint32_t));
            s.q_pos += static_cast<size_t>(m);
        }
        if (s.q_pos == s.q_step.size()) {
            s.q_step.clear();
            s.q_src.clear();
            s.q_pos = 0;
        }
        return m;
    }

    void set_recording(int slot, int64_t start, int32_t every, uint32_t na, int64_t cap) {
        wait();
        NetSlot& s = slots[slot];
        if (s.rec_alloc != nullptr) {
            HIP_CHECK(hipFree(s.rec_alloc));
            s.rec_alloc = nullptr;
        }
        DevNet& n = s.dn;
        n.rec_start = start;
        n.rec_every = every;
        n.rec_n = na;
        n.rec_cap = cap;
        n.rec_buf = nullptr;
        if (na > 0 && cap > 0) {
            HIP_CHECK(hipMalloc(&s.rec_alloc, static_cast<size_t>(cap) * 2 * na * sizeof(double)));
            n.rec_buf = static_cast<double*>(s.rec_alloc);
        } else {
            n.rec_n = 0;
        }
        s.rec_read = 0;
        // Only the recording fields change; write them alone so the device's
        // step counters are not overwritten with the host's copy.
        HIP_CHECK(hipMemcpyAsync(&d_nets[slot].rec_start, &n.rec_start,
                                 offsetof(DevNet, rec_buf) + sizeof(double*) -
                                     offsetof(DevNet, rec_start),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Records available = recording steps in [rec_start, step).
    int64_t read_trace(int slot, double* out, int64_t* steps, int64_t max_rec, bool& overflow) {
        wait();
        NetSlot& s = slots[slot];
        const DevNet& n = s.dn;
        overflow = false;
        if (n.rec_n == 0 || s.step <= n.rec_start) return 0;
        const int64_t avail = (s.step - 1 - n.rec_start) / n.rec_every + 1;
        if (avail - s.rec_read > n.rec_cap) {
            overflow = true;
            return 0;
        }
        const int64_t m = std::min(avail - s.rec_read, max_rec);
        const size_t rec_doubles = 2 * static_cast<size_t>(n.rec_n);
        for (int64_t q = 0; q < m; ++q) {
            const int64_t r = s.rec_read + q;
            HIP_CHECK(hipMemcpyAsync(out + q * rec_doubles,
                                     n.rec_buf + static_cast<size_t>(r % n.rec_cap) * rec_doubles,
                                     rec_doubles * sizeof(double), hipMemcpyDeviceToHost, stream));
            if (steps != nullptr) steps[q] = n.rec_start + r * n.rec_every;
        }
        HIP_CHECK(hipStreamSynchronize(stream));
        s.rec_read += m;
        return m;
    }

    void read_state(int slot, int field, int64_t off, int64_t cnt, double* out) {
        wait();
        const DevNet& n = slots[slot].dn;
        if (cnt <= 0) return;
        const double* src = nullptr;
        const float* fsrc = nullptr;
        const bool f32 = n.precision == ASG_FP32;
        switch (field) {
            case ASG_FIELD_V: src = n.V; break;
            case ASG_FIELD_W: src = n.w; break;
            case ASG_FIELD_G_EX: src = n.gex; fsrc = n.fgex; break;
            case ASG_FIELD_DG_EX: src = n.dgex; fsrc = n.fdgex; break;
            case ASG_FIELD_G_IN: src = n.gin; fsrc = n.fgin; break;
            case ASG_FIELD_DG_IN: src = n.dgin; fsrc = n.fdgin; break;
            case ASG_FIELD_I_SIC: src = n.Isic; fsrc = n.fIsic; break;
            case ASG_FIELD_CA: src = n.Ca; break;
            case ASG_FIELD_IP3: src = n.IP3; break;
            case ASG_FIELD_H: src = n.hh; break;
            case ASG_FIELD_STP_X: src = n.stp_x; fsrc = nullptr; break;
            case ASG_FIELD_STP_U: src = n.stp_u; fsrc = nullptr; break;
            case ASG_FIELD_STP_TLAST: src = n.stp_tl; fsrc = nullptr; break;
            default: break;
        }
        if (f32 && fsrc != nullptr) {
            std::vector<float> tmp(static_cast<size_t>(cnt));
            HIP_CHECK(hipMemcpyAsync(tmp.data(), fsrc + off, cnt * sizeof(float),
                                     hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            for (int64_t q = 0; q < cnt; ++q) out[q] = tmp[static_cast<size_t>(q)];
            return;
        }
        HIP_CHECK(hipMemcpyAsync(out, src + off, cnt * sizeof(double), hipMemcpyDeviceToHost,
                                 stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
};

}  // namespace asg

// ------------------------------------------------------------------ C ABI --

using asg::Engine;
using asg::fail;
using asg::guarded;

struct asg_net {
    std::unique_ptr<Engine> eng;
};

struct asg_batch {
    std::unique_ptr<Engine> eng;
};

struct asg_prepared {
    asg::Prepared p;
};

namespace {

const char* overflow_message(int bits) {
    if (bits & 2) {
        return "a window had more spikes than the delivery capacity of this network "
               "(its deliveries were incomplete; results from that window on are invalid)";
    }
    return "raster buffer overflowed; drain spikes more often";
}

int64_t field_size(const asg::NetSlot& s, int32_t field) {
    if (field >= ASG_FIELD_CA && field <= ASG_FIELD_H) return s.dn.n_astro;
    return s.dn.n_neur;
}

asg_status read_state_common(Engine& e, int32_t slot, int32_t field, int64_t offset,
                             int64_t count, double* out) {
    if (slot < 0 || slot >= e.max_nets || !e.slots[slot].used) {
        return fail(ASG_ERR_INVALID, "no network in this slot");
    }
    if (field < 0 || field >= ASG_FIELD_COUNT) return fail(ASG_ERR_INVALID, "unknown state field");
    const int64_t size = field_size(e.slots[slot], field);
    if (offset < 0 || count < 0 || offset + count > size || (count > 0 && out == nullptr)) {
        return fail(ASG_ERR_INVALID, "state range out of bounds");
    }
    e.read_state(slot, field, offset, count, out);
    return ASG_OK;
}

asg_status counters_common(Engine& e, int32_t slot, asg_counters* out) {
    if (slot < 0 || slot >= e.max_nets || !e.slots[slot].used || out == nullptr) {
        return fail(ASG_ERR_INVALID, "no network in this slot");
    }
    e.wait();
    e.download(slot);
    const asg::DevNet& n = e.slots[slot].dn;
    out->steps = n.c_steps;
    out->spikes = n.c_spikes;
    out->synapse_events = n.c_syn;
    // SIC gathers: one per (astrocyte, target) pair on every arrival step.
    int64_t pend = 0;
    if (n.c_steps > n.d_a2n) {
        pend = (n.c_steps - 1 - n.d_a2n) / n.sic_interval + 1;
    }
    out->sic_entries = pend * n.sic_pairs;
    return ASG_OK;
}

}  // namespace

extern "C" {

const char* asg_last_error(void) { return asg::g_last_error.c_str(); }

asg_status asg_host_fail(asg_status code, const char* what) { return fail(code, what); }

const char* asg_version(void) { return "astrosimgpu-amd 0.1 (gfx950)"; }

int64_t asg_query_footprint(const asg_net_desc* desc) {
    if (desc == nullptr) return -1;
    return static_cast<int64_t>(asg::BlockPool::rounded(asg::plan_of(*desc, nullptr).bytes));
}

asg_status asg_net_create(const asg_net_desc* desc, asg_net** out) {
    if (desc == nullptr || out == nullptr) return fail(ASG_ERR_INVALID, "NULL argument");
    std::string err = asg::validate(*desc);
    if (err.empty()) err = asg::check_fp32_limits(*desc);
    if (!err.empty()) return fail(ASG_ERR_INVALID, err);
    return guarded([&]() -> asg_status {
        auto net = std::make_unique<asg_net>();
        net->eng = std::make_unique<Engine>(1, 0);
        if (net->eng->add(*desc, INT64_MAX) != 0) {
            return fail(ASG_ERR_CAPACITY, "could not place the network");
        }
        *out = net.release();
        return ASG_OK;
    });
}

asg_status asg_net_destroy(asg_net* net) {
    delete net;
    return ASG_OK;
}

asg_status asg_net_step(asg_net* net, int64_t n_steps) {
    if (net == nullptr || n_steps < 0) return fail(ASG_ERR_INVALID, "bad argument");
    return guarded([&]() -> asg_status {
        net->eng->step(n_steps, false);
        if (const int bits = net->eng->overflow_bits(0)) {
            return fail(ASG_ERR_OVERFLOW, overflow_message(bits));
        }
        return ASG_OK;
    });
}

asg_status asg_net_step_async(asg_net* net, int64_t n_steps) {
    if (net == nullptr || n_steps < 0) return fail(ASG_ERR_INVALID, "bad argument");
    return guarded([&]() -> asg_status {
        net->eng->step(n_steps, true);
        return ASG_OK;
    });
}

asg_status asg_net_sync(asg_net* net) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return guarded([&]() -> asg_status {
        net->eng->wait();
        if (const int bits = net->eng->overflow_bits(0)) {
            return fail(ASG_ERR_OVERFLOW, overflow_message(bits));
        }
        return ASG_OK;
    });
}

int64_t asg_net_current_step(const asg_net* net) {
    return net == nullptr ? -1 : net->eng->slots[0].step;
}

asg_status asg_net_spikes(asg_net* net, int64_t* step, uint32_t* source, int64_t capacity,
                          int64_t* n_out) {
    if (net == nullptr || n_out == nullptr || capacity < 0 ||
        (capacity > 0 && (step == nullptr || source == nullptr))) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        *n_out = net->eng->spikes_out(0, step, source, capacity);
        return ASG_OK;
    });
}

asg_status asg_net_read_state(asg_net* net, int32_t field, int64_t offset, int64_t count,
                              double* out) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return guarded([&] { return read_state_common(*net->eng, 0, field, offset, count, out); });
}

asg_status asg_net_counters(asg_net* net, asg_counters* out) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return guarded([&] { return counters_common(*net->eng, 0, out); });
}

namespace {
asg_status set_recording_common(Engine& e, int32_t slot, int64_t start, int32_t every,
                                uint32_t na, int64_t cap) {
    if (slot < 0 || slot >= e.max_nets || !e.slots[slot].used) {
        return fail(ASG_ERR_INVALID, "no network in this slot");
    }
    if (start < 0 || every < 1 || cap < 0 || na > e.slots[slot].dn.n_astro) {
        return fail(ASG_ERR_INVALID, "bad recording setup");
    }
    return guarded([&]() -> asg_status {
        e.set_recording(slot, start, every, na, cap);
        return ASG_OK;
    });
}
asg_status read_trace_common(Engine& e, int32_t slot, double* out, int64_t* steps,
                             int64_t max_rec, int64_t* n_out) {
    if (slot < 0 || slot >= e.max_nets || !e.slots[slot].used || n_out == nullptr ||
        max_rec < 0 || (max_rec > 0 && out == nullptr)) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        bool overflow = false;
        *n_out = e.read_trace(slot, out, steps, max_rec, overflow);
        if (overflow) return fail(ASG_ERR_OVERFLOW, "recording ring wrapped before it was read");
        return ASG_OK;
    });
}
}  // namespace

asg_status asg_net_set_recording(asg_net* net, int64_t start_step, int32_t every,
                                 uint32_t n_astro, int64_t capacity) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return set_recording_common(*net->eng, 0, start_step, every, n_astro, capacity);
}

asg_status asg_net_read_trace(asg_net* net, double* out, int64_t* steps, int64_t max_records,
                              int64_t* n_out) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return read_trace_common(*net->eng, 0, out, steps, max_records, n_out);
}

asg_status asg_batch_set_recording(asg_batch* batch, int32_t slot, int64_t start_step,
                                   int32_t every, uint32_t n_astro, int64_t capacity) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return set_recording_common(*batch->eng, slot, start_step, every, n_astro, capacity);
}

asg_status asg_batch_read_trace(asg_batch* batch, int32_t slot, double* out, int64_t* steps,
                                int64_t max_records, int64_t* n_out) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return read_trace_common(*batch->eng, slot, out, steps, max_records, n_out);
}

asg_status asg_batch_create(int32_t max_nets, int64_t budget_bytes, asg_batch** out) {
    if (max_nets < 1 || out == nullptr || budget_bytes < 0) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        auto b = std::make_unique<asg_batch>();
        b->eng = std::make_unique<Engine>(max_nets, budget_bytes);
        *out = b.release();
        return ASG_OK;
    });
}

asg_status asg_batch_destroy(asg_batch* batch) {
    delete batch;
    return ASG_OK;
}

asg_status asg_batch_add(asg_batch* batch, const asg_net_desc* desc, int64_t total_steps,
                         int32_t* slot) {
    if (batch == nullptr || desc == nullptr || slot == nullptr || total_steps < 0) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    std::string err = asg::validate(*desc);
    if (err.empty()) err = asg::check_fp32_limits(*desc);
    if (!err.empty()) return fail(ASG_ERR_INVALID, err);
    return guarded([&]() -> asg_status {
        const int s = batch->eng->add(*desc, total_steps);
        if (s == -1) return fail(ASG_ERR_CAPACITY, "no free slot");
        if (s == -2) return fail(ASG_ERR_CAPACITY, "device memory budget exceeded");
        *slot = s;
        return ASG_OK;
    });
}

asg_status asg_net_prepare(const asg_net_desc* desc, asg_prepared** out) {
    if (desc == nullptr || out == nullptr) return fail(ASG_ERR_INVALID, "NULL argument");
    std::string err = asg::validate(*desc);
    if (err.empty()) err = asg::check_fp32_limits(*desc);
    if (!err.empty()) return fail(ASG_ERR_INVALID, err);
    try {
        auto prep = std::make_unique<asg_prepared>();
        prep->p = asg::prepare(*desc);
        *out = prep.release();
        return ASG_OK;
    } catch (const std::bad_alloc&) {
        return fail(ASG_ERR_NOMEM, "host allocation failed");
    }
}

void asg_prepared_free(asg_prepared* prep) { delete prep; }

asg_status asg_batch_add_prepared(asg_batch* batch, const asg_net_desc* desc,
                                  const asg_prepared* prep, int64_t total_steps, int32_t* slot) {
    if (batch == nullptr || desc == nullptr || prep == nullptr || slot == nullptr ||
        total_steps < 0) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        const int s = batch->eng->add(*desc, total_steps, &prep->p);
        if (s == -1) return fail(ASG_ERR_CAPACITY, "no free slot");
        if (s == -2) return fail(ASG_ERR_CAPACITY, "device memory budget exceeded");
        *slot = s;
        return ASG_OK;
    });
}

asg_status asg_batch_remove(asg_batch* batch, int32_t slot) {
    if (batch == nullptr || slot < 0 || slot >= batch->eng->max_nets ||
        !batch->eng->slots[slot].used) {
        return fail(ASG_ERR_INVALID, "no network in this slot");
    }
    return guarded([&]() -> asg_status {
        batch->eng->remove(slot);
        return ASG_OK;
    });
}

asg_status asg_batch_step(asg_batch* batch, int64_t n_steps) {
    if (batch == nullptr || n_steps < 0) return fail(ASG_ERR_INVALID, "bad argument");
    return guarded([&]() -> asg_status {
        batch->eng->step(n_steps, false);
        return ASG_OK;
    });
}

asg_status asg_batch_step_async(asg_batch* batch, int64_t n_steps) {
    if (batch == nullptr || n_steps < 0) return fail(ASG_ERR_INVALID, "bad argument");
    return guarded([&]() -> asg_status {
        batch->eng->step(n_steps, true);
        return ASG_OK;
    });
}

int64_t asg_batch_net_step(asg_batch* batch, int32_t slot) {
    if (batch == nullptr || slot < 0 || slot >= batch->eng->max_nets ||
        !batch->eng->slots[slot].used) {
        return -1;
    }
    return batch->eng->slots[slot].step;
}

asg_status asg_batch_spikes(asg_batch* batch, int32_t slot, int64_t* step, uint32_t* source,
                            int64_t capacity, int64_t* n_out) {
    if (batch == nullptr || slot < 0 || slot >= batch->eng->max_nets ||
        !batch->eng->slots[slot].used || n_out == nullptr || capacity < 0 ||
        (capacity > 0 && (step == nullptr || source == nullptr))) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        if (const int bits = batch->eng->overflow_bits(slot)) {
            return fail(ASG_ERR_OVERFLOW, overflow_message(bits));
        }
        *n_out = batch->eng->spikes_out(slot, step, source, capacity);
        return ASG_OK;
    });
}

asg_status asg_batch_spikes_all(asg_batch* batch, int64_t* counts, int64_t* step,
                                uint32_t* source, int64_t capacity, int64_t* n_out) {
    if (batch == nullptr || counts == nullptr || n_out == nullptr || capacity < 0 ||
        (capacity > 0 && (step == nullptr || source == nullptr))) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        if (const int bits = batch->eng->overflow_bits_any()) {
            return fail(ASG_ERR_OVERFLOW, overflow_message(bits));
        }
        *n_out = batch->eng->spikes_out_all(counts, step, source, capacity);
        return ASG_OK;
    });
}

asg_status asg_batch_read_state(asg_batch* batch, int32_t slot, int32_t field, int64_t offset,
                                int64_t count, double* out) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return guarded([&] { return read_state_common(*batch->eng, slot, field, offset, count, out); });
}

asg_status asg_batch_counters(asg_batch* batch, int32_t slot, asg_counters* out) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return guarded([&] { return counters_common(*batch->eng, slot, out); });
}

int64_t asg_batch_bytes_in_use(const asg_batch* batch) {
    return batch == nullptr ? -1 : batch->eng->bytes_in_use();
}

namespace {
asg_status set_graphs(Engine& e, int32_t on) {
    return guarded([&]() -> asg_status {
        e.wait();
        e.use_graphs = on != 0;
        if (!e.use_graphs) e.clear_graphs();
        return ASG_OK;
    });
}
}  // namespace

asg_status asg_net_set_graphs(asg_net* net, int32_t on) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return set_graphs(*net->eng, on);
}

asg_status asg_batch_set_graphs(asg_batch* batch, int32_t on) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return set_graphs(*batch->eng, on);
}

asg_status asg_batch_set_schedule(asg_batch* batch, const asg_schedule* sched) {
    if (batch == nullptr || sched == nullptr) return fail(ASG_ERR_INVALID, "bad argument");
    const int w = sched->cell_waves, g = sched->sic_group, sg = sched->step_group;
    const int dw = sched->deliver_waves, sw = sched->sic_workers, ln = sched->lanes;
    const int tm = sched->teams;
    if ((w != 1 && w != 2 && w != 4) || (g != 0 && g != 1 && g != 2 && g != 4 && g != 8 && g != 16) ||
        sg < 0 || sg > asg::kWindowCap || (dw != 0 && (dw < 64 || dw > 4096)) ||
        (sw != 0 && (sw < 64 || sw > 8192)) || ln < 0 || ln > asg::kMaxLanes || tm < 0 ||
        tm > asg::kMaxTeams) {
        return fail(ASG_ERR_INVALID,
                    "cell_waves must be 1, 2 or 4; sic_group 0, 1, 2, 4, 8 or 16; step_group 0-16; "
                    "deliver_waves 0 or 64-4096; sic_workers 0 or 64-8192; lanes 0-3; "
                    "teams 0-1024");
    }
    return guarded([&]() -> asg_status {
        batch->eng->wait();
        batch->eng->sched_cell_waves = w;
        batch->eng->sched_sic_group = g;
        batch->eng->sched_step_group = sg;
        batch->eng->sched_deliver_waves = dw == 0 ? asg::kDeliverWaves : dw;
        batch->eng->sched_sic_workers = sw;
        batch->eng->set_lanes(ln == 0 ? asg::kLanes : ln);
        batch->eng->set_teams(tm);
        return ASG_OK;
    });
}

asg_status asg_batch_get_schedule(asg_batch* batch, asg_schedule* out) {
    if (batch == nullptr || out == nullptr) return fail(ASG_ERR_INVALID, "bad argument");
    out->cell_waves = batch->eng->sched_cell_waves;
    out->sic_group = batch->eng->sched_sic_group;
    out->step_group = batch->eng->sched_step_group;
    out->deliver_waves = batch->eng->sched_deliver_waves;
    out->sic_workers = batch->eng->sched_sic_workers;
    out->lanes = batch->eng->sched_lanes;
    out->teams = batch->eng->sched_teams;
    out->ms_per_window = batch->eng->tuned_ms_per_window;
    return ASG_OK;
}

asg_status asg_batch_autotune(asg_batch* batch, int32_t windows_per_try, int32_t rounds,
                              asg_schedule* chosen) {
    if (batch == nullptr || windows_per_try <= 0 || rounds <= 0) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    return guarded([&]() -> asg_status {
        batch->eng->autotune(windows_per_try, rounds);
        if (const int bits = batch->eng->overflow_bits_any()) {
            return fail(ASG_ERR_OVERFLOW, overflow_message(bits));
        }
        if (chosen != nullptr) {
            return asg_batch_get_schedule(batch, chosen);
        }
        return ASG_OK;
    });
}

namespace {
asg_status kernel_times(Engine& e, asg_kernel_times* out) {
    if (out == nullptr) return fail(ASG_ERR_INVALID, "NULL output");
    return guarded([&]() -> asg_status {
        e.wait();
        for (int k = 0; k < 5; ++k) out->ms[k] = e.kernel_ms[k];
        out->windows = e.timed_windows;
        return ASG_OK;
    });
}
asg_status set_profiling(Engine& e, int32_t every) {
    if (every < 0) return fail(ASG_ERR_INVALID, "profiling period must be >= 0");
    return guarded([&]() -> asg_status {
        e.wait();
        if (e.profiling != (every != 0)) {
            e.dirty = true;  // profiling runs the FP32 networks in one lane
        }
        e.profiling = every != 0;
        e.prof_every = every > 0 ? every : 1;
        e.prof_seq = 0;
        for (double& v : e.kernel_ms) v = 0.0;
        e.timed_windows = 0;
        return ASG_OK;
    });
}
}  // namespace

asg_status asg_net_set_profiling(asg_net* net, int32_t on) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return set_profiling(*net->eng, on);
}
asg_status asg_net_kernel_times(asg_net* net, asg_kernel_times* out) {
    if (net == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return kernel_times(*net->eng, out);
}
asg_status asg_batch_set_profiling(asg_batch* batch, int32_t on) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return set_profiling(*batch->eng, on);
}
asg_status asg_batch_kernel_times(asg_batch* batch, asg_kernel_times* out) {
    if (batch == nullptr) return fail(ASG_ERR_INVALID, "NULL handle");
    return kernel_times(*batch->eng, out);
}

}  // extern "C"

// ------------------------------------------------- no-allocation variant --

namespace {

// Staging of asg_device_math_noalloc: the inputs (first half) and outputs
// (second half) of one call, in a file-scope device array every GPU the
// library runs on holds until the process ends (16 MiB).
constexpr int64_t kNoAllocMaxValues = int64_t(1) << 20;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxValues];

// The address of `scratch` on each device, looked up once per device with
// that device current (host code must not use the array's name as an
// address: in host code it is only a stand-in for the symbol).
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

asg_status scratch_of_current_device(double** out) {
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        return fail(ASG_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    }
    if (dev < 0 || dev >= kMaxDevices) {
        return fail(ASG_ERR_DEVICE, "device index beyond the scratch table");
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t g = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g != hipSuccess || p == nullptr) {
            return fail(ASG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(g));
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ASG_OK;
}

}  // namespace

extern "C" {

asg_status asg_device_math_noalloc(int32_t fn, const double* x, double* y, int64_t n) {
    if (fn < ASG_MATH_EXP || fn > ASG_MATH_DIV || n < 0 || (n > 0 && (x == nullptr || y == nullptr)) ||
        (fn == ASG_MATH_DIV && n % 2 != 0)) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    if (n > kNoAllocMaxValues) {
        return fail(ASG_ERR_CAPACITY, "asg_device_math_noalloc: n = " + std::to_string(n) +
                                          " exceeds its scratch of " +
                                          std::to_string(kNoAllocMaxValues) + " values");
    }
    if (n == 0) return ASG_OK;
    double* d = nullptr;
    const asg_status st = scratch_of_current_device(&d);
    if (st != ASG_OK) return st;
    return guarded([&]() -> asg_status {
        HIP_CHECK(hipMemcpy(d, x, static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice));
        const int block = 256;
        const int64_t grid = std::min<int64_t>((n + block - 1) / block, 65536);
        hipLaunchKernelGGL(asg::k_math, dim3(static_cast<unsigned>(grid)), dim3(block), 0, nullptr,
                           fn, d, d + kNoAllocMaxValues, n);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(y, d + kNoAllocMaxValues, static_cast<size_t>(n) * sizeof(double),
                            hipMemcpyDeviceToHost));
        return ASG_OK;
    });
}

asg_status asg_device_math(int32_t fn, const double* x, double* y, int64_t n) {
    if (fn < ASG_MATH_EXP || fn > ASG_MATH_DIV || n < 0 || (n > 0 && (x == nullptr || y == nullptr)) ||
        (fn == ASG_MATH_DIV && n % 2 != 0)) {
        return fail(ASG_ERR_INVALID, "bad argument");
    }
    if (n == 0) return ASG_OK;
    return guarded([&]() -> asg_status {
        double* d = nullptr;
        HIP_CHECK(hipMalloc(&d, 2 * static_cast<size_t>(n) * sizeof(double)));
        struct Free {
            double* p;
            ~Free() { (void)hipFree(p); }
        } guard{d};
        HIP_CHECK(hipMemcpy(d, x, static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice));
        const int block = 256;
        const int64_t grid = std::min<int64_t>((n + block - 1) / block, 65536);
        hipLaunchKernelGGL(asg::k_math, dim3(static_cast<unsigned>(grid)), dim3(block), 0, nullptr,
                           fn, d, d + n, n);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(y, d + n, static_cast<size_t>(n) * sizeof(double), hipMemcpyDeviceToHost));
        return ASG_OK;
    });
}

}  // extern "C"

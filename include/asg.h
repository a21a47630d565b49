/*
 * asg.h -- C ABI of the MI355X simulation step for the tripartite
 * neuron-astrocyte network of astrosimgpu.
 *
 * Plain C types only (no HIP, no torch), so a C++17 host, a ctypes binding or
 * any other FFI can call it. Three layers:
 *
 *  1. asg_astro_*  one-for-one replacement of the reference's native CUDA
 *                  astrocyte backend (include/astrosimgpu/astrocyte_cuda.hpp:
 *                  13-27): same arguments, same ownership, same abort-on-error
 *                  behaviour. Only the astrocyte update runs on the device.
 *  2. asg_net_*    the whole time step of one network on the device
 *                  (network.cpp:385-468: arrivals, astrocyte drive and update,
 *                  neuron update, spike delivery, SIC delivery), stepped in
 *                  chunks with the spike raster and state read back on demand.
 *  3. asg_batch_*  many independent networks of uneven size stepped together
 *                  in one set of launches; networks can be added and retired
 *                  between calls (the serving loop's continuous batching).
 *
 * Errors: layers 2 and 3 return an asg_status and keep a per-thread message
 * for asg_last_error(). Setting the environment variable ASG_ABORT_ON_ERROR=1
 * makes every failure print and abort instead, as the reference backend does
 * (astrocyte_cuda.cu:14-24).
 *
 * Threading: one handle belongs to one host thread at a time; each handle has
 * its own HIP stream on the current device.
 */
#ifndef ASG_H
#define ASG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum asg_status {
    ASG_OK = 0,
    ASG_ERR_INVALID = 1,   /* bad argument or description */
    ASG_ERR_NOMEM = 2,     /* device or host allocation failed */
    ASG_ERR_DEVICE = 3,    /* a HIP call failed */
    ASG_ERR_OVERFLOW = 4,  /* raster buffer full: drain it more often */
    ASG_ERR_CONFIG = 5,    /* configuration file rejected */
    ASG_ERR_CAPACITY = 6   /* batch has no room (memory budget or slots) */
} asg_status;

/* Message of the last failing call on this thread ("" if none). */
const char* asg_last_error(void);

/* Library version string. */
const char* asg_version(void);

/* ------------------------------------------------------------------------ */
/* Parameter blocks. Field meanings and units are those of the reference     */
/* structs named in each comment.                                            */
/* ------------------------------------------------------------------------ */

/* NeuronParams, parameters.hpp:40-58 (one population). */
typedef struct asg_neuron_params {
    double C_m, g_L, E_L, V_th, Delta_T, a, b, tau_w;
    double V_reset, V_peak, t_ref, E_ex, E_in, tau_syn_ex, tau_syn_in, I_e;
    double V_m_init;
} asg_neuron_params;

/* InputParams, parameters.hpp:132-138. noise_dt must already be resolved
 * (the reference turns <= 0 into 10 * dt at load, parameters.cpp:169-173). */
typedef struct asg_input_params {
    double poiss_rate, poiss_weight, gauss_noise_std, noise_dt;
    int32_t independent_noise;
} asg_input_params;

/* AstroConstants, astrocyte_kernel.hpp:29-40. */
typedef struct asg_astro_constants {
    double Kd_IP3_1, Kd_IP3_2, Kd_act, Kd_inh, Km_SERCA;
    double k_IP3R, rate_IP3R, rate_L, rate_SERCA, ratio_ER_cyt;
} asg_astro_constants;

/* The rest of AstrocyteParams (parameters.hpp:14-38) the step needs. */
typedef struct asg_astro_params {
    asg_astro_constants c;
    double SIC_th, SIC_scale;
    double Ca_init, IP3_init, h_init;
} asg_astro_params;

/* ConnectionSet, network.hpp:14-27: rows grouped by source. The reference
 * stores one weight and one delay per synapse but always fills each set with
 * a single value (network.cpp:206-223), so they are scalars here. */
typedef struct asg_csr {
    uint32_t n_rows;            /* number of sources */
    const uint32_t* row_start;  /* n_rows + 1 entries */
    const uint32_t* target;     /* row_start[n_rows] entries */
    double weight;
    int32_t delay_steps;        /* >= 1 */
} asg_csr;

/* StpParams, parameters.hpp:61-66. Applies to exc, inh and n2a sets. */
typedef struct asg_stp {
    int32_t enabled;
    double U, tau_rec, tau_fac;
} asg_stp;

enum { ASG_FP64 = 0, ASG_FP32 = 1 };

/* Everything one network's step needs; the host keeps ownership of every
 * array (they are copied at create). Per-neuron arrays have n_exc + n_inh
 * entries, excitatory cells first; per-astrocyte arrays n_astro entries. */
typedef struct asg_net_desc {
    uint32_t n_exc, n_inh, n_astro;
    double dt;
    int32_t substeps;
    uint64_t seed;
    asg_neuron_params neuron[2];      /* [0] excitatory, [1] inhibitory */
    const double* V_reset;            /* per neuron, randomised at build */
    const double* b;                  /* per neuron, randomised at build */
    asg_input_params input_exc, input_inh, input_astro;
    asg_astro_params astro;
    const double* Ca_tot;             /* per astrocyte, randomised at build */
    const double* IP3_0;
    const double* tau_IP3;
    const double* delta_IP3;
    asg_csr exc;                      /* excitatory neuron -> neuron */
    asg_csr inh;                      /* inhibitory neuron -> neuron (weight < 0) */
    asg_csr n2a;                      /* excitatory neuron -> astrocyte */
    asg_csr a2n;                      /* astrocyte -> neuron (SIC) */
    asg_stp stp;
    int32_t sic_interval;             /* >= 1 */
    int32_t precision;                /* ASG_FP64 or ASG_FP32 */
} asg_net_desc;

/* Per-network work counters, cumulative since create (used for the
 * algorithmic-byte count of the roofline). */
typedef struct asg_counters {
    int64_t steps;
    int64_t spikes;
    int64_t synapse_events;     /* delivered exc + inh + n2a synapse entries */
    int64_t sic_entries;        /* (astrocyte, target) gathers with factor != 0 */
} asg_counters;

/* State fields for asg_net_read_state / asg_batch_read_state. */
enum {
    ASG_FIELD_V = 0, ASG_FIELD_W, ASG_FIELD_G_EX, ASG_FIELD_DG_EX, ASG_FIELD_G_IN,
    ASG_FIELD_DG_IN, ASG_FIELD_I_SIC, ASG_FIELD_CA, ASG_FIELD_IP3, ASG_FIELD_H,
    ASG_FIELD_STP_X, ASG_FIELD_STP_U, ASG_FIELD_STP_TLAST, ASG_FIELD_COUNT
};

/* ------------------------------------------------------------------------ */
/* Host side: configuration and network construction (no device needed).   */
/* load_config (parameters.cpp:85-204) followed by Network::build           */
/* (network.cpp:56-228): same JSON keys, same strictness, same network for  */
/* the same seed.                                                            */
/* ------------------------------------------------------------------------ */

typedef struct asg_model asg_model;

/* Command-line overrides applied after the file, as main.cpp:214-233 does
 * (negative = keep the file's value), plus optional "section.key=value"
 * assignments separated by ';' applied to the JSON before it is read. */
typedef struct asg_model_overrides {
    int64_t seed;
    double sim_time;
    double pre_time;
    int64_t n_astro;
    const char* assignments;  /* may be NULL */
} asg_model_overrides;

typedef struct asg_model_info {
    /* NetworkStats, network.hpp:47-55 */
    uint32_t n_astro, n_exc, n_inh;
    uint32_t n_primary_exc, n_primary_inh, n_neuron_to_astro, n_astro_to_neuron;
    /* time grid and run length */
    double dt, pre_sim_time, sim_time;
    int32_t substeps;
    int64_t pre_steps, sim_steps;
    uint64_t seed;
    /* recording (ModelConfig, parameters.hpp:185-191) */
    int32_t record_spikes, record_astro, record_neuron, record_every;
    uint32_t record_astro_max, record_neuron_max;
    /* analysis (AnalysisParams, parameters.hpp:160-166); ca_threshold < 0
     * means the astrocyte's SIC_th */
    int32_t analysis_enabled;
    double ca_threshold, merge_gap, hist_window, hist_shift, SIC_th;
    int32_t ring_slots;       /* max delay + 1, as network.cpp:73-79 */
    char output_dir[256];
} asg_model_info;

/* Build from a file / from JSON text. ASG_ERR_CONFIG with the reference's
 * message on a malformed file, an unknown key or an invalid value. */
asg_status asg_model_load(const char* path, const asg_model_overrides* ov, asg_model** out);
asg_status asg_model_load_text(const char* json_text, const asg_model_overrides* ov,
                               asg_model** out);
void asg_model_free(asg_model* model);

/* Description of the built network (valid while the model lives). */
const asg_net_desc* asg_model_desc(const asg_model* model);
asg_status asg_model_get_info(const asg_model* model, asg_model_info* out);

/* Post-run calcium analysis (analysis.cpp; its use in main.cpp:266-317). */
typedef struct asg_calcium_summary {
    int32_t cells, active;        /* astrocytes analysed / with >= 1 transient */
    int64_t transients, pairs;    /* events after merging / correlated pairs */
    double mean_duration, sd_duration, mean_correlation;
    double threshold, width, shift;  /* the threshold and window actually used */
} asg_calcium_summary;

/* ca is [n_samples][n_cells] sampled at times[n_samples] (ms) over the
 * recorded window [pre_sim_time, pre_sim_time + sim_time]. round6 != 0
 * rounds times and values to 6 significant digits first, as the
 * reference's CSV round trip does. */
asg_status asg_analyze_calcium(const double* times, const double* ca, int32_t n_cells,
                               int64_t n_samples, double threshold, double merge_gap,
                               double pre_sim_time, double sim_time, double hist_window,
                               double hist_shift, int32_t round6, asg_calcium_summary* out);

/* The same over per-cell traces of their own lengths, as main.cpp:46-74
 * reads them from astrocytes.csv: cell c's samples are
 * times/ca[offsets[c] .. offsets[c + 1]), offsets has n_cells + 1 entries. */
asg_status asg_analyze_calcium_cells(const double* times, const double* ca,
                                     const int64_t* offsets, int32_t n_cells, double threshold,
                                     double merge_gap, double pre_sim_time, double sim_time,
                                     double hist_window, double hist_shift, int32_t round6,
                                     asg_calcium_summary* out);

/* detect_events, analysis.cpp:10-64: onsets/offsets of threshold crossings,
 * merged when closer than max_gap (if merge != 0). */
asg_status asg_detect_events(const double* times, const double* values, int64_t n,
                             double threshold, int32_t merge, double max_gap, double* onset,
                             double* offset, int64_t capacity, int64_t* n_events);

/* pearson, analysis.cpp:98-123 (0 when either series has no variance). */
double asg_pearson(const double* a, const double* b, int64_t n);

/* sliding_window_rate, analysis.cpp:66-96: events per ms per window. */
asg_status asg_sliding_window_rate(const double* events, int64_t n, double start, double end,
                                   double width, double shift, double* out, int64_t capacity,
                                   int64_t* n_out);

/* ------------------------------------------------------------------------ */
/* Layer 1: astrocyte backend (drop-in for astrocyte_cuda.hpp:13-27).        */
/* FP64, default stream semantics: every call completes before returning.    */
/* Failures print file:line and abort, as the reference does.                */
/* ------------------------------------------------------------------------ */

typedef struct asg_astro asg_astro;

/* cuda_astro_create, astrocyte_cuda.hpp:15-17. n == 0 gives a handle with no
 * buffers; every later call on it is a no-op. */
asg_astro* asg_astro_create(uint32_t n, const double* Ca, const double* IP3, const double* h,
                            const double* Ca_tot, const double* IP3_0, const double* tau_IP3,
                            const double* delta_IP3);

/* cuda_astro_destroy, astrocyte_cuda.hpp:19: copies Ca, IP3, h back, frees. */
void asg_astro_destroy(asg_astro* state, double* Ca, double* IP3, double* h);

/* cuda_astro_push_input, astrocyte_cuda.hpp:21: host -> device, n doubles. */
void asg_astro_push_input(asg_astro* state, const double* ip3_input);

/* cuda_astro_pull_calcium, astrocyte_cuda.hpp:23: device -> host, n doubles. */
void asg_astro_pull_calcium(asg_astro* state, double* Ca);

/* cuda_astro_update, astrocyte_cuda.hpp:25-27: one step of every cell
 * (optional independent Box-Muller noise, IP3 jump, RK4 substeps, clamps),
 * then zeroes the device ip3_input. */
void asg_astro_update(asg_astro* state, const asg_astro_constants* c, double h_step,
                      int32_t substeps, double noise_std, int32_t independent_noise,
                      double shared_noise, uint64_t noise_seed, uint64_t noise_index);

/* Isolated astrocyte update at scale (the astrocyte half of the step,
 * astro_update_kernel of astrocyte_cuda.cu:27-58): n cells in SoA arrays in
 * device memory, one update launch per step reading each cell's state and
 * presynaptic spike count (Poisson(lambda), int32, drawn into device memory
 * by a separate launch) and writing the state back. No counterpart in the
 * reference's API; it measures the update kernel's HBM throughput. Initial
 * state and per-cell parameters are drawn from `seed` (see
 * csrc/asg_astro_scale.hip); both precisions start from the same state. */
typedef struct asg_astro_scaling_params {
    double Ca_tot, IP3_0, tau_IP3, delta_IP3;   /* nominal per-cell parameters */
    double dt;                                  /* step (ms), one RK4 substep */
    double lambda;                              /* mean spike count per step, <= 30 */
    int32_t randomized;                         /* per-cell parameters drawn like */
    double lower, upper, var;                   /* draw_scaled, astrocyte.cpp:13-23 */
} asg_astro_scaling_params;

typedef struct asg_astro_scaling_result {
    double update_ms;        /* device time of the update launches (HIP events) */
    double total_ms;         /* first update start to last update end */
    double bytes_per_step;   /* algorithmic HBM bytes of one update launch */
    int64_t steps;
} asg_astro_scaling_result;

/* Runs `steps` steps in `precision` and copies the final Ca, IP3, h (as
 * doubles; any may be NULL) to the host. */
asg_status asg_astro_scaling(uint32_t n, int32_t steps, int32_t precision,
                             const asg_astro_constants* constants,
                             const asg_astro_scaling_params* params, uint64_t seed,
                             double* Ca_out, double* IP3_out, double* h_out,
                             asg_astro_scaling_result* out);

/* Algorithmic bytes of one update launch: per cell Ca, IP3, h read and
 * written, the int32 count read, and the four per-cell parameters read when
 * randomized. */
int64_t asg_astro_scaling_bytes(uint32_t n, int32_t precision, int32_t randomized);

/* ------------------------------------------------------------------------ */
/* Layer 2: one network, whole step on the device (replaces the body of      */
/* Network::run's step loop, network.cpp:385-468).                           */
/* ------------------------------------------------------------------------ */

typedef struct asg_net asg_net;

/* Copies the description to the device; the network starts at step 0 with
 * the reference's initial state (network.cpp:56-108, 43-52). */
asg_status asg_net_create(const asg_net_desc* desc, asg_net** out);
asg_status asg_net_destroy(asg_net* net);

/* Advance n_steps >= 0 steps. Returns after the device has finished. */
asg_status asg_net_step(asg_net* net, int64_t n_steps);

/* Same as asg_net_step but returns once the work is queued on the handle's
 * stream, so several networks can run on the device at once; asg_net_sync
 * (or any other call on the handle) waits for it. */
asg_status asg_net_step_async(asg_net* net, int64_t n_steps);
asg_status asg_net_sync(asg_net* net);

/* Number of steps taken so far. */
int64_t asg_net_current_step(const asg_net* net);

/* Move up to `capacity` recorded spikes (step, source) to the host, in the
 * reference's order (step ascending, then source ascending); *n_out is the
 * number written. Spikes not moved stay queued for the next call. */
asg_status asg_net_spikes(asg_net* net, int64_t* step, uint32_t* source, int64_t capacity,
                          int64_t* n_out);

/* Copy `count` entries starting at `offset` of a state field to the host. */
asg_status asg_net_read_state(asg_net* net, int32_t field, int64_t offset, int64_t count,
                              double* out);

asg_status asg_net_counters(asg_net* net, asg_counters* out);

/* Astrocyte recording (Network::run, network.cpp:449-456): from step
 * start_step on, every `every` steps, the Ca and IP3 of astrocytes
 * [0, n_astro) after that step's update are written on the device into a
 * ring of `capacity` records. A record is 2 * n_astro doubles: Ca then IP3.
 * Replaces any previous recording setup. */
asg_status asg_net_set_recording(asg_net* net, int64_t start_step, int32_t every,
                                 uint32_t n_astro, int64_t capacity);

/* Move up to `max_records` recorded samples (oldest first) to `out`
 * (max_records * 2 * n_astro doubles) and their step numbers to `steps`;
 * *n_out is the count. ASG_ERR_OVERFLOW if the ring wrapped before it was
 * drained. */
asg_status asg_net_read_trace(asg_net* net, double* out, int64_t* steps, int64_t max_records,
                              int64_t* n_out);

/* Bytes of device memory asg_net_create would take for `desc` (its planned
 * arrays, rounded up to the 2 MiB blocks they are carved from): what batch
 * admission counts against the budget. */
int64_t asg_query_footprint(const asg_net_desc* desc);

/* ------------------------------------------------------------------------ */
/* Layer 3: batches of independent networks (no counterpart in the           */
/* reference, which runs one network per process).                           */
/* ------------------------------------------------------------------------ */

typedef struct asg_batch asg_batch;

/* An empty batch that may hold up to max_nets networks using at most
 * budget_bytes of device memory (0 = no limit). */
asg_status asg_batch_create(int32_t max_nets, int64_t budget_bytes, asg_batch** out);
asg_status asg_batch_destroy(asg_batch* batch);

/* Admit a network that should run for total_steps steps; *slot receives its
 * id. ASG_ERR_CAPACITY if the footprint does not fit the budget now. */
asg_status asg_batch_add(asg_batch* batch, const asg_net_desc* desc, int64_t total_steps,
                         int32_t* slot);

/* Admission in two parts, for serving loops (no counterpart in the
 * reference). asg_net_prepare does the host-side work of asg_batch_add --
 * validation, the deduplicated neuron->astrocyte rows and the SIC gather
 * layout, the device plan -- with no device call, so it can run on any host
 * thread (a job's builder) while the device thread keeps stepping.
 * asg_batch_add_prepared then only allocates and uploads: it is
 * asg_batch_add for a desc that was prepared (the same desc, its host arrays
 * unchanged), with the same results. Free the handle with
 * asg_prepared_free once the network is added (or not needed). */
typedef struct asg_prepared asg_prepared;
asg_status asg_net_prepare(const asg_net_desc* desc, asg_prepared** out);
void asg_prepared_free(asg_prepared* prep);
asg_status asg_batch_add_prepared(asg_batch* batch, const asg_net_desc* desc,
                                  const asg_prepared* prep, int64_t total_steps, int32_t* slot);

/* Free a network's device memory (finished or not). */
asg_status asg_batch_remove(asg_batch* batch, int32_t slot);

/* Advance every unfinished network by up to n_steps steps in shared
 * launches. Networks stop at their total_steps. */
asg_status asg_batch_step(asg_batch* batch, int64_t n_steps);

/* Same as asg_batch_step but returns without waiting for the device; the
 * next call on the batch waits. */
asg_status asg_batch_step_async(asg_batch* batch, int64_t n_steps);

/* Steps taken by a network so far, or -1 for an empty slot. */
int64_t asg_batch_net_step(asg_batch* batch, int32_t slot);

asg_status asg_batch_spikes(asg_batch* batch, int32_t slot, int64_t* step, uint32_t* source,
                            int64_t capacity, int64_t* n_out);

/* The drain of asg_batch_spikes for every slot in one call (one device
 * round trip instead of one call per network). The pending (step, source)
 * pairs are written back to back in slot order, each slot's in its own
 * order; counts[s] (max_nets entries) receives how many belong to slot s
 * (0 for an empty slot) and *n_out the total. With fewer than the pending
 * pairs' room, the first `capacity` are written (the last slot written may
 * be partial) and the rest stay queued for the next call, so callers repeat
 * while *n_out == capacity. ASG_ERR_OVERFLOW, and nothing drained, if any
 * network's raster overflowed. */
asg_status asg_batch_spikes_all(asg_batch* batch, int64_t* counts, int64_t* step,
                                uint32_t* source, int64_t capacity, int64_t* n_out);
asg_status asg_batch_read_state(asg_batch* batch, int32_t slot, int32_t field, int64_t offset,
                                int64_t count, double* out);
asg_status asg_batch_counters(asg_batch* batch, int32_t slot, asg_counters* out);
asg_status asg_batch_set_recording(asg_batch* batch, int32_t slot, int64_t start_step,
                                   int32_t every, uint32_t n_astro, int64_t capacity);
asg_status asg_batch_read_trace(asg_batch* batch, int32_t slot, double* out, int64_t* steps,
                                int64_t max_records, int64_t* n_out);

/* Device bytes currently held by the batch's networks. */
int64_t asg_batch_bytes_in_use(const asg_batch* batch);

/* Launch schedule of the window pipeline: cell-kernel workgroup size in
 * wavefronts (1, 2 or 4), SIC group of the FP64 networks (window steps per
 * SIC wavefront: 1, 2, 4, 8 or 16; 0 = the largest that still fills the
 * device), step group of the FP64 networks (window steps per spike-marking
 * and per summing wavefront, 1-16; 0 = the whole window for large batches,
 * 1 otherwise), the wavefronts that share out the delivery list (64-4096,
 * 0 = 512), the FP32 post pass's SIC wavefronts (64-8192; 0 = one per
 * neuron block, the default), the lanes (1-3; 0 = 2): groups of the FP32
 * networks of about equal cell count whose windows run on streams of their
 * own, so that one group's cell pass overlaps another's post pass, and as
 * many groups of the FP64 networks, so that one group's cell pass overlaps
 * another's compaction, delivery and sums; and the FP32 teams (0 = off; 1-1024): with teams, each step call runs all
 * of its windows of the FP32 networks in one persistent cooperative launch
 * (grid sized by the occupancy API), the networks split into that many teams
 * (at most one per network), each with its own workgroups and its own
 * barrier between passes, so teams never wait for one another; lanes and
 * graph replay then do not apply to the FP32 networks. A barrier that waits
 * longer than 2 s stops the launch, and the step call (or the next call
 * that waits for it) returns ASG_ERR_DEVICE. Results do not depend on the
 * schedule. ms_per_window is the autotuner's measured time (output only). */
typedef struct asg_schedule {
    int32_t cell_waves;
    int32_t sic_group;
    int32_t step_group;
    int32_t deliver_waves;
    int32_t sic_workers;
    int32_t lanes;
    int32_t teams;
    float ms_per_window;
} asg_schedule;
asg_status asg_batch_set_schedule(asg_batch* batch, const asg_schedule* sched);
asg_status asg_batch_get_schedule(asg_batch* batch, asg_schedule* out);

/* Autotuner: tunes the schedule one knob at a time (how the FP32 networks'
 * windows are launched -- 1-3 lanes, or a persistent launch of 4, 16, 64 or
 * 256 teams -- when the batch has FP32 networks, cell waves, then the FP64 networks'
 * SIC and step groups when the batch has FP64 networks, then the delivery
 * wavefront count when it has FP32 networks), timing each candidate as step
 * calls of windows_per_try windows, `rounds` times round-robin, and keeping
 * the smallest median (written to *chosen if not NULL); a candidate that
 * changes the lanes first runs one untimed call of the same length. The
 * networks advance by the windows run, exactly as step calls of that length
 * would advance them. */
asg_status asg_batch_autotune(asg_batch* batch, int32_t windows_per_try, int32_t rounds,
                              asg_schedule* chosen);

/* Replay each step call's window sequence from a hipGraph captured on first
 * use (per window count, schedule and batch layout; recaptured after
 * networks are added or removed). Off by default; per-kernel profiling keeps
 * the launches eager. */
asg_status asg_batch_set_graphs(asg_batch* batch, int32_t on);
asg_status asg_net_set_graphs(asg_net* net, int32_t on);

/* ------------------------------------------------------------------------ */
/* Diagnostics.                                                              */
/* ------------------------------------------------------------------------ */

/* Device time per kernel of the window pipeline, measured with HIP events on
 * the handle's own stream while profiling is on: [0] cell update, [1] spike
 * compaction, [2] SIC sums + plasticity + synapse marking, [3] ordered delivery
 * sums, [4] external drive draws (Poisson counts, noise current). Sums over
 * the windows timed. */
typedef struct asg_kernel_times {
    double ms[5];
    int64_t windows;   /* windows timed */
} asg_kernel_times;

/* every = 0 switches profiling off; every = n >= 1 times one window in n
 * (the first window after the call, then every n-th), so the events' own
 * cost stays off the other windows. Clears the accumulated times. */
asg_status asg_net_set_profiling(asg_net* net, int32_t every);
asg_status asg_net_kernel_times(asg_net* net, asg_kernel_times* out);
asg_status asg_batch_set_profiling(asg_batch* batch, int32_t every);
asg_status asg_batch_kernel_times(asg_batch* batch, asg_kernel_times* out);

enum {
    ASG_MATH_EXP = 0,
    ASG_MATH_LOG = 1,
    ASG_MATH_SIN = 2,
    ASG_MATH_COS = 3,
    ASG_MATH_SQRT = 4,
    ASG_MATH_EXP2_F32 = 5, /* the FP32 path's 2^y of the AdEx spike current */
    ASG_MATH_DIV = 6       /* the FP64 step's division by a constant through its
                              reciprocal: x holds n / 2 pairs (a, b) (n even),
                              y[2i] = a / b, y[2i + 1] = b */
};

/* Evaluates one of the step's functions on the device for n host inputs
 * (no counterpart in the reference; used to show the device rounds the FP64
 * functions exactly as the host build does, and to bound the FP32 path's
 * 2^y). */
asg_status asg_device_math(int32_t fn, const double* x, double* y, int64_t n);

/* asg_device_math without allocating: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the inputs and
 * outputs are staged through a file-scope device array of 2 x 2^20 doubles
 * (16 MiB, held by every GPU the library runs on until the process ends).
 * n above 2^20 returns ASG_ERR_CAPACITY. Calls that overlap on one GPU race
 * for the array; the last call's contents stay in it. */
asg_status asg_device_math_noalloc(int32_t fn, const double* x, double* y, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* ASG_H */

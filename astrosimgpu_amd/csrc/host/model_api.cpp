// C ABI of the host-side model construction (declared in include/asg.h).
#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "asg.h"
#include "builder.h"
#include "config.h"

// Defined with the device code so every layer reports through the same
// asg_last_error().
extern "C" asg_status asg_host_fail(asg_status code, const char* what);

struct asg_model {
    asg_host::Model m;
};

namespace {

asg_host::Overrides to_overrides(const asg_model_overrides* ov) {
    asg_host::Overrides o;
    if (ov != nullptr) {
        o.seed = ov->seed;
        o.sim_time = ov->sim_time;
        o.pre_time = ov->pre_time;
        o.n_astro = ov->n_astro;
        if (ov->assignments != nullptr) o.assignments = ov->assignments;
    }
    return o;
}

template <typename Load>
asg_status make(Load&& load, asg_model** out) {
    if (out == nullptr) return asg_host_fail(ASG_ERR_INVALID, "NULL output pointer");
    try {
        auto model = std::make_unique<asg_model>();
        model->m.cfg = load();
        asg_host::build_model(model->m);
        *out = model.release();
        return ASG_OK;
    } catch (const std::bad_alloc&) {
        return asg_host_fail(ASG_ERR_NOMEM, "host allocation failed while building the network");
    } catch (const std::exception& e) {
        return asg_host_fail(ASG_ERR_CONFIG, e.what());
    }
}

}  // namespace

extern "C" {

asg_status asg_model_load(const char* path, const asg_model_overrides* ov, asg_model** out) {
    if (path == nullptr) return asg_host_fail(ASG_ERR_INVALID, "NULL path");
    const asg_host::Overrides o = to_overrides(ov);
    return make([&] { return asg_host::load_config(path, o); }, out);
}

asg_status asg_model_load_text(const char* json_text, const asg_model_overrides* ov,
                               asg_model** out) {
    if (json_text == nullptr) return asg_host_fail(ASG_ERR_INVALID, "NULL text");
    const asg_host::Overrides o = to_overrides(ov);
    return make([&] { return asg_host::load_config_text(json_text, "<text>", o); }, out);
}

void asg_model_free(asg_model* model) { delete model; }

const asg_net_desc* asg_model_desc(const asg_model* model) {
    return model == nullptr ? nullptr : &model->m.desc;
}

asg_status asg_model_get_info(const asg_model* model, asg_model_info* out) {
    if (model == nullptr || out == nullptr) return asg_host_fail(ASG_ERR_INVALID, "NULL argument");
    const asg_host::Model& m = model->m;
    const asg_host::ModelConfig& c = m.cfg;
    std::memset(out, 0, sizeof(*out));
    out->n_astro = m.stats.n_astro;
    out->n_exc = m.stats.n_exc;
    out->n_inh = m.stats.n_inh;
    out->n_primary_exc = m.stats.n_primary_exc;
    out->n_primary_inh = m.stats.n_primary_inh;
    out->n_neuron_to_astro = m.stats.n_neuron_to_astro;
    out->n_astro_to_neuron = m.stats.n_astro_to_neuron;
    out->dt = c.dt;
    out->pre_sim_time = c.pre_sim_time;
    out->sim_time = c.sim_time;
    out->substeps = c.substeps;
    out->pre_steps = c.pre_steps();
    out->sim_steps = c.sim_steps();
    out->seed = c.seed;
    out->record_spikes = c.record_spikes;
    out->record_astro = c.record_astro;
    out->record_neuron = c.record_neuron;
    out->record_every = c.record_every;
    out->record_astro_max = c.record_astro_max;
    out->record_neuron_max = c.record_neuron_max;
    out->analysis_enabled = c.analysis.enabled;
    out->ca_threshold = c.analysis.ca_threshold;
    out->merge_gap = c.analysis.merge_gap;
    out->hist_window = c.analysis.hist_window;
    out->hist_shift = c.analysis.hist_shift;
    out->SIC_th = c.astro.SIC_th;
    out->ring_slots = std::max(std::max(1, m.d_exc), std::max(m.d_inh, m.d_a2n)) + 1;
    std::strncpy(out->output_dir, c.output_dir.c_str(), sizeof(out->output_dir) - 1);
    return ASG_OK;
}

}  // extern "C"

// ------------------------------------------------------------- analysis ---

#include "analysis.h"

namespace {
void copy_summary(const asg_host::CalciumSummary& s, asg_calcium_summary* out) {
    out->cells = s.cells;
    out->active = s.active;
    out->transients = s.transients;
    out->pairs = s.pairs;
    out->mean_duration = s.mean_duration;
    out->sd_duration = s.sd_duration;
    out->mean_correlation = s.mean_correlation;
    out->threshold = s.threshold;
    out->width = s.width;
    out->shift = s.shift;
}
}  // namespace

extern "C" {

asg_status asg_analyze_calcium(const double* times, const double* ca, int32_t n_cells,
                               int64_t n_samples, double threshold, double merge_gap,
                               double pre_sim_time, double sim_time, double hist_window,
                               double hist_shift, int32_t round6, asg_calcium_summary* out) {
    if (out == nullptr || n_cells < 0 || n_samples < 0 ||
        (n_cells > 0 && n_samples > 0 && (times == nullptr || ca == nullptr))) {
        return asg_host_fail(ASG_ERR_INVALID, "bad argument");
    }
    const asg_host::CalciumSummary s = asg_host::analyze_calcium(
        times, ca, n_cells, n_samples, threshold, merge_gap, pre_sim_time, sim_time, hist_window,
        hist_shift, round6 != 0);
    copy_summary(s, out);
    return ASG_OK;
}

asg_status asg_analyze_calcium_cells(const double* times, const double* ca,
                                     const int64_t* offsets, int32_t n_cells, double threshold,
                                     double merge_gap, double pre_sim_time, double sim_time,
                                     double hist_window, double hist_shift, int32_t round6,
                                     asg_calcium_summary* out) {
    if (out == nullptr || n_cells < 0 || (n_cells > 0 && offsets == nullptr)) {
        return asg_host_fail(ASG_ERR_INVALID, "bad argument");
    }
    for (int32_t c = 0; c < n_cells; ++c) {
        if (offsets[c] < 0 || offsets[c + 1] < offsets[c]) {
            return asg_host_fail(ASG_ERR_INVALID, "offsets must be non-decreasing from 0");
        }
    }
    if (n_cells > 0 && offsets[n_cells] > 0 && (times == nullptr || ca == nullptr)) {
        return asg_host_fail(ASG_ERR_INVALID, "bad argument");
    }
    copy_summary(asg_host::analyze_calcium_cells(times, ca, offsets, n_cells, threshold,
                                                 merge_gap, pre_sim_time, sim_time, hist_window,
                                                 hist_shift, round6 != 0),
                 out);
    return ASG_OK;
}

asg_status asg_detect_events(const double* times, const double* values, int64_t n,
                             double threshold, int32_t merge, double max_gap, double* onset,
                             double* offset, int64_t capacity, int64_t* n_events) {
    if (n < 0 || n_events == nullptr || (n > 0 && (times == nullptr || values == nullptr))) {
        return asg_host_fail(ASG_ERR_INVALID, "bad argument");
    }
    const asg_host::EventList ev =
        asg_host::find_events(times, values, n, threshold, merge != 0, max_gap);
    const int64_t m = static_cast<int64_t>(ev.onset.size());
    *n_events = m;
    if (m > capacity) return asg_host_fail(ASG_ERR_CAPACITY, "output arrays too small");
    std::copy(ev.onset.begin(), ev.onset.end(), onset);
    std::copy(ev.offset.begin(), ev.offset.end(), offset);
    return ASG_OK;
}

double asg_pearson(const double* a, const double* b, int64_t n) {
    if (n <= 0 || a == nullptr || b == nullptr) return 0.0;
    return asg_host::pearson(a, b, n);
}

asg_status asg_sliding_window_rate(const double* events, int64_t n, double start, double end,
                                   double width, double shift, double* out, int64_t capacity,
                                   int64_t* n_out) {
    if (n < 0 || n_out == nullptr || (n > 0 && events == nullptr)) {
        return asg_host_fail(ASG_ERR_INVALID, "bad argument");
    }
    const std::vector<double> h = asg_host::onset_rate_bins(
        std::vector<double>(events, events + n), start, end, width, shift);
    *n_out = static_cast<int64_t>(h.size());
    if (*n_out > capacity) return asg_host_fail(ASG_ERR_CAPACITY, "output array too small");
    for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
    return ASG_OK;
}

}  // extern "C"

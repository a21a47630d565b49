#include "analysis.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace asg_host {

namespace {

// Feeds one trace's samples in time order; records its events.
class RunDetector {
public:
    RunDetector(double threshold, bool merge, double max_gap)
        : th_(threshold), merge_(merge), gap_(max_gap) {}

    void sample(double t, double v) {
        ++samples_;
        if (v >= th_) {
            if (!in_run_) {
                in_run_ = true;
                run_on_ = t;
            }
            run_last_ = t;
        } else if (in_run_) {
            close_run();
        }
    }

    // Ends the trace; returns its events.
    EventList finish() {
        if (in_run_) {
            close_run();
        }
        if (open_) {
            ev_.onset.push_back(open_on_);
            ev_.offset.push_back(open_off_);
            open_ = false;
        }
        if (samples_ < 2) {
            ev_ = EventList();
        }
        return std::move(ev_);
    }

private:
    void close_run() {
        in_run_ = false;
        if (!merge_) {
            ev_.onset.push_back(run_on_);
            ev_.offset.push_back(run_last_);
            return;
        }
        if (open_ && run_on_ - open_off_ < gap_) {
            open_off_ = run_last_;  // closer than max_gap: the same transient
            return;
        }
        if (open_) {
            ev_.onset.push_back(open_on_);
            ev_.offset.push_back(open_off_);
        }
        open_ = true;
        open_on_ = run_on_;
        open_off_ = run_last_;
    }

    double th_;
    bool merge_;
    double gap_;
    int64_t samples_ = 0;
    bool in_run_ = false;
    double run_on_ = 0.0, run_last_ = 0.0;
    bool open_ = false;  // a merged event whose end may still move
    double open_on_ = 0.0, open_off_ = 0.0;
    EventList ev_;
};

// The calcium block of main.cpp:264-313 over the cells' event lists.
CalciumSummary summarise_cells(std::vector<EventList>& cells, double threshold,
                               double pre_sim_time, double sim_time, double hist_window,
                               double hist_shift) {
    CalciumSummary out;
    out.cells = static_cast<int32_t>(cells.size());
    out.threshold = threshold;
    // durations in cell order, then time order (the reference appends them so)
    std::vector<double> dur;
    for (const EventList& e : cells) {
        if (!e.onset.empty()) {
            ++out.active;
        }
        for (size_t k = 0; k < e.onset.size(); ++k) {
            const double d = e.offset[k] - e.onset[k];
            if (d >= 0.0) {
                dur.push_back(d);
            }
        }
    }
    out.transients = static_cast<int64_t>(dur.size());
    if (!dur.empty()) {
        double s = 0.0;
        for (const double d : dur) s += d;
        out.mean_duration = s / static_cast<double>(dur.size());
        double q = 0.0;
        for (const double d : dur) q += (d - out.mean_duration) * (d - out.mean_duration);
        out.sd_duration = std::sqrt(q / static_cast<double>(dur.size()));
    }
    // A window wider than the recording leaves no bins; the reference then
    // uses a tenth of the recorded span (main.cpp:294-299).
    double width = hist_window, shift = hist_shift;
    if (width >= sim_time && sim_time > 0.0) {
        width = sim_time / 10.0;
        shift = width / 5.0;
    }
    out.width = width;
    out.shift = shift;

    // Rows of centred onset rates of the cells with any onset in the window.
    const double start = pre_sim_time, end = pre_sim_time + sim_time;
    std::vector<std::vector<double>> row;
    std::vector<double> sq;  // sum of squares of each row
    std::vector<size_t> bins;
    for (EventList& e : cells) {
        std::vector<double> r = onset_rate_bins(std::move(e.onset), start, end, width, shift);
        double total = 0.0;
        for (const double v : r) total += v;
        if (!(total > 0.0)) {
            continue;  // no onset in the window: takes part in no pair
        }
        bins.push_back(r.size());
        double mean = 0.0;
        for (const double v : r) mean += v;
        mean /= static_cast<double>(r.size());
        double s2 = 0.0;
        for (double& v : r) {
            v -= mean;
            s2 += v * v;
        }
        sq.push_back(s2);
        row.push_back(std::move(r));
    }
    double sum = 0.0;
    int64_t pairs = 0;
    for (size_t i = 0; i + 1 < row.size(); ++i) {
        for (size_t j = i + 1; j < row.size(); ++j) {
            double r = 0.0;
            const size_t n = std::min(bins[i], bins[j]);
            if (n >= 2 && sq[i] > 0.0 && sq[j] > 0.0) {
                double cov = 0.0;
                const double* a = row[i].data();
                const double* b = row[j].data();
                for (size_t k = 0; k < n; ++k) cov += a[k] * b[k];
                r = cov / std::sqrt(sq[i] * sq[j]);
            }
            sum += r;
            ++pairs;
        }
    }
    out.pairs = pairs;
    out.mean_correlation = pairs > 0 ? sum / static_cast<double>(pairs) : 0.0;
    return out;
}

}  // namespace

EventList find_events(const double* times, const double* values, int64_t n, double threshold,
                      bool merge, double max_gap) {
    RunDetector det(threshold, merge, max_gap);
    for (int64_t k = 0; k < n; ++k) {
        det.sample(times[k], values[k]);
    }
    return det.finish();
}

std::vector<double> onset_rate_bins(std::vector<double> onsets, double start, double end,
                                    double width, double shift) {
    std::vector<double> out;
    if (shift <= 0.0 || width <= 0.0 || end <= start) {
        return out;
    }
    const auto bins = static_cast<int64_t>(std::floor((end - start - width) / shift));
    if (bins <= 0) {
        return out;
    }
    out.assign(static_cast<size_t>(bins), 0.0);
    std::sort(onsets.begin(), onsets.end());
    // Both window ends only move forward: lo_i is the first onset >= lo,
    // hi_i the first onset > hi.
    size_t lo_i = 0, hi_i = 0;
    for (int64_t b = 0; b < bins; ++b) {
        const double lo = start + static_cast<double>(b) * shift;
        const double hi = lo + width;
        while (lo_i < onsets.size() && onsets[lo_i] < lo) ++lo_i;
        if (hi_i < lo_i) hi_i = lo_i;
        while (hi_i < onsets.size() && onsets[hi_i] <= hi) ++hi_i;
        out[static_cast<size_t>(b)] = static_cast<double>(hi_i - lo_i) / width;
    }
    return out;
}

double pearson(const double* a, const double* b, int64_t n) {
    if (n < 2) {
        return 0.0;
    }
    double ma = 0.0, mb = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        ma += a[k];
        mb += b[k];
    }
    ma /= static_cast<double>(n);
    mb /= static_cast<double>(n);
    double cov = 0.0, va = 0.0, vb = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        const double x = a[k] - ma, y = b[k] - mb;
        cov += x * y;
        va += x * x;
        vb += y * y;
    }
    return (va > 0.0 && vb > 0.0) ? cov / std::sqrt(va * vb) : 0.0;
}

double round_6_digits(double x) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%g", x);
    return std::strtod(buf, nullptr);
}

CalciumSummary analyze_calcium(const double* times, const double* ca, int32_t n_cells,
                               int64_t n_samples, double threshold, double merge_gap,
                               double pre_sim_time, double sim_time, double hist_window,
                               double hist_shift, bool round6) {
    // One pass over the record in its row order, a detector per cell.
    std::vector<RunDetector> det(static_cast<size_t>(n_cells),
                                 RunDetector(threshold, true, merge_gap));
    for (int64_t k = 0; k < n_samples; ++k) {
        const double t = round6 ? round_6_digits(times[k]) : times[k];
        const double* rowp = ca + k * n_cells;
        for (int32_t c = 0; c < n_cells; ++c) {
            det[static_cast<size_t>(c)].sample(t, round6 ? round_6_digits(rowp[c]) : rowp[c]);
        }
    }
    std::vector<EventList> cells;
    cells.reserve(det.size());
    for (RunDetector& d : det) cells.push_back(d.finish());
    return summarise_cells(cells, threshold, pre_sim_time, sim_time, hist_window, hist_shift);
}

CalciumSummary analyze_calcium_cells(const double* times, const double* ca, const int64_t* offsets,
                                     int32_t n_cells, double threshold, double merge_gap,
                                     double pre_sim_time, double sim_time, double hist_window,
                                     double hist_shift, bool round6) {
    std::vector<EventList> cells;
    cells.reserve(static_cast<size_t>(n_cells));
    for (int32_t c = 0; c < n_cells; ++c) {
        RunDetector d(threshold, true, merge_gap);
        for (int64_t k = offsets[c]; k < offsets[c + 1]; ++k) {
            d.sample(round6 ? round_6_digits(times[k]) : times[k],
                     round6 ? round_6_digits(ca[k]) : ca[k]);
        }
        cells.push_back(d.finish());
    }
    return summarise_cells(cells, threshold, pre_sim_time, sim_time, hist_window, hist_shift);
}

}  // namespace asg_host

#include "analysis.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <limits>

namespace asg_host {

Events detect_events(const std::vector<double>& times, const std::vector<double>& values,
                     double threshold, bool merge, double max_gap) {
    Events out;
    const size_t n = std::min(times.size(), values.size());
    if (n < 2) return out;
    std::vector<double> on, off;
    for (size_t i = 1; i < n; ++i) {
        if (values[i] >= threshold && values[i - 1] < threshold) on.push_back(times[i]);
        if (values[i] < threshold && values[i - 1] >= threshold) off.push_back(times[i - 1]);
    }
    // An event running at either edge of the record is kept.
    if (values[0] >= threshold) on.insert(on.begin(), times[0]);
    if (values[n - 1] >= threshold) off.push_back(times[n - 1]);
    if (merge && !on.empty() && !off.empty()) {
        std::vector<double> m_on{on.front()}, m_off;
        const size_t pairs = std::min(on.size(), off.size());
        for (size_t i = 0; i + 1 < pairs; ++i) {
            if (on[i + 1] - off[i] >= max_gap) {
                m_off.push_back(off[i]);
                m_on.push_back(on[i + 1]);
            }
        }
        m_off.push_back(off.back());
        on.swap(m_on);
        off.swap(m_off);
    }
    out.onset = on;
    out.offset = off;
    const size_t pairs = std::min(on.size(), off.size());
    for (size_t i = 0; i < pairs; ++i) {
        const double d = off[i] - on[i];
        if (d >= 0.0) out.duration.push_back(d);
    }
    return out;
}

std::vector<double> sliding_window_rate(const std::vector<double>& events, double start, double end,
                                        double width, double shift) {
    std::vector<double> hist;
    if (shift <= 0.0 || width <= 0.0 || end <= start) return hist;
    const auto steps = static_cast<int64_t>(std::floor((end - start - width) / shift));
    if (steps <= 0) return hist;
    hist.assign(static_cast<size_t>(steps), 0.0);
    if (events.empty()) return hist;
    std::vector<double> sorted(events);
    std::sort(sorted.begin(), sorted.end());
    for (int64_t i = 0; i < steps; ++i) {
        const double lo = start + static_cast<double>(i) * shift;
        const double hi = lo + width;
        const auto first = std::lower_bound(sorted.begin(), sorted.end(), lo);
        const auto last = std::upper_bound(sorted.begin(), sorted.end(), hi);
        hist[static_cast<size_t>(i)] = static_cast<double>(std::distance(first, last)) / width;
    }
    return hist;
}

double pearson(const std::vector<double>& a, const std::vector<double>& b) {
    const size_t n = std::min(a.size(), b.size());
    if (n < 2) return 0.0;
    double ma = 0.0, mb = 0.0;
    for (size_t i = 0; i < n; ++i) {
        ma += a[i];
        mb += b[i];
    }
    ma /= static_cast<double>(n);
    mb /= static_cast<double>(n);
    double cov = 0.0, va = 0.0, vb = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double da = a[i] - ma, db = b[i] - mb;
        cov += da * db;
        va += da * da;
        vb += db * db;
    }
    if (va <= 0.0 || vb <= 0.0) return 0.0;
    return cov / std::sqrt(va * vb);
}

std::vector<double> pairwise_correlation(const std::vector<std::vector<double>>& per_cell_events,
                                         double start, double end, double width, double shift) {
    const size_t n = per_cell_events.size();
    std::vector<std::vector<double>> hists(n);
    std::vector<char> usable(n, 0);
    for (size_t i = 0; i < n; ++i) {
        hists[i] = sliding_window_rate(per_cell_events[i], start, end, width, shift);
        double sum = 0.0;
        for (const double v : hists[i]) sum += v;
        usable[i] = sum > 0.0 ? 1 : 0;
    }
    std::vector<double> out;
    for (size_t i = 0; i + 1 < n; ++i) {
        if (!usable[i]) continue;
        for (size_t j = i + 1; j < n; ++j) {
            if (!usable[j]) continue;
            out.push_back(pearson(hists[i], hists[j]));
        }
    }
    return out;
}

std::vector<double> nearest_event_distances(const std::vector<std::vector<double>>& per_cell) {
    std::vector<double> out;
    std::vector<std::vector<double>> sorted(per_cell);
    for (auto& s : sorted) std::sort(s.begin(), s.end());
    for (size_t i = 0; i + 1 < sorted.size(); ++i) {
        if (sorted[i].empty()) continue;
        for (size_t j = i + 1; j < sorted.size(); ++j) {
            if (sorted[j].empty()) continue;
            for (const double e : sorted[i]) {
                const auto it = std::lower_bound(sorted[j].begin(), sorted[j].end(), e);
                double best = std::numeric_limits<double>::max();
                if (it != sorted[j].end()) best = std::min(best, std::abs(*it - e));
                if (it != sorted[j].begin()) best = std::min(best, std::abs(*std::prev(it) - e));
                out.push_back(best);
            }
        }
    }
    return out;
}

Summary summarise(const std::vector<double>& values) {
    Summary s;
    s.count = values.size();
    if (values.empty()) return s;
    double sum = 0.0;
    for (const double v : values) sum += v;
    s.mean = sum / static_cast<double>(values.size());
    double acc = 0.0;
    for (const double v : values) {
        const double d = v - s.mean;
        acc += d * d;
    }
    s.stddev = std::sqrt(acc / static_cast<double>(values.size()));
    return s;
}

double round_6_digits(double x) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%g", x);
    return std::strtod(buf, nullptr);
}

namespace {

// The calcium block of main.cpp:264-313 over per-cell traces: transients per
// cell (merged when closer than merge_gap), durations, then the pairwise
// correlation of the cells' onset histograms over the recorded window.
CalciumSummary analyze_traces(const std::vector<std::vector<double>>& times,
                              const std::vector<std::vector<double>>& ca, double threshold,
                              double merge_gap, double pre_sim_time, double sim_time,
                              double hist_window, double hist_shift) {
    CalciumSummary out;
    out.cells = static_cast<int32_t>(ca.size());
    out.threshold = threshold;
    std::vector<std::vector<double>> onsets;
    std::vector<double> durations;
    for (size_t c = 0; c < ca.size(); ++c) {
        const Events ev = detect_events(times[c], ca[c], threshold, true, merge_gap);
        onsets.push_back(ev.onset);
        durations.insert(durations.end(), ev.duration.begin(), ev.duration.end());
        if (!ev.onset.empty()) ++out.active;
    }
    // A window wider than the recording leaves no bins; the reference then
    // falls back to a tenth of the recorded span (main.cpp:294-299).
    double width = hist_window, shift = hist_shift;
    if (width >= sim_time && sim_time > 0.0) {
        width = sim_time / 10.0;
        shift = width / 5.0;
    }
    out.width = width;
    out.shift = shift;
    const std::vector<double> coeff =
        pairwise_correlation(onsets, pre_sim_time, pre_sim_time + sim_time, width, shift);
    out.pairs = static_cast<int64_t>(coeff.size());
    if (!coeff.empty()) {
        double sum = 0.0;
        for (const double v : coeff) sum += v;
        out.mean_correlation = sum / static_cast<double>(coeff.size());
    }
    const Summary dur = summarise(durations);
    out.transients = static_cast<int64_t>(dur.count);
    out.mean_duration = dur.mean;
    out.sd_duration = dur.stddev;
    return out;
}

}  // namespace

CalciumSummary analyze_calcium(const double* times, const double* ca, int32_t n_cells,
                               int64_t n_samples, double threshold, double merge_gap,
                               double pre_sim_time, double sim_time, double hist_window,
                               double hist_shift, bool round6) {
    std::vector<double> t(static_cast<size_t>(n_samples));
    for (int64_t k = 0; k < n_samples; ++k) {
        t[static_cast<size_t>(k)] = round6 ? round_6_digits(times[k]) : times[k];
    }
    std::vector<std::vector<double>> tc(static_cast<size_t>(n_cells), t), vc(static_cast<size_t>(n_cells));
    for (int32_t c = 0; c < n_cells; ++c) {
        std::vector<double>& trace = vc[static_cast<size_t>(c)];
        trace.resize(static_cast<size_t>(n_samples));
        for (int64_t k = 0; k < n_samples; ++k) {
            const double v = ca[k * n_cells + c];
            trace[static_cast<size_t>(k)] = round6 ? round_6_digits(v) : v;
        }
    }
    return analyze_traces(tc, vc, threshold, merge_gap, pre_sim_time, sim_time, hist_window,
                          hist_shift);
}

CalciumSummary analyze_calcium_cells(const double* times, const double* ca, const int64_t* offsets,
                                     int32_t n_cells, double threshold, double merge_gap,
                                     double pre_sim_time, double sim_time, double hist_window,
                                     double hist_shift, bool round6) {
    std::vector<std::vector<double>> tc(static_cast<size_t>(n_cells)), vc(static_cast<size_t>(n_cells));
    for (int32_t c = 0; c < n_cells; ++c) {
        for (int64_t k = offsets[c]; k < offsets[c + 1]; ++k) {
            tc[static_cast<size_t>(c)].push_back(round6 ? round_6_digits(times[k]) : times[k]);
            vc[static_cast<size_t>(c)].push_back(round6 ? round_6_digits(ca[k]) : ca[k]);
        }
    }
    return analyze_traces(tc, vc, threshold, merge_gap, pre_sim_time, sim_time, hist_window,
                          hist_shift);
}

}  // namespace asg_host

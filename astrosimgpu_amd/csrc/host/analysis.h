// Post-run calcium analysis (reference src/analysis.cpp, include/astrosimgpu/
// analysis.hpp): threshold-crossing event detection with merging, sliding-
// window onset rates, Pearson correlation, pairwise mean, summary statistics.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace asg_host {

struct Events {
    std::vector<double> onset, offset, duration;
};

// analysis.cpp:10-64.
Events detect_events(const std::vector<double>& times, const std::vector<double>& values,
                     double threshold, bool merge, double max_gap);
// analysis.cpp:66-96 (window [start, end]).
std::vector<double> sliding_window_rate(const std::vector<double>& events, double start, double end,
                                        double width, double shift);
// analysis.cpp:98-123.
double pearson(const std::vector<double>& a, const std::vector<double>& b);
// analysis.cpp:136-166; returns every coefficient, the mean is taken by the caller.
std::vector<double> pairwise_correlation(const std::vector<std::vector<double>>& per_cell_events,
                                         double start, double end, double width, double shift);
// analysis.cpp:168-199.
std::vector<double> nearest_event_distances(const std::vector<std::vector<double>>& per_cell);

struct Summary {
    double mean = 0.0, stddev = 0.0;
    size_t count = 0;
};
// analysis.cpp:201-220.
Summary summarise(const std::vector<double>& values);

// What the reference prints under "Calcium analysis" (main.cpp:266-317) for
// sampled traces: per cell events with merging, then the pairwise
// correlation of windowed onset rates, with the short-run window fallback.
struct CalciumSummary {
    int32_t cells = 0, active = 0;
    int64_t transients = 0, pairs = 0;
    double mean_duration = 0.0, sd_duration = 0.0, mean_correlation = 0.0;
    double threshold = 0.0, width = 0.0, shift = 0.0;
};

// ca is [n_samples][n_cells]; times[n_samples] in ms. round6 rounds times and
// values to 6 significant digits first, as the reference's CSV round trip
// does (recorder.cpp:47-50 writes with the default stream precision, main.cpp
// reads the file back).
CalciumSummary analyze_calcium(const double* times, const double* ca, int32_t n_cells,
                               int64_t n_samples, double threshold, double merge_gap,
                               double pre_sim_time, double sim_time, double hist_window,
                               double hist_shift, bool round6);

// The same over per-cell traces of their own lengths: cell c's samples are
// times/ca[offsets[c] .. offsets[c + 1]) (main.cpp:46-74 reads each cell's
// rows into its own series).
CalciumSummary analyze_calcium_cells(const double* times, const double* ca, const int64_t* offsets,
                                     int32_t n_cells, double threshold, double merge_gap,
                                     double pre_sim_time, double sim_time, double hist_window,
                                     double hist_shift, bool round6);

double round_6_digits(double x);

}  // namespace asg_host

// Post-run calcium analysis: what the reference's command line prints under
// "Calcium analysis" (src/main.cpp:264-313, on the functions of
// src/analysis.cpp). Its results are reproduced bit for bit; the passes are
// this file's own:
//  * transients of every cell in one sweep over the [samples][cells] record
//    (a run detector per cell, merging as it goes);
//  * onset-rate bins by a two-pointer sweep of each cell's sorted onsets;
//  * the pairwise Pearson coefficients from rows of centred bins computed
//    once per cell, so a pair costs one dot product. Every sum keeps the
//    reference's order (bins ascending, pairs in (i, j) order), which is what
//    the bits depend on.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace asg_host {

// Threshold-crossing events of one trace (analysis.cpp:10-64): an event is a
// maximal run of samples at or above the threshold, from the time of its
// first sample to the time of its last; with merge, a run that starts less
// than max_gap after the previous one ended extends it. Fewer than two
// samples give no events.
struct EventList {
    std::vector<double> onset, offset;
};
EventList find_events(const double* times, const double* values, int64_t n, double threshold,
                      bool merge, double max_gap);

// Onset rate in sliding windows [start + i shift, start + i shift + width]
// (both ends included), i < floor((end - start - width) / shift), in events
// per ms (analysis.cpp:66-96).
std::vector<double> onset_rate_bins(std::vector<double> onsets, double start, double end,
                                    double width, double shift);

// Pearson's r of two series of n values; 0 when n < 2 or either has no
// variance (analysis.cpp:98-123).
double pearson(const double* a, const double* b, int64_t n);

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

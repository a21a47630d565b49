// TEST INFRASTRUCTURE ONLY: runs the reference's own calcium analysis
// functions (src/analysis.cpp, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/analysis_driver) on traces read from a
// binary file, and prints the quantities of the reference's "Calcium
// analysis" block (src/main.cpp:266-313) at full precision, so that tests can
// compare the product's analysis with the reference's bit for bit.
//
// Input (little-endian): int32 n_cells; double threshold, merge_gap,
// pre_sim_time, sim_time, hist_window, hist_shift; int64 offsets[n_cells + 1];
// double times[offsets[n]]; double ca[offsets[n]].
// Output: one line "active transients pairs mean_duration sd_duration
// mean_correlation width shift" with the doubles in %a.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "astrosimgpu/analysis.hpp"

using namespace astrosimgpu;

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: analysis_driver INPUT\n");
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t n = 0;
    double p[6];
    if (std::fread(&n, 4, 1, f) != 1 || std::fread(p, 8, 6, f) != 6) return 2;
    std::vector<int64_t> off(static_cast<size_t>(n) + 1);
    if (std::fread(off.data(), 8, off.size(), f) != off.size()) return 2;
    const size_t total = static_cast<size_t>(off[n]);
    std::vector<double> t(total), v(total);
    if (std::fread(t.data(), 8, total, f) != total || std::fread(v.data(), 8, total, f) != total)
        return 2;
    std::fclose(f);
    const double threshold = p[0], max_gap = p[1], pre = p[2], sim = p[3];
    double width = p[4], shift = p[5];
    // main.cpp:276-305
    vec<vec<real>> onsets;
    vec<real> durations;
    std::size_t active = 0;
    for (int32_t c = 0; c < n; ++c) {
        const vec<real> tc(t.begin() + off[c], t.begin() + off[c + 1]);
        const vec<real> vc(v.begin() + off[c], v.begin() + off[c + 1]);
        const EventTrain ev = detect_events(tc, vc, threshold, true, max_gap);
        onsets.push_back(ev.onset);
        durations.insert(durations.end(), ev.duration.begin(), ev.duration.end());
        if (!ev.onset.empty()) ++active;
    }
    Interval window{pre, pre + sim};
    if (width >= sim && sim > 0.0) {
        width = sim / 10.0;
        shift = width / 5.0;
    }
    const PairwiseCorrelation corr = pairwise_correlation(onsets, window, width, shift);
    const Summary dur = summarise(durations);
    std::printf("%zu %zu %zu %a %a %a %a %a\n", active, dur.count, corr.coefficient.size(),
                dur.mean, dur.stddev, corr.mean(), width, shift);
    return 0;
}

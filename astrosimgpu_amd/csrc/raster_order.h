// Restores the raster order -- step, then source (the reference sorts each
// step's spikes by source, neuron.cpp:243-248) -- of the entries an FP32
// network's delivery pass appended during one or more windows. Host code.
//
// The delivery pass reserves raster space with one atomic per (wavefront,
// step), so a window's entries come in whatever order its wavefronts ran.
// Sparse firing makes nearly every entry a reservation of its own, so the
// entries are sorted outright: each (step, source) pair is packed into one
// integer key, step offset above source, and the keys go through an LSD
// radix sort, 8 bits a pass -- linear in the entries, where a comparison
// sort of the same entries costs several times more.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace asg {

// Sorts step[0, n) / src[0, n) by (step, source). keys and tmp are scratch
// the caller keeps between calls.
inline void order_raster(int64_t* step, uint32_t* src, size_t n, std::vector<uint64_t>& keys,
                         std::vector<uint64_t>& tmp) {
    if (n < 2) {
        return;
    }
    int64_t smin = step[0], smax = step[0];
    uint32_t rmax = src[0];
    bool sorted = true;
    for (size_t i = 1; i < n; ++i) {
        smin = std::min(smin, step[i]);
        smax = std::max(smax, step[i]);
        rmax = std::max(rmax, src[i]);
        if (step[i] < step[i - 1] || (step[i] == step[i - 1] && src[i] < src[i - 1])) {
            sorted = false;
        }
    }
    if (sorted) {
        return;
    }
    auto bits = [](uint64_t v) {  // bits to hold 0 .. v
        int b = 1;
        while (b < 64 && (v >> b) != 0) ++b;
        return b;
    };
    const int sb = bits(rmax);
    const uint64_t span = static_cast<uint64_t>(smax) - static_cast<uint64_t>(smin);
    const int total = sb + bits(span);
    if (total > 64 || n < 64) {
        std::vector<std::pair<int64_t, uint32_t>> e(n);
        for (size_t q = 0; q < n; ++q) e[q] = {step[q], src[q]};
        std::sort(e.begin(), e.end());
        for (size_t q = 0; q < n; ++q) {
            step[q] = e[q].first;
            src[q] = e[q].second;
        }
        return;
    }
    keys.resize(n);
    tmp.resize(n);
    for (size_t i = 0; i < n; ++i) {
        keys[i] = ((static_cast<uint64_t>(step[i]) - static_cast<uint64_t>(smin)) << sb) | src[i];
    }
    uint64_t* a = keys.data();
    uint64_t* b = tmp.data();
    for (int shift = 0; shift < total; shift += 8) {
        size_t count[257] = {};
        for (size_t i = 0; i < n; ++i) ++count[((a[i] >> shift) & 0xFF) + 1];
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (size_t i = 0; i < n; ++i) b[count[(a[i] >> shift) & 0xFF]++] = a[i];
        std::swap(a, b);
    }
    const uint64_t src_mask = (sb == 64) ? ~0ULL : ((1ULL << sb) - 1ULL);
    for (size_t i = 0; i < n; ++i) {
        step[i] = static_cast<int64_t>(static_cast<uint64_t>(smin) + (a[i] >> sb));
        src[i] = static_cast<uint32_t>(a[i] & src_mask);
    }
}

}  // namespace asg

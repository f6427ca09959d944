// Guide tables for the area-light triangle search.
//
// Distribution1D::SampleReuse (basic.hpp:469-475) picks triangle
// upper_bound(cdf, u) - 1 of an area light.  A binary search over the whole
// CDF is a chain of dependent loads (9 for a 320-triangle light), and a
// wavefront waits for its longest chain.  The guide of a CDF with m = n + 1
// entries has G + 1 entries for G = 2^bits >= n buckets:
//
//     guide[j] = upper_bound(cdf, j / G),  j = 0..G.
//
// u * G is exact in FP32 and FP64 (G is a power of two, at most 2^24), so
// bucket j = floor(u G) holds exactly the u in [j/G, (j+1)/G), and since the
// predicate cdf[k] > u is monotone in k for a sorted CDF,
// guide[j] <= upper_bound(cdf, u) <= guide[j + 1].  A binary search of
// [guide[j], guide[j + 1]) then returns upper_bound(cdf, u) itself (or
// guide[j + 1] when no entry in the range exceeds u, which is then the
// answer), in about one step for lights of similar triangle areas.  A CDF
// that is not sorted or holds a NaN (an emitter of zero area) gets bits = 0,
// one bucket [0, m): the unguided search with its own probe sequence.
#pragma once

#include <stdint.h>

#include "ngi_common.h"

namespace ngi {

#define NGI_GUIDE_MAX_BITS 24

// Bucket bits of the guide of a sorted, NaN-free CDF over n triangles.
NGI_HD uint32_t guide_bits(uint32_t n) {
    uint32_t b = n <= 1u ? 0u : 32u - (uint32_t)__builtin_clz(n - 1u);
    return b > NGI_GUIDE_MAX_BITS ? NGI_GUIDE_MAX_BITS : b;
}

// True when cdf[0..m) is non-decreasing and NaN-free (the guide applies).
template <typename T>
NGI_HD bool guide_sorted(const T* cdf, uint32_t m) {
    for (uint32_t k = 0; k < m; ++k) {
        if (!(cdf[k] == cdf[k])) return false;
        if (k > 0 && cdf[k] < cdf[k - 1]) return false;
    }
    return true;
}

// Fills guide[0..2^bits] for cdf[0..m) (one monotone sweep; bits = 0 gives
// the single bucket [0, m)).
template <typename T>
NGI_HD void guide_build(const T* cdf, uint32_t m, uint32_t bits, uint32_t* guide) {
    const uint32_t G = 1u << bits;
    if (bits == 0) {
        guide[0] = 0;
        guide[1] = m;
        return;
    }
    uint32_t k = 0;
    for (uint32_t j = 0; j <= G; ++j) {
        const T x = (T)j / (T)G;
        while (k < m && !(cdf[k] > x)) ++k;
        guide[j] = k;
    }
}

// upper_bound(cdf, cdf + m, u) - cdf for u in [0, 1), through the guide.
template <typename T>
NGI_HD uint32_t guided_upper_bound(const T* cdf, const uint32_t* guide, uint32_t bits, T u) {
    const uint32_t G = 1u << bits;
    const T x = u * (T)G;
    uint32_t j = x > T(0) ? (uint32_t)x : 0u;
    j = j < G ? j : G - 1u;
    uint32_t lo = guide[j], hi = guide[j + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

}  // namespace ngi

// Combinatorial description of the subset-DP plan (host and device).
//
// The reference builds the plan by a DFS over reachable bitmasks, sorts them by (popcount, mask)
// and ranks them with a Dict (src/dp.jl:248-304).  The same array can be written without any set
// or map, because the reachable subsets of one size have a closed form:
//
//   a non-empty S with p = min(S) and r = K - |S| indices removed is reachable iff
//   p - 1 <= r <= 2(p - 1)                                               (src/dp.jl:16-27)
//
// For a fixed size m = |S| (r = K - m, even), the upper bound p <= r + 1 holds for every
// m-subset with min p, so the reachable m-subsets are exactly ALL m-subsets of the ground set
// {p_lo, ..., K} with p_lo = ceil(r/2) + 1.  Ordering same-size masks by value is colexicographic
// order, whose rank is the combinatorial number system: for positions x_1 < ... < x_m inside the
// ground set, rank = sum_i C(x_i, i).  A child S \ {p, j} lives in the level two sizes down,
// whose ground set starts one index higher, so every position shifts down by one.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define EGG_HD __host__ __device__
#else
#define EGG_HD
#endif

namespace eggman {

constexpr int kMaxK = 64;

// C(n, k) for 0 <= n, k <= 64; every entry fits in uint64 (the largest is C(64, 32)).
struct Binom {
    uint64_t c[kMaxK + 1][kMaxK + 1];
};

// Pascal's triangle, evaluated at compile time (usable for a __constant__ table).
constexpr Binom make_binom() {
    Binom b{};
    for (int n = 0; n <= kMaxK; n++) {
        for (int k = 0; k <= kMaxK; k++) b.c[n][k] = 0;
        b.c[n][0] = 1;
        for (int k = 1; k <= n; k++) b.c[n][k] = b.c[n - 1][k - 1] + (k <= n - 1 ? b.c[n - 1][k] : 0);
    }
    return b;
}

// 1-based flat upper-triangle index of the pair (i < j), src/unrolled.jl:60.
EGG_HD inline int64_t pair_index(int64_t i, int64_t j, int64_t K) {
    return (i - 1) * K - (i * (i + 1)) / 2 + j;
}

// First index of the ground set of level size m (degree K): ceil((K - m) / 2) + 1.
EGG_HD inline int ground_lo(int K, int m) { return (K - m + 1) / 2 + 1; }

// Unrank: the positions x[0..m-1] (ascending, 0-based inside the ground set) of the colex
// rank `rank` among m-subsets of an n-set.
EGG_HD inline void colex_unrank(const uint64_t (*C)[kMaxK + 1], int n, int m, uint64_t rank,
                                int *x) {
    int c = n - 1;
    for (int i = m; i >= 1; i--) {
        while (C[c][i] > rank) c--;
        x[i - 1] = c;
        rank -= C[c][i];
        c--;
    }
}

// Advance x (m positions inside an n-set) to the next subset in colex order.
EGG_HD inline void colex_next(int m, int *x) {
    int i = 0;
    while (i + 1 < m && x[i] + 1 == x[i + 1]) i++;
    x[i]++;
    for (int t = 0; t < i; t++) x[t] = t;
}

// Transitions of one state, in the reference's order (ascending partner j, src/dp.jl:283-292).
// x holds the state's m positions inside its ground set starting at index `lo` (1-based);
// `child_first` is the 1-based number of the first state of the child level.  Writes
// (child << bits) | pair for the m - 1 transitions into out[0..m-2].
template <typename W>
EGG_HD inline void state_transitions(const uint64_t (*C)[kMaxK + 1], int K, int m, int lo,
                                     const int *x, int64_t child_first, int bits, W *out) {
    // child rank for removing position t (1-based t = 2..m):
    //   sum_{i=2}^{t-1} C(x_i - 1, i - 1) + sum_{i=t+1}^{m} C(x_i - 1, i - 2)
    uint64_t suffix = 0;
    for (int i = 3; i <= m; i++) suffix += C[x[i - 1] - 1][i - 2];
    uint64_t prefix = 0;
    const int64_t p = lo + x[0];
    for (int t = 2; t <= m; t++) {
        const int64_t j = lo + x[t - 1];
        const int64_t child = child_first + (int64_t)(prefix + suffix);
        out[t - 2] = (W)((child << bits) | pair_index(p, j, K));
        // move position t from the suffix sum into the prefix sum for the next t
        prefix += C[x[t - 1] - 1][t - 1];
        if (t + 1 <= m) suffix -= C[x[t] - 1][t - 1];
    }
}

}  // namespace eggman

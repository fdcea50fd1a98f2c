/*
 * dp_oracle.c -- CPU restatement of TheEggman.jl's subset-DP hafnian path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the tests, the smoke check and the
 * CPU-baseline leg of the benchmark load this library, and only as the checker (or as the
 * timed CPU baseline).  The product library never links or calls it.
 *
 * What is restated, with the reference lines each function follows:
 *   oracle_dp_counts      src/dp.jl:192-208   closed-form (nstates, ntrans)
 *   oracle_build_plan     src/dp.jl:223-304   guards, DFS over reachable masks, sort by
 *                                             (popcount, mask), rank map (a hash map, as the
 *                                             reference's Dict), packed CSR, levels
 *   pair_index            src/unrolled.jl:60  flat upper-triangle index, 1-based
 *   dp_states_*           src/dp.jl:344-363   the inner loop, plan order, no FMA
 *   haf_dp_one_*          src/dp.jl:408-442   level driver with the _dp_level_chunks policy
 *                                             (src/dp.jl:370-388)
 *   oracle_haf_dp (B>1)   src/hafnian.jl:1226-1242  _batch_cpu!: B split into min(nthreads,B)
 *                                             contiguous chunks, each instance serial
 * Complex multiply is written as (ac-bd, ad+bc), which is Julia's `*` for Complex, and the
 * file is compiled with -ffp-contract=off so no multiply-add is fused (Julia does not contract
 * a*b+c by default).
 *
 * Extra element types used only as checkers: a 128-bit wrapping integer walk (exact for integer
 * matrices whose hafnian fits in 127 bits, since the DP is a ring computation) and a
 * __float128 walk (113-bit mantissa) standing in for the BigFloat oracle of
 * test/runtests.jl:318-332.  oracle_perm_ryser gives exact permanents for the identity
 * haf([0 B; B^T 0]) = perm(B) quoted at README.md:238-244.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <quadmath.h>
#include <omp.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

enum {
    OR_F32 = 1, OR_F64 = 2, OR_C32 = 3, OR_C64 = 4, OR_I128 = 5, OR_Q = 6, OR_CQ = 7
};
enum { OR_OK = 0, OR_EINVAL = -1, OR_ENOMEM = -2, OR_EPLAN = -3 };

/* ---------------------------------------------------------------- counts (dp.jl:192-208) */

static u128 binom_u128(int n, int k) {
    if (k < 0 || k > n) return 0;
    u128 c = 1;
    for (int i = 1; i <= k; i++) c = c * (u128)(n - k + i) / (u128)i;
    return c;
}

int oracle_dp_counts(int K, int64_t *nstates_out, int64_t *ntrans_out) {
    if (K < 0 || K > 64) return OR_EINVAL;
    u128 nstates = 1, ntrans = 0;                    /* the empty set */
    for (int p = 1; p <= K; p++) {
        int lo = p - 1;
        int hi = 2 * (p - 1) < K - 1 ? 2 * (p - 1) : K - 1;
        int r = (lo % 2 == 0) ? lo : lo + 1;
        while (r <= hi) {
            int k = K - r - 1;                       /* |S|-1 = transitions out of S */
            u128 c = binom_u128(K - p, k);
            nstates += c;
            ntrans += c * (u128)k;
            r += 2;
        }
    }
    *nstates_out = (int64_t)nstates;
    *ntrans_out = (int64_t)ntrans;
    return OR_OK;
}

/* ------------------------------------------------------- packing (dp.jl:170-180, unrolled:60) */

static inline int pair_bits(int index_bytes) { return index_bytes == 4 ? 9 : 12; }
static inline int64_t pair_index(int64_t i, int64_t j, int64_t K) {
    return (i - 1) * K - (i * (i + 1)) / 2 + j;
}

/* ------------------------------------------------------------------ a set of uint64 masks */

typedef struct {
    uint64_t *keys;          /* 0 marks an empty slot; the mask 0 is tracked by has_zero */
    uint64_t cap, len;
    int has_zero;
} maskset;

static inline uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
    return x;
}

/* returns 1 if inserted, 0 if present */
static int maskset_insert(maskset *s, uint64_t m) {
    if (m == 0) {
        if (s->has_zero) return 0;
        s->has_zero = 1; s->len++;
        return 1;
    }
    uint64_t h = mix64(m) & (s->cap - 1);
    for (;;) {
        if (s->keys[h] == 0) { s->keys[h] = m; s->len++; return 1; }
        if (s->keys[h] == m) return 0;
        h = (h + 1) & (s->cap - 1);
    }
}

static int cmp_popmask(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    int px = __builtin_popcountll(x), py = __builtin_popcountll(y);
    if (px != py) return px < py ? -1 : 1;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* slot of a nonzero mask present in the set (read-only, so threads may share it) */
static uint64_t maskset_slot(const maskset *s, uint64_t m) {
    uint64_t h = mix64(m) & (s->cap - 1);
    while (s->keys[h] != m) h = (h + 1) & (s->cap - 1);
    return h;
}

static void st(void *a, int w, int64_t i, int64_t v) {
    if (w == 4) ((int32_t *)a)[i] = (int32_t)v; else ((int64_t *)a)[i] = v;
}
static int64_t ld(const void *a, int w, int64_t i) {
    return w == 4 ? (int64_t)((const int32_t *)a)[i] : ((const int64_t *)a)[i];
}

/*
 * Guards of dp.jl:223-243 (the memory guard uses this machine's physical memory), then the
 * enumeration of dp.jl:248-304.  `starts` (nstates+1), `trans` (ntrans) and `levels` (K/2+2)
 * are caller-allocated arrays of int32 (index_bytes 4) or int64 (8), filled with the
 * reference's 1-based values.
 */
int oracle_build_plan(int K, int index_bytes, void *starts, void *trans, void *levels) {
    if (K < 0 || K > 64) return OR_EINVAL;
    if (index_bytes != 4 && index_bytes != 8) return OR_EINVAL;
    int64_t nstates, ntrans;
    oracle_dp_counts(K, &nstates, &ntrans);
    int bits = pair_bits(index_bytes);
    int64_t npairs = (int64_t)K * (K - 1) / 2;
    int64_t imax = index_bytes == 4 ? INT32_MAX : INT64_MAX;
    if (npairs > ((int64_t)1 << bits) - 1) return OR_EPLAN;
    if (nstates > (imax >> bits)) return OR_EPLAN;
    if (ntrans > imax) return OR_EPLAN;
    double bytes = (double)(ntrans + nstates + 1) * index_bytes;
    double total = (double)sysconf(_SC_PHYS_PAGES) * (double)sysconf(_SC_PAGESIZE);
    if (2.0 * bytes >= total) return OR_ENOMEM;

    uint64_t full = K == 0 ? 0 : (~(uint64_t)0) >> (64 - K);
    maskset set;
    set.cap = 16;
    while (set.cap < (uint64_t)nstates * 2) set.cap <<= 1;
    set.keys = (uint64_t *)calloc(set.cap, sizeof(uint64_t));
    set.len = 0; set.has_zero = 0;
    uint64_t *stack = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nstates + 1));
    if (!set.keys || !stack) { free(set.keys); free(stack); return OR_ENOMEM; }
    int64_t sp = 0;
    maskset_insert(&set, full);
    stack[sp++] = full;
    while (sp > 0) {
        uint64_t m = stack[--sp];
        if (m == 0) continue;
        int i = __builtin_ctzll(m) + 1;
        uint64_t rest = m & ~((uint64_t)1 << (i - 1));
        uint64_t r = rest;
        while (r != 0) {
            int j = __builtin_ctzll(r) + 1;
            r &= ~((uint64_t)1 << (j - 1));
            uint64_t sub = rest & ~((uint64_t)1 << (j - 1));
            if (maskset_insert(&set, sub)) stack[sp++] = sub;
        }
    }
    free(stack);
    if ((int64_t)set.len != nstates) { free(set.keys); return OR_EPLAN; }

    uint64_t *masks = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nstates);
    if (!masks) { free(set.keys); return OR_ENOMEM; }
    int64_t n = 0;
    if (set.has_zero) masks[n++] = 0;
    for (uint64_t h = 0; h < set.cap; h++)
        if (set.keys[h]) masks[n++] = set.keys[h];
    qsort(masks, (size_t)nstates, sizeof(uint64_t), cmp_popmask);

    /* the rank map `number` of dp.jl:272 (a Dict{M,I} there): the set's own hash table, with a
       1-based state number per slot; the empty mask is state 1 (it sorts first) */
    int64_t *number = (int64_t *)malloc(sizeof(int64_t) * (size_t)set.cap);
    if (!number) { free(set.keys); free(masks); return OR_ENOMEM; }
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nstates; s++)
        if (masks[s] != 0) number[maskset_slot(&set, masks[s])] = s + 1;

    /* starts: a state S has |S| - 1 transitions (none for the empty set) */
    st(starts, index_bytes, 0, 1);
    for (int64_t s = 0; s < nstates; s++) {
        int pc = __builtin_popcountll(masks[s]);
        st(starts, index_bytes, s + 1, ld(starts, index_bytes, s) + (pc > 0 ? pc - 1 : 0));
    }
    /* transitions in ascending j (dp.jl:286-292); every state writes its own slice, so the
       states are emitted in parallel with the same result */
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t s = 0; s < nstates; s++) {
        uint64_t m = masks[s];
        if (m == 0) continue;
        int64_t t = ld(starts, index_bytes, s) - 1;
        int i = __builtin_ctzll(m) + 1;
        uint64_t rest = m & ~((uint64_t)1 << (i - 1));
        uint64_t r = rest;
        while (r != 0) {
            int j = __builtin_ctzll(r) + 1;
            r &= ~((uint64_t)1 << (j - 1));
            uint64_t sub = rest & ~((uint64_t)1 << (j - 1));
            int64_t c = sub == 0 ? 1 : number[maskset_slot(&set, sub)];
            st(trans, index_bytes, t, (c << bits) | pair_index(i, j, K));
            t++;
        }
    }
    free(number);
    free(set.keys);
    int64_t nl = 0;
    st(levels, index_bytes, nl++, 1);
    for (int64_t s = 1; s < nstates; s++)
        if (__builtin_popcountll(masks[s]) != __builtin_popcountll(masks[s - 1]))
            st(levels, index_bytes, nl++, s + 1);
    st(levels, index_bytes, nl++, nstates + 1);
    free(masks);
    return nl == K / 2 + 2 ? OR_OK : OR_EPLAN;
}

/* -------------------------------------------------------------------- element types */

typedef struct { double re, im; } c64;
typedef struct { float re, im; } c32;
typedef struct { __float128 re, im; } cq;

#define ZERO_REAL(T) ((T)0)
static inline c64 c64_muladd(c64 acc, c64 p, c64 h) {
    double re = p.re * h.re - p.im * h.im;       /* Julia: (ac - bd, ad + bc) */
    double im = p.re * h.im + p.im * h.re;
    acc.re = acc.re + re; acc.im = acc.im + im;
    return acc;
}
static inline c32 c32_muladd(c32 acc, c32 p, c32 h) {
    float re = p.re * h.re - p.im * h.im;
    float im = p.re * h.im + p.im * h.re;
    acc.re = acc.re + re; acc.im = acc.im + im;
    return acc;
}
static inline cq cq_muladd(cq acc, cq p, cq h) {
    __float128 re = p.re * h.re - p.im * h.im;
    __float128 im = p.re * h.im + p.im * h.re;
    acc.re = acc.re + re; acc.im = acc.im + im;
    return acc;
}

/*
 * _dp_states! (dp.jl:344-363) for one element type and one index width.
 *   H, P    : 0-based arrays of T; H[s-1] is state s, P[q-1] is pair q.
 *   lo, hi  : 1-based inclusive state range.
 */
#define DEF_STATES(NAME, T, IT, BITS, ZERO, STEP)                                          \
    static void NAME(T *H, const T *P, const IT *starts, const IT *trans, int64_t lo,      \
                     int64_t hi) {                                                         \
        const IT mask = ((IT)1 << (BITS)) - 1;                                             \
        for (int64_t s = lo; s <= hi; s++) {                                               \
            T acc = ZERO;                                                                  \
            for (int64_t k = (int64_t)starts[s - 1]; k <= (int64_t)starts[s] - 1; k++) {   \
                IT t = trans[k - 1];                                                       \
                const T p = P[(int64_t)(t & mask) - 1];                                    \
                const T h = H[(int64_t)(t >> (BITS)) - 1];                                 \
                STEP;                                                                      \
            }                                                                              \
            H[s - 1] = acc;                                                                \
        }                                                                                  \
    }

#define STEP_REAL acc = acc + p * h
#define STEP_C64 acc = c64_muladd(acc, p, h)
#define STEP_C32 acc = c32_muladd(acc, p, h)
#define STEP_CQ acc = cq_muladd(acc, p, h)

static const c64 c64_zero = {0.0, 0.0};
static const c32 c32_zero = {0.0f, 0.0f};
static const cq cq_zero = {0, 0};

DEF_STATES(states_f64_i32, double, int32_t, 9, 0.0, STEP_REAL)
DEF_STATES(states_f64_i64, double, int64_t, 12, 0.0, STEP_REAL)
DEF_STATES(states_f32_i32, float, int32_t, 9, 0.0f, STEP_REAL)
DEF_STATES(states_f32_i64, float, int64_t, 12, 0.0f, STEP_REAL)
DEF_STATES(states_c64_i32, c64, int32_t, 9, c64_zero, STEP_C64)
DEF_STATES(states_c64_i64, c64, int64_t, 12, c64_zero, STEP_C64)
DEF_STATES(states_c32_i32, c32, int32_t, 9, c32_zero, STEP_C32)
DEF_STATES(states_c32_i64, c32, int64_t, 12, c32_zero, STEP_C32)
DEF_STATES(states_u128_i32, u128, int32_t, 9, (u128)0, STEP_REAL)
DEF_STATES(states_u128_i64, u128, int64_t, 12, (u128)0, STEP_REAL)
DEF_STATES(states_q_i32, __float128, int32_t, 9, (__float128)0, STEP_REAL)
DEF_STATES(states_q_i64, __float128, int64_t, 12, (__float128)0, STEP_REAL)
DEF_STATES(states_cq_i32, cq, int32_t, 9, cq_zero, STEP_CQ)
DEF_STATES(states_cq_i64, cq, int64_t, 12, cq_zero, STEP_CQ)

/*
 * _haf_dp (dp.jl:408-442): H[1] = one; for each level L = 2..length(levels)-1 evaluate
 * levels[L]:levels[L+1]-1, split into _dp_level_chunks(ntrans_L, n_L, nthreads) equal chunks
 * when nthreads > 1 and the plan has at least _DP_PARALLEL_MIN = 60000 transitions.
 * Returns H[nstates].  The chunk boundaries never change a state's sum, so the result is the
 * same for every thread count.
 */
#define DEF_HAF_ONE(NAME, T, IT, STATES, ONE)                                             \
    static int NAME(const IT *starts, const IT *trans, const IT *levels, int64_t nlevels,  \
                    int64_t nstates, const T *P, int nthreads, T *out) {                   \
        T *H = (T *)malloc(sizeof(T) * (size_t)nstates);                                   \
        if (!H) return OR_ENOMEM;                                                          \
        H[0] = ONE;                                                                        \
        int64_t ntr = (int64_t)starts[nstates] - 1;                                        \
        int parallel = nthreads > 1 && ntr >= 60000;                                       \
        for (int64_t L = 2; L <= nlevels - 1; L++) {                                       \
            int64_t lo = (int64_t)levels[L - 1];                                           \
            int64_t hi = (int64_t)levels[L] - 1;                                           \
            int64_t n = hi - lo + 1;                                                       \
            int64_t ntl = (int64_t)starts[hi] - (int64_t)starts[lo - 1];                   \
            int64_t nch = 1;                                                               \
            if (parallel) {                                                                \
                int64_t cap = nthreads < n ? nthreads : n;                                 \
                nch = ntl / 2000;                                                          \
                if (nch < 1) nch = 1;                                                      \
                if (nch > cap) nch = cap;                                                  \
            }                                                                              \
            if (nch == 1) {                                                                \
                STATES(H, P, starts, trans, lo, hi);                                       \
            } else {                                                                       \
                _Pragma("omp parallel for num_threads(nch) schedule(static, 1)")           \
                for (int64_t c = 1; c <= nch; c++) {                                       \
                    int64_t clo = lo + ((c - 1) * n) / nch;                                \
                    int64_t chi = lo + (c * n) / nch - 1;                                  \
                    if (clo <= chi) STATES(H, P, starts, trans, clo, chi);                 \
                }                                                                          \
            }                                                                              \
        }                                                                                  \
        *out = H[nstates - 1];                                                             \
        free(H);                                                                           \
        return OR_OK;                                                                      \
    }

static const c64 c64_one = {1.0, 0.0};
static const c32 c32_one = {1.0f, 0.0f};
static const cq cq_one = {1, 0};

DEF_HAF_ONE(haf_f64_i32, double, int32_t, states_f64_i32, 1.0)
DEF_HAF_ONE(haf_f64_i64, double, int64_t, states_f64_i64, 1.0)
DEF_HAF_ONE(haf_f32_i32, float, int32_t, states_f32_i32, 1.0f)
DEF_HAF_ONE(haf_f32_i64, float, int64_t, states_f32_i64, 1.0f)
DEF_HAF_ONE(haf_c64_i32, c64, int32_t, states_c64_i32, c64_one)
DEF_HAF_ONE(haf_c64_i64, c64, int64_t, states_c64_i64, c64_one)
DEF_HAF_ONE(haf_c32_i32, c32, int32_t, states_c32_i32, c32_one)
DEF_HAF_ONE(haf_c32_i64, c32, int64_t, states_c32_i64, c32_one)
DEF_HAF_ONE(haf_u128_i32, u128, int32_t, states_u128_i32, (u128)1)
DEF_HAF_ONE(haf_u128_i64, u128, int64_t, states_u128_i64, (u128)1)
DEF_HAF_ONE(haf_q_i32, __float128, int32_t, states_q_i32, (__float128)1)
DEF_HAF_ONE(haf_q_i64, __float128, int64_t, states_q_i64, (__float128)1)
DEF_HAF_ONE(haf_cq_i32, cq, int32_t, states_cq_i32, cq_one)
DEF_HAF_ONE(haf_cq_i64, cq, int64_t, states_cq_i64, cq_one)

/* size in bytes of one P element and of one output element, per dtype */
static int p_size(int dtype) {
    switch (dtype) {
    case OR_F32: return 4;
    case OR_F64: return 8;
    case OR_C32: return 8;
    case OR_C64: return 16;
    case OR_I128: return 8;        /* int64 matrix entries */
    case OR_Q: return 8;           /* double entries, promoted */
    case OR_CQ: return 16;         /* complex double entries, promoted */
    }
    return 0;
}
static int out_size(int dtype) {
    switch (dtype) {
    case OR_F32: return 4;
    case OR_F64: return 8;
    case OR_C32: return 8;
    case OR_C64: return 16;
    case OR_I128: return 16;       /* (lo, hi) uint64 limbs, two's complement */
    case OR_Q: return 16;          /* (hi, lo) double-double */
    case OR_CQ: return 32;         /* (re_hi, re_lo, im_hi, im_lo) */
    }
    return 0;
}

/* one instance: P0 is its contiguous pair vector (npairs elements of the dtype's P type) */
static int haf_one(int index_bytes, const void *starts, const void *trans, const void *levels,
                   int64_t nlevels, int64_t nstates, int64_t npairs, int dtype, const void *P0,
                   int nthreads, void *out) {
    int w8 = index_bytes == 8;
#define CALL(F32N, F64N, P, O) (w8 ? F64N((const int64_t *)starts, (const int64_t *)trans,    \
                                          (const int64_t *)levels, nlevels, nstates, P,       \
                                          nthreads, O)                                        \
                                   : F32N((const int32_t *)starts, (const int32_t *)trans,    \
                                          (const int32_t *)levels, nlevels, nstates, P,       \
                                          nthreads, O))
    switch (dtype) {
    case OR_F64: return CALL(haf_f64_i32, haf_f64_i64, (const double *)P0, (double *)out);
    case OR_F32: return CALL(haf_f32_i32, haf_f32_i64, (const float *)P0, (float *)out);
    case OR_C64: return CALL(haf_c64_i32, haf_c64_i64, (const c64 *)P0, (c64 *)out);
    case OR_C32: return CALL(haf_c32_i32, haf_c32_i64, (const c32 *)P0, (c32 *)out);
    case OR_I128: {
        u128 *Pu = (u128 *)malloc(sizeof(u128) * (size_t)npairs);
        if (!Pu) return OR_ENOMEM;
        for (int64_t q = 0; q < npairs; q++) Pu[q] = (u128)(i128)((const int64_t *)P0)[q];
        u128 v = 0;
        int rc = CALL(haf_u128_i32, haf_u128_i64, Pu, &v);
        free(Pu);
        ((uint64_t *)out)[0] = (uint64_t)v;
        ((uint64_t *)out)[1] = (uint64_t)(v >> 64);
        return rc;
    }
    case OR_Q: {
        __float128 *Pq = (__float128 *)malloc(sizeof(__float128) * (size_t)npairs);
        if (!Pq) return OR_ENOMEM;
        for (int64_t q = 0; q < npairs; q++) Pq[q] = (__float128)((const double *)P0)[q];
        __float128 v = 0;
        int rc = CALL(haf_q_i32, haf_q_i64, Pq, &v);
        free(Pq);
        double hi = (double)v;
        ((double *)out)[0] = hi;
        ((double *)out)[1] = (double)(v - (__float128)hi);
        return rc;
    }
    case OR_CQ: {
        cq *Pq = (cq *)malloc(sizeof(cq) * (size_t)npairs);
        if (!Pq) return OR_ENOMEM;
        for (int64_t q = 0; q < npairs; q++) {
            Pq[q].re = (__float128)((const double *)P0)[2 * q];
            Pq[q].im = (__float128)((const double *)P0)[2 * q + 1];
        }
        cq v = cq_zero;
        int rc = CALL(haf_cq_i32, haf_cq_i64, Pq, &v);
        free(Pq);
        double rh = (double)v.re, ih = (double)v.im;
        ((double *)out)[0] = rh;
        ((double *)out)[1] = (double)(v.re - (__float128)rh);
        ((double *)out)[2] = ih;
        ((double *)out)[3] = (double)(v.im - (__float128)ih);
        return rc;
    }
    }
#undef CALL
    return OR_EINVAL;
}

/*
 * Hafnians of B instances sharing one plan.  P is laid out [pair][batch] with the batch index
 * fastest -- exactly Julia's column-major Matrix{T}(B, npairs) from _gather_pairs_batch
 * (src/dp.jl:464-512).  B == 1 follows _haf_dp's level chunking with `nthreads`; B > 1 follows
 * _batch_cpu! (src/hafnian.jl:1226-1242): min(nthreads, B) contiguous chunks of instances, each
 * instance evaluated serially (hafnian(As[b]; nthreads = 1), src/hafnian.jl:1298-1300).
 */
int oracle_haf_dp(int K, int index_bytes, const void *starts, const void *trans,
                  const void *levels, int dtype, const void *P, int64_t B, int nthreads,
                  void *out) {
    if (K < 2 || K > 64 || (K & 1)) return OR_EINVAL;
    if (index_bytes != 4 && index_bytes != 8) return OR_EINVAL;
    int ps = p_size(dtype), os = out_size(dtype);
    if (!ps || B < 0) return OR_EINVAL;
    if (nthreads < 1) nthreads = 1;
    int64_t nstates, ntrans;
    oracle_dp_counts(K, &nstates, &ntrans);
    int64_t npairs = (int64_t)K * (K - 1) / 2;
    int64_t nlevels = K / 2 + 2;
    if (B == 0) return OR_OK;
    if (B == 1)
        return haf_one(index_bytes, starts, trans, levels, nlevels, nstates, npairs, dtype, P,
                       nthreads, out);
    int64_t nch = nthreads < B ? nthreads : B;
    int rc_all = OR_OK;
#pragma omp parallel for num_threads(nch) schedule(static, 1)
    for (int64_t c = 1; c <= nch; c++) {
        int64_t lo = ((c - 1) * B) / nch + 1, hi = (c * B) / nch;
        char *P0 = (char *)malloc((size_t)ps * (size_t)npairs);
        if (!P0) { rc_all = OR_ENOMEM; continue; }
        for (int64_t b = lo; b <= hi; b++) {
            for (int64_t q = 0; q < npairs; q++)
                memcpy(P0 + q * ps, (const char *)P + (q * B + (b - 1)) * ps, (size_t)ps);
            int rc = haf_one(index_bytes, starts, trans, levels, nlevels, nstates, npairs, dtype,
                             P0, 1, (char *)out + (b - 1) * os);
            if (rc) rc_all = rc;
        }
        free(P0);
    }
    return rc_all;
}

/*
 * Exact permanent of an n x n integer matrix (row-major int64), by Ryser's formula with a Gray
 * code, in wrapping 128-bit arithmetic: perm(B) = (-1)^n sum_S (-1)^|S| prod_i sum_{j in S} b_ij.
 * Used only for the identity haf([0 B; B^T 0]) = perm(B).  out = (lo, hi) limbs.
 */
int oracle_perm_ryser(const int64_t *Bm, int n, uint64_t *out) {
    if (n < 0 || n > 30) return OR_EINVAL;
    if (n == 0) { out[0] = 1; out[1] = 0; return OR_OK; }
    i128 rowsum[30];
    for (int i = 0; i < n; i++) rowsum[i] = 0;
    i128 total = 0;
    uint64_t gray_prev = 0;
    for (uint64_t k = 1; k < ((uint64_t)1 << n); k++) {
        uint64_t gray = k ^ (k >> 1);
        uint64_t diff = gray ^ gray_prev;
        int j = __builtin_ctzll(diff);
        int add = (gray & diff) != 0;
        for (int i = 0; i < n; i++) {
            if (add) rowsum[i] += Bm[i * n + j]; else rowsum[i] -= Bm[i * n + j];
        }
        u128 prod = 1;
        for (int i = 0; i < n; i++) prod *= (u128)rowsum[i];
        int sz = __builtin_popcountll(gray);
        if (((n - sz) & 1) == 0) total += (i128)prod; else total -= (i128)prod;
        gray_prev = gray;
    }
    out[0] = (uint64_t)(u128)total;
    out[1] = (uint64_t)((u128)total >> 64);
    return OR_OK;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

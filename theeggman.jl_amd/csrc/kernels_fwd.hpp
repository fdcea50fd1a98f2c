// Forward level kernels of the subset DP on gfx950.
//
// One launch evaluates one level: every state s of the level and every instance b of the batch,
//
//     H[s][b] = sum_{k in plan order} P[pair_k][b] * H[child_k][b]          (src/dp.jl:354-361)
//
// Data layout: state values are [state][batch] and pair values [pair][batch], batch fastest,
// the layout of the reference's device path (ext/TheEggmanKernelAbstractionsExt.jl:56-64).
// A group of BT consecutive lanes (BT = 1..64, a power of two) owns one state and BT
// consecutive instances of it.  With BT = 64 a whole wave64 follows one state, so the packed
// transition word is wave-uniform and is fetched with scalar loads while the P and H reads are
// 64 consecutive elements (one 512-byte burst for FP64).  Smaller BT serve small batches; BT = 1
// is the plain path with one lane per state.  Each lane accumulates its own instance serially in
// plan order, so results do not depend on BT, on the chunking of the batch or on the launch
// shape, and equal the reference CPU walk bit for bit.
//
// All states of a level have the same out-degree k (src/dp.jl:384-386), so the transitions of
// state r of the level are trans_lvl[r*k .. r*k+k-1]: the device never reads `starts`.
#pragma once
#include "common.hpp"

#pragma clang fp contract(off)

namespace eggman {

// Grid-stride bound for level launches: enough blocks to fill 256 CUs many times over.
constexpr int64_t kMaxBlocks = 1 << 16;
constexpr int kBlock = 256;

__device__ inline int64_t wave_uniform(int64_t v) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)(uint64_t)v);
    const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

template <typename W> struct Packing;
template <> struct Packing<int32_t> { static constexpr int bits = 9; };    // src/dp.jl:170
template <> struct Packing<int64_t> { static constexpr int bits = 12; };   // src/dp.jl:171

// H of the level's first state row `Hc` is child state `cfirst` (1-based number); rows are `ldh`
// elements apart.  Ho receives this level's rows, `ldo` apart.  P rows are `ldp` apart and
// P/Hc/Ho already point at the chunk's first instance.
template <int DT, typename W, int BT>
__global__ __launch_bounds__(kBlock) void k_level_fwd(
    const W *__restrict__ trans_lvl, int64_t nst, int k,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t cfirst, int64_t ldh,
    typename Elt<DT>::H *__restrict__ Ho, int64_t ldo, int64_t nb, int64_t ntiles) {
    using E = Elt<DT>;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    const int lane = (int)(threadIdx.x % BT);
    const int64_t ngroups = nst * ntiles;
    const int64_t gstride = (int64_t)gridDim.x * (kBlock / BT);
    for (int64_t g = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / BT; g < ngroups; g += gstride) {
        int64_t gg = g;
        if (BT == 64) gg = wave_uniform(g);
        const int64_t s = ntiles == 1 ? gg : gg / ntiles;
        const int64_t tile = gg - s * ntiles;
        const int64_t b = tile * BT + lane;
        if (b >= nb) continue;
        const W *tw = trans_lvl + s * (int64_t)k;
        typename E::H acc = E::zero();
        constexpr int U = 8;   // loads issued together before they are summed in order
        auto batch = [&](int i, auto n) {
            constexpr int N = decltype(n)::value;
            W t[N];
            typename E::P pv[N];
            typename E::H hv[N];
#pragma unroll
            for (int u = 0; u < N; u++) t[u] = tw[i + u];
#pragma unroll
            for (int u = 0; u < N; u++) {
                pv[u] = P[((int64_t)(t[u] & mask) - 1) * ldp + b];
                hv[u] = Hc[((int64_t)(t[u] >> bits) - cfirst) * ldh + b];
            }
#pragma unroll
            for (int u = 0; u < N; u++) acc = E::muladd(acc, pv[u], hv[u]);
        };
        int i = 0;
        for (; i + U <= k; i += U) batch(i, std::integral_constant<int, U>{});
        batch_tail<U>(k - i, [&](auto n) { batch(i, n); });
        Ho[s * ldo + b] = acc;
    }
}

// One or two instances (BT = 1 or 2 lanes per state, one instance each), with the instances' pair
// values staged in LDS: a lane's pair reads are scattered over the pair rows, and from LDS they
// leave the vector memory path to the plan words and the child values.  Each lane sums its
// (state, instance) in plan order, so the values are the other kernels' bits.  P (pair rows ldp
// apart), Hc and Ho (rows ldh and ldo apart) point at the first instance.  (Staging the block's
// plan words in LDS as well, to coalesce their reads, cost more in barriers than it saved.)
constexpr int kPlainBlock = 256;

template <int DT, typename W, int BT>
__global__ __launch_bounds__(kPlainBlock) void k_level_fwd_plain(
    const W *__restrict__ trans_lvl, int64_t nst, int k,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp, int npairs,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t cfirst, int64_t ldh,
    typename Elt<DT>::H *__restrict__ Ho, int64_t ldo, int64_t nb) {
    using E = Elt<DT>;
    using PT = typename E::P;
    using HT = typename E::H;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    PT *Ps = reinterpret_cast<PT *>(lds_raw);   // pair q of instance b at Ps[q * BT + b]
    for (int e = (int)threadIdx.x; e < npairs * BT; e += kPlainBlock) {
        const int q = e / BT, bb = e % BT;
        Ps[e] = bb < nb ? P[(int64_t)q * ldp + bb] : PT{};
    }
    __syncthreads();
    const int b = (int)(threadIdx.x % BT);
    const PT *Pl = Ps + b - BT;                    // pair row q (1-based) of this lane: Pl[q * BT]
    const int64_t per_block = kPlainBlock / BT;
    for (int64_t s = (int64_t)blockIdx.x * per_block + threadIdx.x / BT; s < nst;
         s += (int64_t)gridDim.x * per_block) {
        if (b >= nb) continue;
        const W *tw = trans_lvl + s * (int64_t)k;
        HT acc = E::zero();
        constexpr int U = 16;   // child loads issued together before they are summed in order
        auto batch = [&](int i, auto n) {
            constexpr int N = decltype(n)::value;
            W t[N];
            HT h[N];
#pragma unroll
            for (int u = 0; u < N; u++) t[u] = tw[i + u];
#pragma unroll
            for (int u = 0; u < N; u++) h[u] = Hc[((int64_t)(t[u] >> bits) - cfirst) * ldh + b];
#pragma unroll
            for (int u = 0; u < N; u++) acc = E::muladd(acc, Pl[(int)(t[u] & mask) * BT], h[u]);
        };
        int i = 0;
        for (; i + U <= k; i += U) batch(i, std::integral_constant<int, U>{});
        batch_tail<U>(k - i, [&](auto n) { batch(i, n); });
        Ho[s * ldo + b] = acc;
    }
}

// ---------------------------------------------------------------------------------------
// Batched level kernel with the level's pair values staged in LDS and an XCD-aware mapping.
//
// Every state of level L has min(S) = p in [p_lo, p_hi] (plan_rank.hpp), so its transitions only
// read the pair rows p_lo..p_hi, which are one contiguous range [qlo, qhi] of pair indices --
// 105..140 of the 378 pairs at K = 28.  A workgroup owns one 64-instance tile, copies that range
// of P for its tile into LDS once, and then walks many states of the level with one wave per
// state: the only global reads left per transition are the wave-uniform plan word (scalar) and
// the 64 consecutive child values.
//
// Workgroups are dealt to the 8 XCDs round-robin (workgroup j runs on XCD j % 8).  The mapping
// below gives each XCD whole tiles when there are at least 8, and a contiguous share of the
// states of a tile otherwise, so a tile's child rows are only ever cached in one XCD's L2.  The
// workgroups of one XCD sweep the states of a tile in plan order, which is colexicographic in
// the subset, and consecutive states share most of their children -- that sweep is what lets
// the L2 serve most child reads.
struct LdsMap {
    int64_t ntiles;      // instance tiles in the chunk (64 wide, or 16 for the narrow kernel)
    int64_t per_xcd;     // workgroups per XCD (R)
    int64_t tiles_xcd;   // tiles per XCD when ntiles >= 8 (ceil(ntiles / 8)), else 0
};

// Arguments of one level of one group of instances sharing a plan.
template <int DT, typename W>
struct LevelArgs {
    const W *trans_lvl;                       // the level's first transition
    int64_t nst;                              // states in the level
    int k;                                    // transitions per state
    int qlo, nq;                              // pair rows [qlo, qlo + nq) read by the level
    const typename Elt<DT>::P *P;             // pair values, rows ldp apart
    int64_t ldp;
    const typename Elt<DT>::H *Hc;            // the level below; row 0 is state cfirst
    int64_t cfirst, ldh;
    typename Elt<DT>::H *Ho;                  // this level, rows ldo apart
    int64_t ldo, nb;                          // instances in the group
    LdsMap map;
    // child addresses fit 32-bit arithmetic: every packed child number is below 2^24 and
    // (child number + 1) * ldh * sizeof(H) below 2^32 (checked by the launcher)
    bool addr32;
    // parent pairs (k_level_fwd_lds_vec2 only): the level's items, item i = state number within
    // the level | (second parent follows) << 31; the degree and the level's smallest min(S)
    const uint32_t *items;
    int64_t nitems;
    int K, plo;
    // the 16-byte-load kernels store a lane's instances as one 16-byte store: Ho is 16-byte
    // aligned and ldo a multiple of 16 / sizeof(H) (checked by the launcher)
    bool vstore;
};

// Launch-time check behind LevelArgs::addr32 and the other kernels' 32-bit address paths: a
// child number c <= nstates of the plan, read from a transition word, addresses its row with
// one 24-bit multiply, c * row_bytes + lane offset, when that stays below 2^32.
inline bool child_addr32(int64_t nstates, int64_t row_bytes) {
    return nstates < (int64_t(1) << 24) && row_bytes < (int64_t(1) << 24) &&
           (nstates + 2) * row_bytes < (int64_t(1) << 32);
}

// The body of the LDS kernel for workgroup `j` of a launch that gives this level
// 8 * map.per_xcd workgroups (j counts from a multiple of 8, so j % 8 is still the XCD).
template <int DT, typename W, int WPB>
__device__ inline void level_lds_body(const LevelArgs<DT, W> &a, int64_t j,
                                      typename Elt<DT>::P *lds) {
    using E = Elt<DT>;
    using PT = typename E::P;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    const LdsMap &map = a.map;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = (int)(threadIdx.x >> 6);
    const int64_t xcd = j & 7, r = j >> 3;
    int64_t tile, s_begin, s_end, bi, bstride;
    if (map.tiles_xcd > 0) {
        const int64_t ti = r % map.tiles_xcd;
        tile = xcd + 8 * ti;
        bi = r / map.tiles_xcd;
        bstride = map.per_xcd / map.tiles_xcd;
        s_begin = 0;
        s_end = a.nst;
        if (tile >= map.ntiles) return;
    } else {
        tile = xcd % map.ntiles;
        const int64_t group = xcd / map.ntiles;
        const int64_t ngroups = (8 - tile + map.ntiles - 1) / map.ntiles;
        s_begin = group * a.nst / ngroups;
        s_end = (group + 1) * a.nst / ngroups;
        bi = r;
        bstride = map.per_xcd;
    }
    const int64_t b = tile * 64 + lane;
    const bool live = b < a.nb;

    // stage pair rows [qlo, qlo + nq) of this tile
    for (int i = wave; i < a.nq; i += WPB)
        lds[i * 64 + lane] = live ? a.P[(int64_t)(a.qlo + i) * a.ldp + b] : PT{};
    __syncthreads();

    const int64_t step = bstride * WPB;
    // child loads issued together before they are summed in order; wider batches (8, 12, 16)
    // measured slower on MI355X at K = 24..32 when the tail still went one transition at a time
    constexpr int U = 4;
    for (int64_t s0 = s_begin + bi * WPB; s0 < s_end; s0 += step) {
        const int64_t s = wave_uniform(s0 + wave);
        if (s >= s_end) break;
        const W *tw = a.trans_lvl + s * (int64_t)a.k;
        typename E::H acc = E::zero();
        if (live) {
            // N transitions from i: their N child loads issued together, then summed in order
            auto batch = [&](int i, auto n) {
                constexpr int N = decltype(n)::value;
                W t[N];
                typename E::H h[N];
#pragma unroll
                for (int u = 0; u < N; u++) t[u] = tw[i + u];
#pragma unroll
                for (int u = 0; u < N; u++)
                    h[u] = a.Hc[((int64_t)(t[u] >> bits) - a.cfirst) * a.ldh + b];
#pragma unroll
                for (int u = 0; u < N; u++)
                    acc = E::muladd(acc, lds[((int)(t[u] & mask) - 1 - a.qlo) * 64 + lane], h[u]);
            };
            int i = 0;
            for (; i + U <= a.k; i += U) batch(i, std::integral_constant<int, U>{});
            // the last k mod U transitions as one batch too
            batch_tail<U>(a.k - i, [&](auto n) { batch(i, n); });
            a.Ho[s * a.ldo + b] = acc;
        }
    }
}

// Narrow-tile variant: 16-instance tiles, a wave covering 4 consecutive states x 16 instances
// (lane 16j + i: state s0 + j, instance i of the tile).  One state's row of a tile is then 16
// values -- one 128-byte line for 8-byte values -- instead of 64, so an XCD's L2 holds four
// times as many child rows, and with 8 tiles per pass each XCD owns one tile and walks every
// state of it.  The level is bound by L2 misses, which this cuts.  Each lane still sums one
// (state, instance) in plan order: the results are the same bits as the 64-wide kernel's.
template <int DT, typename W, int WPB, bool A32>
__device__ inline void level_lds_narrow_body_a(const LevelArgs<DT, W> &a, int64_t j,
                                               typename Elt<DT>::P *lds) {
    using E = Elt<DT>;
    using PT = typename E::P;
    using HT = typename E::H;
    constexpr int bits = Packing<W>::bits;
    constexpr int TW = 16;
    const LdsMap &map = a.map;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int slot = lane >> 4, li = lane & 15;
    const int64_t xcd = j & 7, r = j >> 3;
    int64_t tile, s_begin, s_end, bi, bstride;
    if (map.tiles_xcd > 0) {
        const int64_t ti = r % map.tiles_xcd;
        tile = xcd + 8 * ti;
        bi = r / map.tiles_xcd;
        bstride = map.per_xcd / map.tiles_xcd;
        s_begin = 0;
        s_end = a.nst;
        if (tile >= map.ntiles) return;
    } else {
        tile = xcd % map.ntiles;
        const int64_t group = xcd / map.ntiles;
        const int64_t ngroups = (8 - tile + map.ntiles - 1) / map.ntiles;
        s_begin = group * a.nst / ngroups;
        s_end = (group + 1) * a.nst / ngroups;
        bi = r;
        bstride = map.per_xcd;
    }
    const int64_t b = tile * TW + li;
    const bool live = b < a.nb;
    const int64_t bl = live ? b : a.nb - 1;   // dead lanes read a valid column, store nothing

    // stage pair rows [qlo, qlo + nq) of this tile: row i at lds[i * TW]
    for (int e = (int)threadIdx.x; e < a.nq * TW; e += WPB * 64) {
        const int64_t bb = tile * TW + (e & (TW - 1));
        lds[e] = bb < a.nb ? a.P[(int64_t)(a.qlo + e / TW) * a.ldp + bb] : PT{};
    }
    __syncthreads();

    const int64_t step = bstride * WPB * 4;
    constexpr int U = 4;
    // child row c (1-based packed number) of this lane: hbase + c * hrow + hlane (32-bit form);
    // pair row q (1-based) of this lane in LDS: lbase + q * TW * sizeof(P)
    const char *hbase = reinterpret_cast<const char *>(a.Hc) - a.cfirst * a.ldh * (int64_t)sizeof(HT);
    const uint32_t hrow = (uint32_t)(a.ldh * (int64_t)sizeof(HT));
    const uint32_t hlane = (uint32_t)(bl * (int64_t)sizeof(HT));
    const char *lbase = reinterpret_cast<const char *>(lds) +
                        ((int64_t)li - (int64_t)(1 + a.qlo) * TW) * (int64_t)sizeof(PT);
    // a state's words are read once, lane li holding words li and 16 + li, and handed to the
    // group's lanes by a shuffle within the 16-lane group (k <= 31 for the Int32 layout; the
    // Int64 layout, k <= 63, reads four words per lane)
    constexpr int NW = sizeof(W) == 4 ? 2 : 4;
    for (int64_t s0 = s_begin + (bi * WPB + wave) * 4; s0 < s_end; s0 += step) {
        const int64_t s = s0 + slot;
        const bool ok = s < s_end;
        const W *tw = a.trans_lvl + (ok ? s : s_end - 1) * (int64_t)a.k;
        W wv[NW];
#pragma unroll
        for (int c = 0; c < NW; c++) wv[c] = tw[min(16 * c + li, a.k - 1)];
        auto word = [&](int t) -> W {   // t uniform
            W w = 0;
#pragma unroll
            for (int c = 0; c < NW; c++)
                if ((t >> 4) == c) w = (W)__shfl(wv[c], t & 15, 16);
            return w;
        };
        auto load_child = [&](W t) -> HT {
            if constexpr (A32) {
                const uint32_t off = __umul24((uint32_t)((uint64_t)t >> bits), hrow) + hlane;
                return *reinterpret_cast<const HT *>(hbase + (uint64_t)off);
            } else {
                return a.Hc[((int64_t)(t >> bits) - a.cfirst) * a.ldh + bl];
            }
        };
        auto pair_val = [&](W t) -> PT {
            const uint32_t q = __builtin_amdgcn_ubfe((uint32_t)t, 0, bits);
            return *reinterpret_cast<const PT *>(lbase + q * (uint32_t)(TW * sizeof(PT)));
        };
        HT acc = E::zero();
        // N transitions from i: their N child loads issued together, then summed in order
        auto batch = [&](int i, auto n) {
            constexpr int N = decltype(n)::value;
            W t[N];
            HT h[N];
#pragma unroll
            for (int u = 0; u < N; u++) t[u] = word(i + u);
#pragma unroll
            for (int u = 0; u < N; u++) h[u] = load_child(t[u]);
#pragma unroll
            for (int u = 0; u < N; u++) acc = E::muladd(acc, pair_val(t[u]), h[u]);
        };
        int i = 0;
        for (; i + U <= a.k; i += U) batch(i, std::integral_constant<int, U>{});
        // the last k mod U transitions as one batch too
        batch_tail<U>(a.k - i, [&](auto n) { batch(i, n); });
        if (ok && live) a.Ho[s * a.ldo + b] = acc;
    }
}

// The narrow body with 32-bit child addresses when the launch allows them.  The 64-bit form
// spends about nine integer instructions per transition on the address (three of them
// quarter-rate multiplies), which bounds the kernel by instruction issue; the 32-bit form is a
// shift and a 24-bit multiply-add from a uniform base.
template <int DT, typename W, int WPB>
__device__ inline void level_lds_narrow_body(const LevelArgs<DT, W> &a, int64_t j,
                                             typename Elt<DT>::P *lds) {
    if (a.addr32) level_lds_narrow_body_a<DT, W, WPB, true>(a, j, lds);
    else level_lds_narrow_body_a<DT, W, WPB, false>(a, j, lds);
}

// ---------------------------------------------------------------------------------------
// Vector variant of the narrow kernel: every load is 16 bytes.  A lane holds VEC = 16 / sizeof(H)
// consecutive instances (two for FP64), 8 lanes cover a state's 128-byte row of a tile of
// TW = 8 * VEC instances, and a wave walks 8 consecutive states.  The L1 and its address unit
// process a load four lanes per cycle whatever the width, which bounds the 8-byte kernels:
// 16-byte loads move twice the data per cycle and halve the address work per value.  Each
// lane keeps VEC accumulators, every one summed in plan order, so the results are the same bits
// as the other kernels'.
//
// The level rows must be 16-byte aligned with TW | ldh, and a row's padding past the batch must
// be readable: lanes past the batch read it and store nothing.
template <int DT, typename W, int WPB, bool A32, int UB>
__device__ inline void level_lds_vec_body_a(const LevelArgs<DT, W> &a, int64_t j,
                                            typename Elt<DT>::P *lds) {
    using E = Elt<DT>;
    using PT = typename E::P;
    using HT = typename E::H;
    constexpr int bits = Packing<W>::bits;
    constexpr int VEC = 16 / (int)sizeof(HT);
    constexpr int LPS = 8;                  // lanes per state
    constexpr int TW = LPS * VEC;           // instances per tile
    constexpr int SPW = 64 / LPS;           // states per wave
    struct alignas(16) HV { HT v[VEC]; };
    struct alignas(VEC * sizeof(PT)) PV { PT v[VEC]; };
    const LdsMap &map = a.map;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int slot = lane / LPS, li = lane % LPS;
    const int64_t xcd = j & 7, r = j >> 3;
    int64_t tile, s_begin, s_end, bi, bstride;
    if (map.tiles_xcd > 0) {
        const int64_t ti = r % map.tiles_xcd;
        tile = xcd + 8 * ti;
        bi = r / map.tiles_xcd;
        bstride = map.per_xcd / map.tiles_xcd;
        s_begin = 0;
        s_end = a.nst;
        if (tile >= map.ntiles) return;
    } else {
        tile = xcd % map.ntiles;
        const int64_t group = xcd / map.ntiles;
        const int64_t ngroups = (8 - tile + map.ntiles - 1) / map.ntiles;
        s_begin = group * a.nst / ngroups;
        s_end = (group + 1) * a.nst / ngroups;
        bi = r;
        bstride = map.per_xcd;
    }
    const int64_t b = tile * TW + li * VEC;   // this lane's first instance

    // stage pair rows [qlo, qlo + nq) of this tile: row i at lds[i * TW]
    for (int e = (int)threadIdx.x; e < a.nq * TW; e += WPB * 64) {
        const int64_t bb = tile * TW + (e % TW);
        lds[e] = bb < a.nb ? a.P[(int64_t)(a.qlo + e / TW) * a.ldp + bb] : PT{};
    }
    __syncthreads();

    // child row c (1-based packed number) of this lane: hbase + c * hrow + hlane (32-bit form);
    // pair row q (1-based) of this lane in LDS: lbase + q * TW * sizeof(P)
    const char *hbase = reinterpret_cast<const char *>(a.Hc) - a.cfirst * a.ldh * (int64_t)sizeof(HT);
    const uint32_t hrow = (uint32_t)(a.ldh * (int64_t)sizeof(HT));
    const uint32_t hlane = (uint32_t)(b * (int64_t)sizeof(HT));
    const char *lbase = reinterpret_cast<const char *>(lds) +
                        ((int64_t)li * VEC - (int64_t)(1 + a.qlo) * TW) * (int64_t)sizeof(PT);
    auto load_child = [&](W t) -> HV {
        if constexpr (A32) {
            const uint32_t off = __umul24((uint32_t)((uint64_t)t >> bits), hrow) + hlane;
            return *reinterpret_cast<const HV *>(hbase + (uint64_t)off);
        } else {
            return *reinterpret_cast<const HV *>(a.Hc + ((int64_t)(t >> bits) - a.cfirst) * a.ldh + b);
        }
    };
    auto pair_val = [&](W t) -> PV {
        const uint32_t q = __builtin_amdgcn_ubfe((uint32_t)t, 0, bits);
        return *reinterpret_cast<const PV *>(lbase + q * (uint32_t)(TW * sizeof(PT)));
    };

    const int64_t step = bstride * WPB * SPW;
    constexpr int U = UB;   // child loads of a state issued together (4, or 6 for Tune::deep)
    // a state's words are read once, lane li holding words li, li + 8, ..., and handed to the
    // group's lanes by a shuffle within the 8-lane group
    constexpr int NW = sizeof(W) == 4 ? 4 : 8;   // k <= 31 (Int32 layout) or 63 (Int64)
    for (int64_t s0 = s_begin + (bi * WPB + wave) * SPW; s0 < s_end; s0 += step) {
        const int64_t s = s0 + slot;
        const bool ok = s < s_end;
        const W *tw = a.trans_lvl + (ok ? s : s_end - 1) * (int64_t)a.k;
        W wv[NW];
#pragma unroll
        for (int c = 0; c < NW; c++)
            if (c * LPS < a.k) wv[c] = tw[min(c * LPS + li, a.k - 1)];
        auto word = [&](int t) -> W {   // t uniform
            W w = 0;
#pragma unroll
            for (int c = 0; c < NW; c++)
                if (t / LPS == c) w = (W)__shfl(wv[c], t % LPS, LPS);
            return w;
        };
        HT acc[VEC];
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[v] = E::zero();
        // N transitions from i: their N child loads issued together, then summed in order
        auto batch = [&](int i, auto n) {
            constexpr int N = decltype(n)::value;
            W t[N];
            HV h[N];
#pragma unroll
            for (int u = 0; u < N; u++) t[u] = word(i + u);
#pragma unroll
            for (int u = 0; u < N; u++) h[u] = load_child(t[u]);
#pragma unroll
            for (int u = 0; u < N; u++) {
                const PV p = pair_val(t[u]);
#pragma unroll
                for (int v = 0; v < VEC; v++) acc[v] = E::muladd(acc[v], p.v[v], h[u].v[v]);
            }
        };
        int i = 0;
        for (; i + U <= a.k; i += U) batch(i, std::integral_constant<int, U>{});
        // the last k mod U transitions as one batch too (k is uniform, so is the branch)
        batch_tail<U>(a.k - i, [&](auto n) { batch(i, n); });
        if (ok) {
            if (a.vstore && b + VEC <= a.nb) {
                HV o;
#pragma unroll
                for (int v = 0; v < VEC; v++) o.v[v] = acc[v];
                *reinterpret_cast<HV *>(a.Ho + s * a.ldo + b) = o;
            } else {
#pragma unroll
                for (int v = 0; v < VEC; v++)
                    if (b + v < a.nb) a.Ho[s * a.ldo + b + v] = acc[v];
            }
        }
    }
}

template <int DT, typename W, int WPB, int UB = 4>
__device__ inline void level_lds_vec_body(const LevelArgs<DT, W> &a, int64_t j,
                                          typename Elt<DT>::P *lds) {
    if (a.addr32) level_lds_vec_body_a<DT, W, WPB, true, UB>(a, j, lds);
    else level_lds_vec_body_a<DT, W, WPB, false, UB>(a, j, lds);
}

// ---------------------------------------------------------------------------------------
// Parents that share their children.  A state S = {p} u T (p = min S) has the children
// T \ {j}, j in T ascending, which depend on T only; the states {p} u T of one T are consecutive
// in the plan (colex order), with p = p_lo, p_lo + 1, ... from the level's smallest p_lo.  Items
// pair them up -- (p_lo, p_lo + 1), (p_lo + 2, p_lo + 3), ... -- and the 16-byte-load kernel below
// walks items instead of states: an 8-lane slot loads an item's children once and accumulates
// both parents, each in its own plan order, so every state's sum is the same bits as the
// single-parent kernels'.  Parent p + 1's pair for partner j is parent p's plus K - p - 1
// (pair_index(i, j) = (i-1)K - i(i+1)/2 + j).  Across N = 24..32 this removes 26% of the child
// and plan-word loads.
//
// Item flags of one level: bit 0 = state r (0-based in the level) starts an item (p - p_lo even),
// bit 1 = the next state is the second parent of that item (same T: its p is p + 1).
__device__ inline int pair_row_of(int64_t q, int K, int plo) {   // q: 1-based pair index
    int p = plo;
    while (p + 1 < K && (int64_t)p * K - (int64_t)(p + 1) * (p + 2) / 2 + p + 2 <= q) p++;
    return p;
}

template <typename W>
__global__ __launch_bounds__(256) void k_item_flags(const W *__restrict__ trans_lvl, int64_t nst,
                                                    int k, int K, int plo,
                                                    uint8_t *__restrict__ flags) {
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nst) return;
    const int p = pair_row_of((int64_t)(trans_lvl[r * k] & mask), K, plo);
    const bool start = ((p - plo) & 1) == 0;
    bool two = false;
    if (start && r + 1 < nst)
        two = pair_row_of((int64_t)(trans_lvl[(r + 1) * k] & mask), K, plo) == p + 1;
    flags[r] = (uint8_t)((start ? 1 : 0) | (two ? 2 : 0));
}

template <int DT, typename W, int WPB, bool A32, int UB>
__device__ inline void level_lds_vec2_body_a(const LevelArgs<DT, W> &a, int64_t j,
                                             typename Elt<DT>::P *lds) {
    using E = Elt<DT>;
    using PT = typename E::P;
    using HT = typename E::H;
    constexpr int bits = Packing<W>::bits;
    constexpr int VEC = 16 / (int)sizeof(HT);
    constexpr int LPS = 8;                  // lanes per item
    constexpr int TW = LPS * VEC;           // instances per tile
    constexpr int SPW = 64 / LPS;           // items per wave
    struct alignas(16) HV { HT v[VEC]; };
    struct alignas(VEC * sizeof(PT)) PV { PT v[VEC]; };
    const LdsMap &map = a.map;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int slot = lane / LPS, li = lane % LPS;
    const int64_t xcd = j & 7, r = j >> 3;
    int64_t tile, i_begin, i_end, bi, bstride;
    if (map.tiles_xcd > 0) {
        const int64_t ti = r % map.tiles_xcd;
        tile = xcd + 8 * ti;
        bi = r / map.tiles_xcd;
        bstride = map.per_xcd / map.tiles_xcd;
        i_begin = 0;
        i_end = a.nitems;
        if (tile >= map.ntiles) return;
    } else {
        tile = xcd % map.ntiles;
        const int64_t group = xcd / map.ntiles;
        const int64_t ngroups = (8 - tile + map.ntiles - 1) / map.ntiles;
        i_begin = group * a.nitems / ngroups;
        i_end = (group + 1) * a.nitems / ngroups;
        bi = r;
        bstride = map.per_xcd;
    }
    const int64_t b = tile * TW + li * VEC;   // this lane's first instance

    // stage pair rows [qlo, qlo + nq) of this tile: row i at lds[i * TW]
    for (int e = (int)threadIdx.x; e < a.nq * TW; e += WPB * 64) {
        const int64_t bb = tile * TW + (e % TW);
        lds[e] = bb < a.nb ? a.P[(int64_t)(a.qlo + e / TW) * a.ldp + bb] : PT{};
    }
    __syncthreads();

    const char *hbase = reinterpret_cast<const char *>(a.Hc) - a.cfirst * a.ldh * (int64_t)sizeof(HT);
    const uint32_t hrow = (uint32_t)(a.ldh * (int64_t)sizeof(HT));
    const uint32_t hlane = (uint32_t)(b * (int64_t)sizeof(HT));
    const char *lbase = reinterpret_cast<const char *>(lds) +
                        ((int64_t)li * VEC - (int64_t)(1 + a.qlo) * TW) * (int64_t)sizeof(PT);
    auto load_child = [&](W t) -> HV {
        if constexpr (A32) {
            const uint32_t off = __umul24((uint32_t)((uint64_t)t >> bits), hrow) + hlane;
            return *reinterpret_cast<const HV *>(hbase + (uint64_t)off);
        } else {
            return *reinterpret_cast<const HV *>(a.Hc + ((int64_t)(t >> bits) - a.cfirst) * a.ldh + b);
        }
    };
    auto pair_at = [&](uint32_t q) -> PV {
        return *reinterpret_cast<const PV *>(lbase + q * (uint32_t)(TW * sizeof(PT)));
    };

    const int64_t step = bstride * WPB * SPW;
    constexpr int U = UB;
    constexpr int NW = sizeof(W) == 4 ? 4 : 8;   // k <= 31 (Int32 layout) or 63 (Int64)
    for (int64_t i0 = i_begin + (bi * WPB + wave) * SPW; i0 < i_end; i0 += step) {
        const int64_t it = i0 + slot;
        const bool ok = it < i_end;
        const uint32_t item = a.items[ok ? it : i_end - 1];
        const int64_t s = (int64_t)(item & 0x7fffffffu);
        const bool two = ok && (item >> 31) != 0;
        const W *tw = a.trans_lvl + s * (int64_t)a.k;
        W wv[NW];
#pragma unroll
        for (int c = 0; c < NW; c++)
            if (c * LPS < a.k) wv[c] = tw[min(c * LPS + li, a.k - 1)];
        auto word = [&](int t) -> W {   // t uniform
            W w = 0;
#pragma unroll
            for (int c = 0; c < NW; c++)
                if (t / LPS == c) w = (W)__shfl(wv[c], t % LPS, LPS);
            return w;
        };
        // the second parent's pair offset: K - p - 1, p = min S of the first parent
        const int p = pair_row_of((int64_t)__builtin_amdgcn_ubfe((uint32_t)word(0), 0, bits), a.K, a.plo);
        const uint32_t d = (uint32_t)(a.K - p - 1);
        HT acc0[VEC], acc1[VEC];
#pragma unroll
        for (int v = 0; v < VEC; v++) acc0[v] = acc1[v] = E::zero();
        auto batch = [&](int i, auto n) {
            constexpr int N = decltype(n)::value;
            W t[N];
            HV h[N];
#pragma unroll
            for (int u = 0; u < N; u++) t[u] = word(i + u);
#pragma unroll
            for (int u = 0; u < N; u++) h[u] = load_child(t[u]);
#pragma unroll
            for (int u = 0; u < N; u++) {
                const uint32_t q = __builtin_amdgcn_ubfe((uint32_t)t[u], 0, bits);
                const PV p0 = pair_at(q);
#pragma unroll
                for (int v = 0; v < VEC; v++) acc0[v] = E::muladd(acc0[v], p0.v[v], h[u].v[v]);
                if (two) {
                    const PV p1 = pair_at(q + d);
#pragma unroll
                    for (int v = 0; v < VEC; v++) acc1[v] = E::muladd(acc1[v], p1.v[v], h[u].v[v]);
                }
            }
        };
        int i = 0;
        for (; i + U <= a.k; i += U) batch(i, std::integral_constant<int, U>{});
        batch_tail<U>(a.k - i, [&](auto n) { batch(i, n); });
        if (ok) {
            if (a.vstore && b + VEC <= a.nb) {
                HV o0, o1;
#pragma unroll
                for (int v = 0; v < VEC; v++) {
                    o0.v[v] = acc0[v];
                    o1.v[v] = acc1[v];
                }
                *reinterpret_cast<HV *>(a.Ho + s * a.ldo + b) = o0;
                if (two) *reinterpret_cast<HV *>(a.Ho + (s + 1) * a.ldo + b) = o1;
            } else {
#pragma unroll
                for (int v = 0; v < VEC; v++)
                    if (b + v < a.nb) a.Ho[s * a.ldo + b + v] = acc0[v];
                if (two) {
#pragma unroll
                    for (int v = 0; v < VEC; v++)
                        if (b + v < a.nb) a.Ho[(s + 1) * a.ldo + b + v] = acc1[v];
                }
            }
        }
    }
}

template <int DT, typename W, int WPB, int UB = 4>
__device__ inline void level_lds_vec2_body(const LevelArgs<DT, W> &a, int64_t j,
                                           typename Elt<DT>::P *lds) {
    if (a.addr32) level_lds_vec2_body_a<DT, W, WPB, true, UB>(a, j, lds);
    else level_lds_vec2_body_a<DT, W, WPB, false, UB>(a, j, lds);
}

template <int DT, typename W, int WPB, int TAG, int UB = 4>
__global__ __launch_bounds__(WPB * 64) void k_level_fwd_lds_vec2(LevelArgs<DT, W> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    level_lds_vec2_body<DT, W, WPB, UB>(a, (int64_t)blockIdx.x,
                                        reinterpret_cast<typename Elt<DT>::P *>(lds_raw));
}

// UB: transitions per batch of child loads, 4 or 6 (the autotuner picks; 6 pays with passes of
// 64 instances at N=28)
template <int DT, typename W, int WPB, int TAG, int UB = 4>
__global__ __launch_bounds__(WPB * 64) void k_level_fwd_lds_vec(LevelArgs<DT, W> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    level_lds_vec_body<DT, W, WPB, UB>(a, (int64_t)blockIdx.x,
                                       reinterpret_cast<typename Elt<DT>::P *>(lds_raw));
}

template <int DT, typename W, int WPB, int TAG>
__global__ __launch_bounds__(WPB * 64) void k_level_fwd_lds_narrow(LevelArgs<DT, W> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    level_lds_narrow_body<DT, W, WPB>(a, (int64_t)blockIdx.x,
                                      reinterpret_cast<typename Elt<DT>::P *>(lds_raw));
}

template <int DT, typename W, int WPB, int TAG>
__global__ __launch_bounds__(WPB * 64) void k_level_fwd_lds(LevelArgs<DT, W> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    level_lds_body<DT, W, WPB>(a, (int64_t)blockIdx.x,
                               reinterpret_cast<typename Elt<DT>::P *>(lds_raw));
}

// Several groups of instances, each with its own plan (degree), in one launch: group g owns
// workgroups [first[g], first[g+1]) of the grid, each range starting at a multiple of 8.  This is
// how a batch of mixed degrees advances one level of every degree per launch.
constexpr int kMaxGroups = 8;

template <int DT, typename W>
struct MultiLevelArgs {
    int n;
    int64_t first[kMaxGroups + 1];
    LevelArgs<DT, W> g[kMaxGroups];
};

template <int DT, typename W, int WPB, int KIND>
__global__ __launch_bounds__(WPB * 64) void k_level_fwd_lds_multi(MultiLevelArgs<DT, W> m) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int64_t j = blockIdx.x;
    int g = 0;
    while (g + 1 < m.n && j >= m.first[g + 1]) g++;
    auto *lds = reinterpret_cast<typename Elt<DT>::P *>(lds_raw);
    if constexpr (KIND == 2)        // 16-byte loads
        level_lds_vec_body<DT, W, WPB>(m.g[g], j - m.first[g], lds);
    else if constexpr (KIND == 1)   // 16-instance tiles
        level_lds_narrow_body<DT, W, WPB>(m.g[g], j - m.first[g], lds);
    else
        level_lds_body<DT, W, WPB>(m.g[g], j - m.first[g], lds);
}

// ---------------------------------------------------------------------------------------
// Whole walk of one instance in one workgroup, every state value held in LDS.
//
// For small degrees the whole state array fits in LDS (10946 states, 86 KB of FP64 at K = 20),
// so one workgroup can run every level of an instance with a workgroup barrier between levels
// instead of a kernel launch: the plain path then costs a single launch.  Instances are spread
// over workgroups (grid-stride); each state is still summed by one lane in plan order.
constexpr int kFusedMaxLevels = 33;

struct FusedLevels {
    int nlev;                           // levels 2 .. nlev (1-based), level 1 is the empty set
    int32_t first[kFusedMaxLevels + 1]; // 0-based first state of level L at [L-1]
    int32_t nst[kFusedMaxLevels + 1];
    int64_t tstart[kFusedMaxLevels + 1];
    int32_t k[kFusedMaxLevels + 1];
};

template <int DT, typename W>
__global__ __launch_bounds__(1024) void k_fused_lds(
    const W *__restrict__ trans, FusedLevels lv, int nstates, int npairs,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    typename Elt<DT>::H *__restrict__ out, int64_t B) {
    using E = Elt<DT>;
    using HT = typename E::H;
    using PT = typename E::P;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    HT *Hs = reinterpret_cast<HT *>(lds_raw);
    PT *Ps = reinterpret_cast<PT *>(lds_raw + (((size_t)nstates * sizeof(HT) + 15) & ~(size_t)15));
    for (int64_t b = blockIdx.x; b < B; b += gridDim.x) {
        for (int q = threadIdx.x; q < npairs; q += blockDim.x) Ps[q] = P[(int64_t)q * ldp + b];
        if (threadIdx.x == 0) Hs[0] = E::one();
        __syncthreads();
        for (int L = 2; L <= lv.nlev; L++) {
            const int first = lv.first[L - 1], n = lv.nst[L - 1], k = lv.k[L - 1];
            const W *tl = trans + lv.tstart[L - 1];
            for (int r = threadIdx.x; r < n; r += blockDim.x) {
                const W *tw = tl + (int64_t)r * k;
                HT acc = E::zero();
#pragma unroll 4
                for (int i = 0; i < k; i++) {
                    const W t = tw[i];
                    acc = E::muladd(acc, Ps[(int)(t & mask) - 1], Hs[(int64_t)(t >> bits) - 1]);
                }
                Hs[first + r] = acc;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[b] = Hs[nstates - 1];
        __syncthreads();
    }
}

// Levels La .. Lb of each instance in one workgroup, for runs of small levels that would each
// cost a launch of their own.  Level La reads its children from Hin (level La-1, rows ldh apart),
// or, when La = 2, from the empty set's one held in LDS; levels La .. Lb-1 are kept in LDS
// (slot = 0-based plan index - base); level Lb is written to Hout (rows ldo apart).  Instances
// are spread over workgroups (grid-stride) and each state is still summed by one lane in plan
// order, so the values are the level kernels' bits.
template <int DT, typename W>
__global__ __launch_bounds__(1024) void k_fused_range(
    const W *__restrict__ trans, FusedLevels lv, int La, int Lb, int base, int nkeep, int npairs,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Hin, int64_t ldh,
    typename Elt<DT>::H *__restrict__ Hout, int64_t ldo, int64_t B) {
    using E = Elt<DT>;
    using HT = typename E::H;
    using PT = typename E::P;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    HT *Hs = reinterpret_cast<HT *>(lds_raw);
    PT *Ps = reinterpret_cast<PT *>(lds_raw + (((size_t)nkeep * sizeof(HT) + 15) & ~(size_t)15));
    const int64_t in_first = La > 2 ? lv.first[La - 2] : 0;   // 0-based plan index of Hin's row 0
    for (int64_t b = blockIdx.x; b < B; b += gridDim.x) {
        for (int q = threadIdx.x; q < npairs; q += blockDim.x) Ps[q] = P[(int64_t)q * ldp + b];
        if (La == 2 && threadIdx.x == 0) Hs[0] = E::one();
        __syncthreads();
        for (int L = La; L <= Lb; L++) {
            const int first = lv.first[L - 1], n = lv.nst[L - 1], k = lv.k[L - 1];
            const W *tl = trans + lv.tstart[L - 1];
            const bool global_in = L == La && La > 2;
            for (int r = threadIdx.x; r < n; r += blockDim.x) {
                const W *tw = tl + (int64_t)r * k;
                HT acc = E::zero();
                if (global_in) {
                    constexpr int U = 8;   // child loads issued together before they are summed
                    auto batch = [&](int i, auto n) {
                        constexpr int N = decltype(n)::value;
                        W t[N];
                        HT h[N];
#pragma unroll
                        for (int u = 0; u < N; u++) t[u] = tw[i + u];
#pragma unroll
                        for (int u = 0; u < N; u++) h[u] = Hin[((int64_t)(t[u] >> bits) - 1 - in_first) * ldh + b];
#pragma unroll
                        for (int u = 0; u < N; u++) acc = E::muladd(acc, Ps[(int)(t[u] & mask) - 1], h[u]);
                    };
                    int i = 0;
                    for (; i + U <= k; i += U) batch(i, std::integral_constant<int, U>{});
                    batch_tail<U>(k - i, [&](auto n) { batch(i, n); });
                } else {
#pragma unroll 4
                    for (int i = 0; i < k; i++) {
                        const W t = tw[i];
                        acc = E::muladd(acc, Ps[(int)(t & mask) - 1], Hs[(int64_t)(t >> bits) - 1 - base]);
                    }
                }
                if (L == Lb) Hout[(int64_t)r * ldo + b] = acc;
                else Hs[first + r - base] = acc;
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------------
// Pair values gathered on the device for the boson-sampling batch shape: one matrix A, many
// index lists (src/dp.jl:479-494).  P[q][b] = A[idx_b[i], idx_b[j]] for the q-th pair i < j in
// the row-major order of _pair_index (src/unrolled.jl:60).  A is column-major (Julia's layout):
// element (r, c) at c*N + r, 0-based.  One thread per (pair, instance), instance fastest, so the
// writes are coalesced.
template <typename PT>
__global__ __launch_bounds__(256) void k_gather_pairs(const PT *__restrict__ A, int64_t N,
                                                      const int32_t *__restrict__ idx, int K,
                                                      int64_t B, PT *__restrict__ P, int64_t ldp) {
    const int64_t npairs = (int64_t)K * (K - 1) / 2;
    const int64_t total = npairs * B;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int64_t q = t / B, b = t - q * B;
        // row i (0-based) of the upper triangle holds pairs [i*K - i(i+1)/2, ...) of length K-1-i
        int i = 0;
        int64_t base = 0;
        while (base + (K - 1 - i) <= q) {
            base += K - 1 - i;
            i++;
        }
        const int j = i + 1 + (int)(q - base);
        const int64_t r = idx[b * K + i] - 1, c = idx[b * K + j] - 1;
        P[q * ldp + b] = A[c * N + r];
    }
}

// Range check of the exact (128-bit) walk, one thread per instance of P ([pair][ldp], int64):
// out[b] = log2 of min((K-1)!! * m^(K/2), sqrt(prod_i max(r_i, 1))), m = max(1, max |p|) and
// r_i = sum_{j != i} |p_ij|, an upper bound on |haf| and on every (K-2)-minor of instance b
// (for entrywise absolute values haf^2 <= perm <= the product of the row sums).
static __global__ __launch_bounds__(256) void k_exact_log2_bound(const long long *__restrict__ P,
                                                          int64_t ldp, int64_t B, int K,
                                                          double *__restrict__ out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double m = 1.0, lr = 0.0;
    for (int i = 0; i < K; i++) {
        double ri = 0.0;
        // pairs (j, i), j < i: q = j*K - j(j+1)/2 + (i - j - 1); pairs (i, j), j > i
        for (int j = 0; j < i; j++) {
            const int64_t q = (int64_t)j * K - (int64_t)j * (j + 1) / 2 + (i - j - 1);
            ri += fabs((double)P[q * ldp + b]);
        }
        const int64_t q0 = (int64_t)i * K - (int64_t)i * (i + 1) / 2;
        for (int j = i + 1; j < K; j++) {
            const double a = fabs((double)P[(q0 + (j - i - 1)) * ldp + b]);
            ri += a;
            m = fmax(m, a);
        }
        lr += log2(fmax(ri, 1.0));
    }
    double lf = 0.5 * K * log2(m);
    for (int k = K - 1; k > 1; k -= 2) lf += log2((double)k);
    out[b] = fmin(lf, 0.5 * lr);
}

// Measurement helper (no counterpart in the reference): the rate at which the device serves
// random reads of whole 128-byte lines, each read the way the 16-byte-load level kernels read a
// child row (8 lanes x 16 bytes), from a working set of `nlines` lines.  Every line index is
// reduced modulo nlines, so every read is in bounds.  The practical ceiling of the level
// kernels' gathers, next to the device-copy rate.
static __global__ __launch_bounds__(512) void k_gather_probe(const double2 *__restrict__ buf,
                                                             uint64_t nlines, int iters, uint64_t seed,
                                                             double *__restrict__ sink) {
    const int li = (int)(threadIdx.x & 7);
    uint64_t x = seed + (uint64_t)(blockIdx.x * 64 + (threadIdx.x >> 3)) * 0x9E3779B97F4A7C15ull;
    double acc = 0.0;
    for (int it = 0; it < iters; it++) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            x ^= x >> 31;
            x *= 0xBF58476D1CE4E5B9ull;
            x ^= x >> 29;
            v[u] = buf[(x % nlines) * 8 + li];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u].x + v[u].y;
    }
    if (acc == 1.25e300) sink[0] = acc;   // never true for the zeroed buffer; keeps the loads
}

// Row of ones for the empty set (H[1] = one, src/dp.jl:421): a plain store kernel, so the
// sequence stays valid inside a captured graph.
template <int DT>
__global__ void k_fill_one(typename Elt<DT>::H *__restrict__ H, int64_t nb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) H[i] = Elt<DT>::one();
}

}  // namespace eggman

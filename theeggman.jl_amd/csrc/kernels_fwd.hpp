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
#pragma unroll 4
        for (int i = 0; i < k; i++) {
            const W t = tw[i];
            const int64_t q = (int64_t)(t & mask) - 1;
            const int64_t c = (int64_t)(t >> bits) - cfirst;
            acc = E::muladd(acc, P[q * ldp + b], Hc[c * ldh + b]);
        }
        Ho[s * ldo + b] = acc;
    }
}

// Row of ones for the empty set (H[1] = one, src/dp.jl:421): a plain store kernel, so the
// sequence stays valid inside a captured graph.
template <int DT>
__global__ void k_fill_one(typename Elt<DT>::H *__restrict__ H, int64_t nb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) H[i] = Elt<DT>::one();
}

}  // namespace eggman

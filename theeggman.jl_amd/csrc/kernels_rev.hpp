// Reverse sweep of the subset DP: d haf / d P for every pair, for a batch of instances.
//
// The forward walk is H[s] = sum_k P[q_k] H[c_k] for the states s of level L and their children
// c_k in level L-1 (src/dp.jl:354-361), ending at haf = H[full set].  Its adjoint, with
// G[s] = d haf / d H[s] and G[full set] = 1, is
//
//     G[c]   = sum over the parent transitions (s, q) of c:   P[q] G[s]           (gather)
//     dP[q]  = sum over every transition (s, c) with pair q:  H[c] G[s]
//
// Both are written as gathers over index structures built once per degree (plan_host.hpp:
// RevPlanHost), so every sum has a fixed order and the result is deterministic: the child-major
// CSR lists each child's parents, and the pair-major segments list each pair's transitions.  A
// pair's segment sums are added in ascending level order by a last small kernel.  No atomics.
//
// Layouts match the forward: [state][batch] and [pair][batch], batch fastest, 64 instances per
// wave.  H keeps every forward level (the sweep reads level L-1 while handling level L); G only
// keeps two levels.
#pragma once
#include "kernels_fwd.hpp"

#pragma clang fp contract(off)

namespace eggman {

struct RevSegDev {
    int64_t start;
    int32_t len;
    int32_t level;
};

// G of level L-1 from G of level L.  One wave per (child, 64-instance tile); the pair rows of
// level L are staged in LDS as in the forward LDS kernel (same qlo/nq), and the workgroup-to-XCD
// mapping is the forward one.
template <int DT, typename W, int WPB, bool USE_LDS>
__global__ __launch_bounds__(WPB * 64) void k_rev_g(
    const W *__restrict__ cent, const int64_t *__restrict__ coff, int64_t nch, int qlo, int nq,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ Gc, int64_t nb, LdsMap map) {
    using E = Elt<DT>;
    using PT = typename E::P;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    PT *lds = reinterpret_cast<PT *>(lds_raw);

    const int lane = (int)(threadIdx.x & 63);
    const int wave = (int)(threadIdx.x >> 6);
    const int64_t j = blockIdx.x;
    const int64_t xcd = j & 7, r = j >> 3;
    int64_t tile, s_begin, s_end, bi, bstride;
    if (map.tiles_xcd > 0) {
        const int64_t ti = r % map.tiles_xcd;
        tile = xcd + 8 * ti;
        bi = r / map.tiles_xcd;
        bstride = map.per_xcd / map.tiles_xcd;
        s_begin = 0;
        s_end = nch;
        if (tile >= map.ntiles) return;
    } else {
        tile = xcd % map.ntiles;
        const int64_t group = xcd / map.ntiles;
        const int64_t ngroups = (8 - tile + map.ntiles - 1) / map.ntiles;
        s_begin = group * nch / ngroups;
        s_end = (group + 1) * nch / ngroups;
        bi = r;
        bstride = map.per_xcd;
    }
    const int64_t b = tile * 64 + lane;
    const bool live = b < nb;
    if (USE_LDS) {
        for (int i = wave; i < nq; i += WPB)
            lds[i * 64 + lane] = live ? P[(int64_t)(qlo + i) * ldp + b] : PT{};
        __syncthreads();
    }

    const int64_t step = bstride * WPB;
    for (int64_t c0 = s_begin + bi * WPB; c0 < s_end; c0 += step) {
        const int64_t c = wave_uniform(c0 + wave);
        if (c >= s_end) break;
        const int64_t e0 = coff[c], e1 = coff[c + 1];
        typename E::H acc = E::zero();
        if (live) {
            constexpr int U = 8;   // gathers issued together before they are summed in order
            int64_t e = e0;
            for (; e + U <= e1; e += U) {
                W t[U];
                typename E::H g[U];
#pragma unroll
                for (int u = 0; u < U; u++) t[u] = cent[e + u];
#pragma unroll
                for (int u = 0; u < U; u++) g[u] = Gs[(int64_t)(t[u] >> bits) * ldg + b];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int q = (int)(t[u] & mask) - 1 - qlo;
                    const PT pv = USE_LDS ? lds[q * 64 + lane] : P[(int64_t)(qlo + q) * ldp + b];
                    acc = E::muladd(acc, pv, g[u]);
                }
            }
            for (; e < e1; e++) {
                const W t = cent[e];
                const int q = (int)(t & mask) - 1 - qlo;
                const int64_t s = (int64_t)(t >> bits);
                const PT pv = USE_LDS ? lds[q * 64 + lane] : P[(int64_t)(qlo + q) * ldp + b];
                acc = E::muladd(acc, pv, Gs[s * ldg + b]);
            }
            Gc[c * nb + b] = acc;
        }
    }
}

// Partial dP sums of one level: one wave per (segment, tile), sum over the segment's
// transitions (parent s, child c) of H[c] * G[s].
template <int DT>
__global__ __launch_bounds__(256) void k_rev_dp_seg(
    const uint64_t *__restrict__ pent, const RevSegDev *__restrict__ segs, int64_t seg0,
    int64_t nseg, const typename Elt<DT>::H *__restrict__ Hc, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ part, int64_t nb, int64_t ntiles) {
    using E = Elt<DT>;
    const int lane = (int)(threadIdx.x & 63);
    const int64_t ntask = nseg * ntiles;
    const int64_t gstride = (int64_t)gridDim.x * 4;
    for (int64_t g = ((int64_t)blockIdx.x * 256 + threadIdx.x) / 64; g < ntask; g += gstride) {
        const int64_t gg = wave_uniform(g);
        const int64_t si = gg / ntiles;
        const int64_t tile = gg - si * ntiles;
        const int64_t b = tile * 64 + lane;
        if (b >= nb) continue;
        const RevSegDev sg = segs[seg0 + si];
        const uint64_t *pe = pent + sg.start;
        typename E::H acc = E::zero();
        constexpr int U = 8;   // gathers issued together before they are summed in order
        int i = 0;
        for (; i + U <= sg.len; i += U) {
            typename E::H hv[U], gv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t w = pe[i + u];
                hv[u] = Hc[(int64_t)(w & 0xffffffffull) * ldh + b];
                gv[u] = Gs[(int64_t)(w >> 32) * ldg + b];
            }
#pragma unroll
            for (int u = 0; u < U; u++) acc = E::muladd_hh(acc, hv[u], gv[u]);
        }
        for (; i < sg.len; i++) {
            const uint64_t w = pe[i];
            const int64_t s = (int64_t)(w >> 32), c = (int64_t)(w & 0xffffffffull);
            acc = E::muladd_hh(acc, Hc[c * ldh + b], Gs[s * ldg + b]);
        }
        part[(seg0 + si) * nb + b] = acc;
    }
}

// dP[q][b] = sum of the pair's segment sums in ascending level order.
template <int DT>
__global__ __launch_bounds__(256) void k_rev_dp_reduce(
    const int64_t *__restrict__ qoff, const int32_t *__restrict__ qseg,
    const typename Elt<DT>::H *__restrict__ part, typename Elt<DT>::H *__restrict__ dP,
    int64_t lddp, int64_t npairs, int64_t nb) {
    using E = Elt<DT>;
    const int64_t total = npairs * nb;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t q = i / nb, b = i - q * nb;
        typename E::H acc = E::zero();
        for (int64_t k = qoff[q]; k < qoff[q + 1]; k++) acc = E::add_h(acc, part[(int64_t)qseg[k] * nb + b]);
        dP[q * lddp + b] = acc;
    }
}

}  // namespace eggman

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
#include "plan_host.hpp"

#pragma clang fp contract(off)

namespace eggman {

struct RevSegDev {
    int64_t start;
    int32_t len;
    int32_t level;
};

// G of level L-1 from G of level L.  One wave per (child, 64-instance tile); the pair rows of
// level L are staged in LDS as in the forward LDS kernel (same qlo/nq), and the workgroup-to-XCD
// mapping is the forward one.  Both G levels have rows ldg apart.
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
            // N entries from e: their N gathers issued together, then summed in list order
            auto batch = [&](int64_t e, auto n) {
                constexpr int N = decltype(n)::value;
                W t[N];
                typename E::H g[N];
#pragma unroll
                for (int u = 0; u < N; u++) t[u] = cent[e + u];
#pragma unroll
                for (int u = 0; u < N; u++) g[u] = Gs[(int64_t)(t[u] >> bits) * ldg + b];
#pragma unroll
                for (int u = 0; u < N; u++) {
                    const int q = (int)(t[u] & mask) - 1 - qlo;
                    const PT pv = USE_LDS ? lds[q * 64 + lane] : P[(int64_t)(qlo + q) * ldp + b];
                    acc = E::muladd(acc, pv, g[u]);
                }
            };
            int64_t e = e0;
            for (; e + U <= e1; e += U) batch(e, std::integral_constant<int, U>{});
            // the last entries (list length mod U) as one batch too
            batch_tail<U>((int)(e1 - e), [&](auto n) { batch(e, n); });
            Gc[c * ldg + b] = acc;
        }
    }
}

// The adjoint gather for small batches: BT lanes (BT < 64, a power of two) per (child, tile of BT
// instances), P read from global memory.  Each lane sums its instance over the child's parent
// list in list order, as k_rev_g does, so both give the same bits.
template <int DT, typename W, int BT>
__device__ inline void rev_g_bt_body(
    int64_t bid, int64_t nbid, const W *__restrict__ cent, const int64_t *__restrict__ coff,
    int64_t nch, const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ Gc, int64_t nb, int64_t ntiles) {
    using E = Elt<DT>;
    constexpr int bits = Packing<W>::bits;
    constexpr W mask = (W)((W(1) << bits) - 1);
    const int lane = (int)(threadIdx.x % BT);
    const int64_t ngroups = nch * ntiles;
    const int64_t gstride = nbid * (kBlock / BT);
    for (int64_t g = (bid * kBlock + threadIdx.x) / BT; g < ngroups; g += gstride) {
        const int64_t c = ntiles == 1 ? g : g / ntiles;
        const int64_t b = (g - c * ntiles) * BT + lane;
        if (b >= nb) continue;
        const int64_t e0 = coff[c], n = coff[c + 1] - e0;
        typename E::H acc = E::zero();
        // eight entries a batch, summed in list order; the next batch's entry words load while
        // this batch's values are in flight (past the end, repeats of the last entry, not added)
        constexpr int U = 8;
        if (n > 0) {
            W t[U];
#pragma unroll
            for (int u = 0; u < U; u++) t[u] = cent[e0 + min((int64_t)u, n - 1)];
            for (int64_t b0 = 0; b0 < n; b0 += U) {
                typename E::P pv[U];
                typename E::H gv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    pv[u] = P[((int64_t)(t[u] & mask) - 1) * ldp + b];
                    gv[u] = Gs[(int64_t)(t[u] >> bits) * ldg + b];
                }
#pragma unroll
                for (int u = 0; u < U; u++) t[u] = cent[e0 + min(b0 + U + u, n - 1)];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (b0 + u < n) acc = E::muladd(acc, pv[u], gv[u]);
            }
        }
        Gc[c * ldg + b] = acc;
    }
}

template <int DT, typename W, int BT>
__global__ __launch_bounds__(kBlock) void k_rev_g_bt(
    const W *__restrict__ cent, const int64_t *__restrict__ coff, int64_t nch,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ Gc, int64_t nb, int64_t ntiles) {
    rev_g_bt_body<DT, W, BT>(blockIdx.x, gridDim.x, cent, coff, nch, P, ldp, Gs, ldg, Gc, nb, ntiles);
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
        auto batch = [&](int i, auto n) {
            constexpr int N = decltype(n)::value;
            typename E::H hv[N], gv[N];
#pragma unroll
            for (int u = 0; u < N; u++) {
                const uint64_t w = pe[i + u];
                hv[u] = Hc[(int64_t)(w & 0xffffffffull) * ldh + b];
                gv[u] = Gs[(int64_t)(w >> 32) * ldg + b];
            }
#pragma unroll
            for (int u = 0; u < N; u++) acc = E::muladd_hh(acc, hv[u], gv[u]);
        };
        int i = 0;
        for (; i + U <= sg.len; i += U) batch(i, std::integral_constant<int, U>{});
        batch_tail<U>(sg.len - i, [&](auto n) { batch(i, n); });
        part[(seg0 + si) * nb + b] = acc;
    }
}

// ---------------------------------------------------------------------------------------
// Parent-major dP for one level.  Every transition of a parent S leaves from p = min(S), so all
// of S's contributions H[child] * G[S] land in row p of dP, at column j - p - 1 for the partner
// j.  One workgroup takes a task (parents of the level with the same p) for one 64-instance
// tile; wave w accumulates the row over its fixed slice of the task's parents in its own LDS
// rows -- one G read per parent and one child read per transition, half the gathers of the
// pair-major segments -- and the waves' rows are then summed in wave order into the task's
// partial rows.  A parent's transitions have distinct partners, so a batch of them reads all its
// LDS slots before writing any: the read-modify-writes of a batch overlap instead of chaining.
// Tasks of a row are summed in a fixed order by k_rev_dp_reduce_runs / k_rev_dp_reduce_pairs.
struct ParTaskDev {
    int64_t start;
    int64_t row0;
    int32_t len;
    int32_t p;
    int32_t level;
    int32_t pad;
};

struct RedRowDev {
    int32_t tb, te, u, pad;
};

// Plan words are int32 here (the parent-major path runs on the Int32 layout).  U is the batch of
// transitions gathered together: parents of out-degree k <= U go two at a time (2U gathers in
// flight); larger out-degrees take ceil(k / U) batches per parent.  Plan words and parent ranks are read with
// vector loads, one word per lane, and broadcast with readlane, so the only LDS-counter traffic
// in the loop is the accumulator's own.
template <int DT, int WPB, int U>
__global__ __launch_bounds__(WPB * 64) void k_rev_dp_par(
    const int32_t *__restrict__ trans_lvl, int k, int K,
    const int32_t *__restrict__ prank, const ParTaskDev *__restrict__ tasks, int64_t task0,
    int64_t ntask, const typename Elt<DT>::H *__restrict__ Hc, int64_t cfirst, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ part, int64_t nb, int64_t ntiles, int rows_per_wave,
    int sub_len) {
    using E = Elt<DT>;
    using HT = typename E::H;
    constexpr int bits = Packing<int32_t>::bits;
    constexpr int32_t mask = (1 << bits) - 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    HT *rows = reinterpret_cast<HT *>(lds_raw);
    const int lane = (int)(threadIdx.x & 63);
    // uniform across the wave
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t wslice = (int64_t)rows_per_wave * 64;
    HT *acc = rows + wave * wslice;
    const int64_t nwork = ntask * ntiles;
    const int klast = k - 1;
    for (int64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int64_t ti = w / ntiles, tile = w - ti * ntiles;
        const ParTaskDev tk = tasks[task0 + ti];
        const int64_t b = tile * 64 + lane;
        // lanes past the batch shadow the last instance (valid addresses); their sums are dropped
        const int64_t bl = b < nb ? b : nb - 1;
        const int nrow = K - tk.p;
        // 0-based index of pair (p, p+1): (p-1)K - p(p+1)/2 + p
        const int qbase = (tk.p - 1) * K - tk.p * (tk.p + 1) / 2 + tk.p;
        for (int u = 0; u < nrow; u++) acc[u * 64 + lane] = E::zero();
        // this wave's parents: [wave * sub_len, (wave + 1) * sub_len) of the task
        const int i0 = wave * sub_len;
        const int cnt = min(tk.len - i0, sub_len);
        if (cnt > 0) {
            const int rk = prank[tk.start + i0 + min(lane, cnt - 1)];   // lane j: rank of parent j
            auto words = [&](int j) {   // lane v: word v of parent j (clamped to the last)
                const int r = __builtin_amdgcn_readlane(rk, j);
                return trans_lvl[(int64_t)r * k + min(lane, klast)];
            };
            auto gval = [&](int j) {
                return Gs[(int64_t)__builtin_amdgcn_readlane(rk, j) * ldg + bl];
            };
            auto gather = [&](int wv, int t, HT *h) {
#pragma unroll
                for (int v = 0; v < U; v++) {
                    const int wd = __builtin_amdgcn_readlane(wv, min(t + v, klast));
                    h[v] = Hc[((int64_t)(wd >> bits) - cfirst) * ldh + bl];
                }
            };
            // `on` false stores back what was read (a no-op written without a branch, so the
            // loads feeding it cannot be sunk below earlier updates)
            auto update = [&](int wv, int t, const HT *h, HT g, bool on) {
                // a short batch repeats its last transition: the repeats read the same slot
                // before any write and store the same value, so they change nothing
                HT a[U];
                int slot[U];
#pragma unroll
                for (int v = 0; v < U; v++) {
                    const int wd = __builtin_amdgcn_readlane(wv, min(t + v, klast));
                    slot[v] = ((wd & mask) - 1 - qbase) * 64 + lane;
                    a[v] = acc[slot[v]];
                }
#pragma unroll
                for (int v = 0; v < U; v++) {
                    const HT r = E::muladd_hh(a[v], h[v], g);
                    acc[slot[v]] = on ? r : a[v];
                }
            };
            if (k <= U) {
                // two parents per step: both parents' gathers are issued together, then the
                // first parent's updates, then the second's (so slots shared by the two parents
                // are summed in parent order, as one parent at a time would)
                for (int j = 0; j < cnt; j += 2) {
                    const bool two = j + 1 < cnt;
                    const int wa = words(j), wb = words(two ? j + 1 : j);
                    const HT ga = gval(j), gb = gval(two ? j + 1 : j);
                    HT ha[U], hb[U];
                    gather(wa, 0, ha);
                    gather(wb, 0, hb);
                    // every gather of the step is issued before the first LDS update
                    __builtin_amdgcn_sched_barrier(0);
                    update(wa, 0, ha, ga, true);
                    update(wb, 0, hb, gb, two);
                }
            } else {
                // out-degree above 64 does not occur (degree <= 64)
                for (int j = 0; j < cnt; j++) {
                    const int wv = words(j);
                    const HT g = gval(j);
                    for (int t = 0; t < k; t += U) {
                        HT h[U];
                        gather(wv, t, h);
                        update(wv, t, h, g, true);
                    }
                }
            }
        }
        __syncthreads();
        // the task's row: the waves' rows summed in wave order, spread over the workgroup
        for (int e = (int)threadIdx.x; e < nrow * 64; e += WPB * 64) {
            const int64_t bb = tile * 64 + (e & 63);
            if (bb < nb) {
                HT sum = rows[e];
#pragma unroll
                for (int v = 1; v < WPB; v++) sum = E::add_h(sum, rows[v * wslice + e]);
                part[(tk.row0 + (e >> 6)) * nb + bb] = sum;
            }
        }
        __syncthreads();
    }
}

// Parent-major dP for one instance: the sums of k_rev_dp_par in the same order, with lanes
// over a parent's transitions instead of over instances.  Slice w of a task (parents
// [w * sub_len, (w + 1) * sub_len), as wave w of k_rev_dp_par takes them) runs on 32 lanes, lane t
// holding transition t of the current parent (k <= 31 on the Int32 layout); a parent's
// transitions have distinct partners, so its lanes update distinct slots of the slice's LDS row,
// parent after parent.  The slices' rows are then added in slice order.  Every slot therefore
// sums the same terms in the same order as the 64-wide kernel, without 63 idle lanes per wave
// and with a few hundred bytes of LDS per task, so many tasks share a CU.
template <int DT, int WPB>
__global__ __launch_bounds__(WPB * 32) void k_rev_dp_par_one(
    const int32_t *__restrict__ trans_lvl, int k, int K, const int32_t *__restrict__ prank,
    const ParTaskDev *__restrict__ tasks, int64_t task0, int64_t ntask,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t cfirst, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg, typename Elt<DT>::H *__restrict__ part,
    int64_t nb, int rows_per_slice, int sub_len) {
    using E = Elt<DT>;
    using HT = typename E::H;
    constexpr int bits = Packing<int32_t>::bits;
    constexpr int32_t mask = (1 << bits) - 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    HT *rows = reinterpret_cast<HT *>(lds_raw);
    const int t = (int)(threadIdx.x & 31);
    const int slice = (int)(threadIdx.x >> 5);
    HT *acc = rows + slice * rows_per_slice;
    for (int64_t ti = blockIdx.x; ti < ntask; ti += gridDim.x) {
        const ParTaskDev tk = tasks[task0 + ti];
        const int nrow = K - tk.p;
        const int qbase = (tk.p - 1) * K - tk.p * (tk.p + 1) / 2 + tk.p;
        for (int u = t; u < nrow; u += 32) acc[u] = E::zero();
        const int i0 = slice * sub_len;
        const int cnt = min(tk.len - i0, sub_len);
        for (int j = 0; j < cnt; j++) {
            const int r = prank[tk.start + i0 + j];
            const HT g = Gs[(int64_t)r * ldg];
            if (t < k) {
                const int32_t wd = trans_lvl[(int64_t)r * k + t];
                const HT h = Hc[((int64_t)(wd >> bits) - cfirst) * ldh];
                const int slot = (wd & mask) - 1 - qbase;
                acc[slot] = E::muladd_hh(acc[slot], h, g);
            }
        }
        __syncthreads();
        // the task's row: the slices' rows summed in slice order
        for (int u = (int)threadIdx.x; u < nrow; u += WPB * 32) {
            HT sum = rows[u];
#pragma unroll
            for (int v = 1; v < WPB; v++) sum = E::add_h(sum, rows[v * rows_per_slice + u]);
            part[(tk.row0 + u) * nb] = sum;
        }
        __syncthreads();
    }
}

// Sum of `n` values src[idx(i)] in ascending i, loads issued eight at a time (the tail as one
// batch).
template <int DT, typename F>
__device__ __forceinline__ typename Elt<DT>::H ordered_sum(int64_t n, F &&load) {
    using E = Elt<DT>;
    typename E::H acc = E::zero();
    int64_t i = 0;
    auto batch = [&](int64_t i, auto m) {
        constexpr int N = decltype(m)::value;
        typename E::H v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = load(i + j);
#pragma unroll
        for (int j = 0; j < N; j++) acc = E::add_h(acc, v[j]);
    };
    for (; i + 8 <= n; i += 8) batch(i, std::integral_constant<int, 8>{});
    batch_tail<8>((int)(n - i), [&](auto m) { batch(i, m); });
    return acc;
}

// Stage one: red[r][b] = sum of column u of the partial rows of tasks rtask[tb .. te), in order.
template <int DT>
__global__ __launch_bounds__(256) void k_rev_dp_reduce_runs(
    const RedRowDev *__restrict__ red, int64_t nred, const int32_t *__restrict__ rtask,
    const ParTaskDev *__restrict__ tasks, const typename Elt<DT>::H *__restrict__ part,
    typename Elt<DT>::H *__restrict__ out, int64_t nb) {
    const int64_t total = nred * nb;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t r = i / nb, b = i - r * nb;
        const RedRowDev d = red[r];
        out[r * nb + b] = ordered_sum<DT>(d.te - d.tb, [&](int64_t t) {
            return part[(tasks[rtask[d.tb + t]].row0 + d.u) * nb + b];
        });
    }
}

// Stage two: dP[q][b] = sum over the runs of row p(q), in order, of their stage-one rows.
template <int DT>
__global__ __launch_bounds__(256) void k_rev_dp_reduce_pairs(
    const int32_t *__restrict__ qrow, const int32_t *__restrict__ qcol,
    const int32_t *__restrict__ srow0, const int32_t *__restrict__ sruns, int K,
    const typename Elt<DT>::H *__restrict__ red_out, typename Elt<DT>::H *__restrict__ dP,
    int64_t lddp, int64_t npairs, int64_t nb) {
    const int64_t total = npairs * nb;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t q = i / nb, b = i - q * nb;
        const int p = qrow[q], u = qcol[q];
        const int64_t r0 = srow0[p] + u, step = K - p;
        dP[q * lddp + b] = ordered_sum<DT>(sruns[p], [&](int64_t g) {
            return red_out[(r0 + g * step) * nb + b];
        });
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

// ---------------------------------------------------------------------------------------
// Fused child-major reverse step of one level (Int32 layout, elements of at most 8 bytes): G of
// the children and the level's dP partial rows in one pass that gathers each parent value once
// per transition, where the adjoint gather and the parent-major dP gather it twice.  The sums
// follow the order stated at FusePlanHost (plan_host.hpp); the small-batch kernels below follow
// it too, so every pass size gives the same bits.
//
// A state value of another lane of the wave.
template <int DT>
__device__ inline typename Elt<DT>::H shfl_h(typename Elt<DT>::H x, int src) {
    using HT = typename Elt<DT>::H;
    if constexpr (std::is_same<HT, float>::value || std::is_same<HT, double>::value) {
        return __shfl(x, src);
    } else if constexpr (std::is_same<HT, float2>::value) {
        return make_float2(__shfl(x.x, src), __shfl(x.y, src));
    } else if constexpr (std::is_same<HT, double2>::value) {
        return make_double2(__shfl(x.x, src), __shfl(x.y, src));
    } else {
        return make_ulonglong2((unsigned long long)__shfl((long long)x.x, src),
                               (unsigned long long)__shfl((long long)x.y, src));
    }
}

// f(n) with n an integral_constant equal to `n` (1 <= n <= N).
template <int N, class F>
__device__ inline void dispatch_n(int n, F &&f) {
    if constexpr (N >= 1) {
        if (n == N) f(std::integral_constant<int, N>{});
        else dispatch_n<N - 1>(n, f);
    }
}

// One workgroup of kFuseWaves waves per (stripe s, 64-instance tile); wave w is wave-stripe
// v = s * kFuseWaves + w and walks its blocks k = v, v + V, ... in every phase p (the waves of all
// stripes of a tile sweep the child level together, so neighbouring children's shared parents are
// read while still in L2), so each child is always handled by the same lanes: G[c] is carried from
// one phase to the next through Gc.  Per phase: the pair rows (p, p+1 .. K) are staged in LDS; a
// wave's children of the phase form one stream, taken STEP at a time (at most 16 gathers a step);
// a step's entry words are loaded two steps ahead and its values one step ahead, every step issuing
// the same loads, so a wave keeps a step of gathers in flight while it updates the previous one;
// each child's G term P * G[parent] goes into G[c] and its dP term H[c] * G[parent] into slot
// j - p - 1 of the wave's LDS total row; at the phase's end the waves' rows are summed in wave order
// into the stripe's partial row.  A child's parents have distinct partners, so its update reads all
// its slots before writing any.
//
// A32: every row offset of Hc, Gs and Gc (row * row bytes + column bytes) fits in 32 bits, so a
// row access is a uniform byte base plus one 32-bit per-lane offset (the saddr form of the loads)
// instead of a 64-bit per-lane address per gather.
template <int DT, bool A32>
__global__ __launch_bounds__(kFuseWaves * 64) void k_rev_fused(
    const int32_t *__restrict__ fent, const int32_t *__restrict__ fco,
    const uint32_t *__restrict__ cmask, int64_t nch, int K, int csize, int plo, int phi, int S,
    const int64_t *__restrict__ prow0, const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, typename Elt<DT>::H *__restrict__ Gc, int64_t ldg,
    typename Elt<DT>::H *__restrict__ part, int64_t nb, int64_t ntiles) {
    using E = Elt<DT>;
    using HT = typename E::H;
    using PT = typename E::P;
    constexpr int bits = Packing<int32_t>::bits;
    constexpr int32_t mask = (1 << bits) - 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int nslot = K - plo;
    PT *prow = reinterpret_cast<PT *>(lds_raw);
    HT *rows = reinterpret_cast<HT *>(lds_raw + (size_t)nslot * 64 * sizeof(PT));
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // workgroup j runs on XCD j mod 8: the stripes of a tile share one XCD (and its L2) when there
    // are 8 tiles or more, else each tile's stripes are dealt to the XCDs that hold the tile
    const int64_t j = blockIdx.x, xcd = j & 7, r = j >> 3;
    int64_t tile, s;
    if (ntiles >= 8) {
        tile = xcd + 8 * (r / S);
        s = r % S;
        if (tile >= ntiles) return;
    } else {
        tile = xcd % ntiles;
        const int64_t g = xcd / ntiles, ng = (8 - tile + ntiles - 1) / ntiles;
        s = r * ng + g;
        if (s >= S) return;
    }
    HT *tr = rows + (size_t)wave * nslot * 64;   // this wave's total row
    const int64_t b = tile * 64 + lane;
    const bool live = b < nb;
    // lanes past the batch shadow the last instance
    const uint32_t col = (uint32_t)(live ? b : nb - 1);
    const uint32_t cbytes = col * (uint32_t)sizeof(HT);
    const uint32_t hrow = (uint32_t)(ldh * (int64_t)sizeof(HT)), grow = (uint32_t)(ldg * (int64_t)sizeof(HT));
    // element (row, col) of a level: base + row * row bytes + column bytes
    auto row_at = [&](const HT *base, int64_t row, int64_t ld, uint32_t rbytes) -> const HT * {
        if constexpr (A32)
            return reinterpret_cast<const HT *>(reinterpret_cast<const char *>(base) +
                                                (uint64_t)((uint32_t)row * rbytes + cbytes));
        else
            return base + row * ld + col;
    };
    const int64_t V = (int64_t)S * kFuseWaves, v = s * kFuseWaves + wave;
    const int64_t nblk = (nch + kFuseT - 1) / kFuseT;
    for (int u = 0; u < nslot; u++) tr[u * 64 + lane] = E::zero();
    int off = 0;   // start of the phase's run inside every child's sorted list
    for (int p = plo; p <= phi; p++) {
        const int nrow = K - p, np = nrow - csize;
        const int qbase = (p - 1) * K - p * (p + 1) / 2 + p;   // 0-based index of pair (p, p + 1)
        const uint32_t below = (1u << p) - 1u;                  // elements 1 .. p
        __syncthreads();
        for (int u = wave; u < nrow; u += kFuseWaves) prow[u * 64 + lane] = (P + (int64_t)(qbase + u) * ldp)[col];
        __syncthreads();
        dispatch_n<16>(np, [&](auto npc) {
            constexpr int NP = decltype(npc)::value;
            constexpr int STEP = NP <= 4 ? 4 : NP <= 8 ? 2 : 1;    // children per step (<= 16 gathers)
            struct Buf {
                HT h[STEP], g0[STEP], g[STEP][NP];
            };
            struct Step {
                int cv;          // lane q: rank of child q (level-relative)
                int cov;         // lane q: start of child q's sorted list
                uint32_t on;     // uniform, bit q: child q is real (else it repeats a real child)
            };
            // The wave's children of the phase as one stream: its blocks in groups of 8 (lane l of
            // a group is child l & 7 of block gk + (l >> 3) V), the next group's masks and offsets
            // loaded one group ahead; a group's children in the phase are the set bits of a ballot,
            // in (block, rank) order.
            auto load_group = [&](int64_t g, uint32_t &cmv, int32_t &cov) {
                const int64_t kl = g + (int64_t)(lane >> 3) * V;
                const int64_t cl = kl * kFuseT + (lane & 7);
                const bool in = kl < nblk && cl < nch;
                cmv = in ? cmask[cl] : 0xffffffffu;
                cov = in ? fco[cl] : 0;
            };
            int64_t gk = v, gkn = v + 8 * V;
            uint32_t cmc, cmn;
            int32_t coc, con;
            load_group(gk, cmc, coc);
            load_group(gkn, cmn, con);
            uint64_t rem = __ballot((cmc & below) == 0);
            int lastc = 0, lastco = 0;
            auto take = [&](int &c, int &co) -> bool {
                while (rem == 0) {
                    if (gkn >= nblk) return false;
                    gk = gkn;
                    coc = con;
                    rem = __ballot((cmn & below) == 0);
                    gkn = gk + 8 * V;
                    load_group(gkn, cmn, con);
                }
                const int li = __builtin_ctzll(rem);
                rem &= rem - 1;
                c = (int)((gk + (int64_t)(li >> 3) * V) * kFuseT + (li & 7));
                co = __builtin_amdgcn_readlane(coc, li);
                lastc = c;
                lastco = co;
                return true;
            };
            // the next STEP children of the stream (past its end, repeats of the last child)
            auto form = [&](Step &d) {
                d.on = 0;
                d.cv = lastc;
                d.cov = lastco;
#pragma unroll
                for (int q = 0; q < STEP; q++) {
                    int c = lastc, co = lastco;
                    if (take(c, co)) d.on |= 1u << q;
                    d.cv = lane == q ? c : d.cv;
                    d.cov = lane == q ? co : d.cov;
                }
            };
            // a step's entry words: lane q * NP + u holds entry u of child q's phase run
            auto entries = [&](const Step &d) -> int {
                const int q = min(lane / NP, STEP - 1);
                const int co = __shfl(d.cov, q);
                return fent[co + off + (lane / NP < STEP ? lane % NP : NP - 1)];
            };
            // a step's loads, the same set every step: H, carried G and the NP parent values of
            // each child
            auto issue = [&](const Step &d, int ew, Buf &bf) {
#pragma unroll
                for (int q = 0; q < STEP; q++) {
                    const int c = __builtin_amdgcn_readlane(d.cv, q);
                    bf.h[q] = *row_at(Hc, c, ldh, hrow);
                    bf.g0[q] = p > plo ? *row_at(Gc, c, ldg, grow) : E::zero();
#pragma unroll
                    for (int u = 0; u < NP; u++)
                        bf.g[q][u] = *row_at(Gs, __builtin_amdgcn_readlane(ew, q * NP + u) >> bits, ldg, grow);
                }
            };
            // a step's updates, children in order: G term in list order, dP terms into the
            // child's distinct slots
            auto update = [&](const Step &d, int ew, const Buf &bf) {
#pragma unroll
                for (int q = 0; q < STEP; q++) {
                    if ((d.on >> q) & 1u) {
                        HT acc = bf.g0[q], a[NP];
                        int sl[NP];
#pragma unroll
                        for (int u = 0; u < NP; u++) {
                            const int q1 = __builtin_amdgcn_readlane(ew, q * NP + u) & mask;
                            sl[u] = (q1 - 1 - qbase) * 64 + lane;
                            a[u] = tr[sl[u]];
                        }
#pragma unroll
                        for (int u = 0; u < NP; u++) {
                            acc = E::muladd(acc, prow[sl[u]], bf.g[q][u]);
                            tr[sl[u]] = E::muladd_hh(a[u], bf.h[q], bf.g[q][u]);
                        }
                        if (live) *const_cast<HT *>(row_at(Gc, __builtin_amdgcn_readlane(d.cv, q), ldg, grow)) = acc;
                    }
                }
            };
            if (rem == 0 && gkn >= nblk) return;   // no child of this phase in the stripe
            // three steps in flight: step n+2's entry words load, step n+1's values load, step n
            // updates.  Each step issues its loads in that order, so every wait leaves the later
            // ones pending.
            Step d0, d1, d2;
            form(d0);
            if (!d0.on) return;
            int e0 = entries(d0);
            form(d1);
            int e1 = entries(d1), e2;
            Buf b0, b1;
            issue(d0, e0, b0);
            auto iter = [&](const Buf &cur, Buf &nxt) {
                form(d2);
                e2 = entries(d2);
                issue(d1, e1, nxt);
                update(d0, e0, cur);
                d0 = d1;
                e0 = e1;
                d1 = d2;
                e1 = e2;
            };
            for (;;) {
                iter(b0, b1);
                if (!d0.on) break;
                iter(b1, b0);
                if (!d0.on) break;
            }
        });
        // the stripe's partial row: the waves' rows in wave order
        __syncthreads();
        const int64_t rb = prow0[p - plo] + s * nrow;
        for (int e = (int)threadIdx.x; e < nrow * 64; e += kFuseWaves * 64) {
            HT sum = rows[e];
#pragma unroll
            for (int w = 1; w < kFuseWaves; w++) sum = E::add_h(sum, rows[(size_t)w * nslot * 64 + e]);
            const int64_t bb = tile * 64 + (e & 63);
            if (bb < nb) part[(rb + (e >> 6)) * nb + bb] = sum;
        }
        __syncthreads();
        for (int u = 0; u < nrow; u++) tr[u * 64 + lane] = E::zero();
        off += np;
    }
}

// The same order without the phase loop, for small batches and small levels.  Four adjacent
// threads take the kFuseWaves wave-stripes of one (phase, stripe s, slot u, instance): each sums
// its term list (child, parent) with list_sum, and the first adds the four
// totals in wave order (shuffles within the wave) and writes the stripe's partial row
// (rows prow0[ph] + s (K - p) + u).  G comes from the adjoint gather over the sorted lists
// (k_rev_g / k_rev_g_bt), which sums each child's list in order.
// Sum of the terms H[c] * G[s] of one term list (child c, parent s) for instance b, in list
// order from zero.  Eight terms a batch; the next batch's words load while this batch's values
// are in flight, and every batch issues the same loads (past the end, repeats of the last term,
// whose products are not added).
template <int DT>
__device__ inline typename Elt<DT>::H list_sum(const int2 *__restrict__ terms, int64_t t0, int64_t t1,
                                               const typename Elt<DT>::H *__restrict__ Hc, int64_t ldh,
                                               const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
                                               int64_t b) {
    using E = Elt<DT>;
    using HT = typename E::H;
    constexpr int N = 8;
    HT tot = E::zero();
    const int64_t n = t1 - t0;
    if (n <= 0) return tot;
    int2 cs[N];
#pragma unroll
    for (int x = 0; x < N; x++) cs[x] = terms[t0 + min((int64_t)x, n - 1)];
    for (int64_t b0 = 0; b0 < n; b0 += N) {
        HT h[N], g[N];
#pragma unroll
        for (int x = 0; x < N; x++) {
            h[x] = Hc[(int64_t)cs[x].x * ldh + b];
            g[x] = Gs[(int64_t)cs[x].y * ldg + b];
        }
#pragma unroll
        for (int x = 0; x < N; x++) cs[x] = terms[t0 + min(b0 + N + x, n - 1)];
#pragma unroll
        for (int x = 0; x < N; x++)
            if (b0 + x < n) tot = E::muladd_hh(tot, h[x], g[x]);
    }
    return tot;
}

template <int DT>
__device__ inline void fstripe_body(
    int64_t bid, int64_t nbid, const int2 *__restrict__ terms, const int64_t *__restrict__ toff,
    const int64_t *__restrict__ prow0, int S, int nslot, int K, int plo, int nph,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg, typename Elt<DT>::H *__restrict__ part,
    int64_t nb) {
    using E = Elt<DT>;
    using HT = typename E::H;
    static_assert(kFuseWaves == 4 && kBlock % 4 == 0, "a stripe's wave-stripes are four adjacent lanes");
    const int64_t V = (int64_t)S * kFuseWaves;
    const int64_t total = (int64_t)nph * S * nslot * nb * kFuseWaves;   // a multiple of 4
    const int64_t stride = nbid * blockDim.x;
    const int64_t start = bid * blockDim.x + threadIdx.x;
    // every thread of a wave runs the same number of iterations (the shuffles need all four)
    for (int64_t i0 = start - (threadIdx.x & 63); i0 < total; i0 += stride) {
        const int64_t i = i0 + (threadIdx.x & 63);
        const bool in = i < total;
        const int w = (int)(i & 3);
        const int64_t q = in ? i >> 2 : 0;
        const int64_t b = q % nb, rest = q / nb;
        const int u = (int)(rest % nslot);
        const int64_t ps = rest / nslot;
        const int64_t s = ps % S;
        const int ph = (int)(ps / S);
        const int nrow = K - plo - ph;
        HT tot = E::zero();
        if (in && u < nrow) {
            const int64_t li = ((int64_t)ph * nslot + u) * V + s * kFuseWaves + w;
            tot = list_sum<DT>(terms, toff[li], toff[li + 1], Hc, ldh, Gs, ldg, b);
        }
        // the four totals in wave order
        const int lane = (int)(threadIdx.x & 63), l0 = lane & ~3;
        HT sum = shfl_h<DT>(tot, l0);
#pragma unroll
        for (int x = 1; x < kFuseWaves; x++) sum = E::add_h(sum, shfl_h<DT>(tot, l0 + x));
        if (in && w == 0 && u < nrow) part[(prow0[ph] + s * nrow + u) * nb + b] = sum;
    }
}

// Orders a wave's LDS writes before its later LDS reads (and reads before later writes): the LDS
// executes one wave's operations in order, so this only keeps the compiler from moving them.
__device__ inline void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The stripe sums for batches of at most 32 instances (k_rev_fsmall), in the order of
// fstripe_body: one workgroup per (phase, stripe s, slot u), wave w summing wave-stripe
// 4s + w's term list for every instance.  A chunk of CT terms is multiplied out with the lanes over
// (term, instance) pairs into LDS, then each instance's lane adds the chunk's products in list
// order; wave 0 adds the four totals in wave order.  The sums are list_sum's to the bit (a product
// is rounded before it meets the total), without its chain of one dependent round trip per eight
// terms, which on the largest levels' lists (up to 288 terms at N=28) set the kernel's time.
template <int DT, int BT>
__device__ inline void fstripe_lds_body(
    int64_t bid, int64_t nbid, const int2 *__restrict__ terms, const int64_t *__restrict__ toff,
    const int64_t *__restrict__ prow0, int S, int nslot, int K, int plo, int nph,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg, typename Elt<DT>::H *__restrict__ part,
    int64_t nb, typename Elt<DT>::H *lds) {
    using E = Elt<DT>;
    using HT = typename E::H;
    static_assert(kBlock == kFuseWaves * 64 && BT <= 32, "one wave per wave-stripe, BT <= 32");
    constexpr int CT = BT <= 16 ? 64 : 32;     // terms per chunk: at most 1024 products per wave
    constexpr int R = (CT * BT + 63) / 64;     // (term, instance) pairs per lane and chunk
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    HT *buf = lds + (size_t)w * CT * BT;
    HT *tw = lds + (size_t)kFuseWaves * CT * BT;   // the four waves' totals, [wave][instance]
    const int64_t V = (int64_t)S * kFuseWaves;
    const int64_t nitems = (int64_t)nph * S * nslot;
    for (int64_t it = bid; it < nitems; it += nbid) {
        const int u = (int)(it % nslot);
        const int64_t ps = it / nslot;
        const int64_t s = ps % S;
        const int ph = (int)(ps / S);
        const int nrow = K - plo - ph;
        if (u >= nrow) continue;   // the same for the whole workgroup
        const int64_t li = ((int64_t)ph * nslot + u) * V + s * kFuseWaves + w;
        const int64_t t0 = toff[li], t1 = toff[li + 1];
        HT acc = E::zero();
        for (int64_t c0 = t0; c0 < t1; c0 += CT) {
            const int n = (int)min((int64_t)CT, t1 - c0);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lane + 64 * r, t = idx / BT, b = idx % BT;
                if (t < n && b < nb) {
                    const int2 cs = terms[c0 + t];
                    buf[t * BT + b] = E::mul_hh(Hc[(int64_t)cs.x * ldh + b], Gs[(int64_t)cs.y * ldg + b]);
                }
            }
            wave_lds_order();
            if (lane < nb)
                for (int t = 0; t < n; t++) acc = E::add_h(acc, buf[t * BT + lane]);
            wave_lds_order();
        }
        if (lane < BT) tw[w * BT + lane] = acc;
        __syncthreads();
        if (w == 0 && lane < nb) {
            HT sum = tw[lane];
#pragma unroll
            for (int x = 1; x < kFuseWaves; x++) sum = E::add_h(sum, tw[x * BT + lane]);
            part[(prow0[ph] + s * nrow + u) * nb + lane] = sum;
        }
        __syncthreads();
    }
}

// The adjoint gather for batches of at most 32 on levels of few children with long parent lists
// (k_rev_fsmall): one wave per child, a chunk of CT list entries multiplied out with the lanes over
// (entry, instance) pairs into LDS, then each instance's lane adds the chunk's products in list
// order.  The sums of rev_g_bt_body to the bit, with one round trip per chunk instead of one per
// eight entries (lists of up to 105 entries at N=28).
template <int DT, int BT>
__device__ inline void rev_g_lanes_body(
    int64_t bid, int64_t nbid, const int32_t *__restrict__ cent, const int64_t *__restrict__ coff,
    int64_t nch, const typename Elt<DT>::P *__restrict__ P, int64_t ldp,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg, typename Elt<DT>::H *__restrict__ Gc,
    int64_t nb, typename Elt<DT>::H *lds) {
    using E = Elt<DT>;
    using HT = typename E::H;
    constexpr int bits = Packing<int32_t>::bits;
    constexpr int32_t mask = (1 << bits) - 1;
    constexpr int CT = BT <= 16 ? 64 : 32;
    constexpr int R = (CT * BT + 63) / 64;
    constexpr int WPB = kBlock / 64;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    HT *buf = lds + (size_t)w * CT * BT;
    for (int64_t c = bid * WPB + w; c < nch; c += nbid * WPB) {
        const int64_t e0 = coff[c], e1 = coff[c + 1];
        HT acc = E::zero();
        for (int64_t c0 = e0; c0 < e1; c0 += CT) {
            const int n = (int)min((int64_t)CT, e1 - c0);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lane + 64 * r, t = idx / BT, b = idx % BT;
                if (t < n && b < nb) {
                    const int32_t e = cent[c0 + t];
                    buf[t * BT + b] = E::mul_ph(P[((int64_t)(e & mask) - 1) * ldp + b],
                                                Gs[(int64_t)(e >> bits) * ldg + b]);
                }
            }
            wave_lds_order();
            if (lane < nb)
                for (int t = 0; t < n; t++) acc = E::add_h(acc, buf[t * BT + lane]);
            wave_lds_order();
        }
        if (lane < nb) Gc[c * ldg + lane] = acc;
    }
}

// LDS bytes of fstripe_lds_body<DT, BT> (rev_g_lanes_body uses a part of them).
template <int DT, int BT>
constexpr size_t fstripe_lds_bytes() {
    return (size_t)kFuseWaves * ((BT <= 16 ? 64 : 32) + 1) * BT * sizeof(typename Elt<DT>::H);
}

template <int DT>
__global__ __launch_bounds__(kBlock) void k_rev_fstripe(
    const int2 *__restrict__ terms, const int64_t *__restrict__ toff, const int64_t *__restrict__ prow0,
    int S, int nslot, int K, int plo, int nph, const typename Elt<DT>::H *__restrict__ Hc,
    int64_t ldh, const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ part, int64_t nb) {
    fstripe_body<DT>(blockIdx.x, gridDim.x, terms, toff, prow0, S, nslot, K, plo, nph, Hc, ldh, Gs, ldg,
                     part, nb);
}

// The small-batch reverse step of one level in one launch: workgroups [0, nga) run the adjoint
// gather over the sorted lists (BT lanes per child, or with `glanes` one wave per child,
// rev_g_lanes_body), the others the stripe sums (one workgroup per
// (phase, stripe, slot), fstripe_lds_body, with fstripe_lds_bytes<DT, BT>() of dynamic LDS); both
// read only G of the parents' level, so they are independent.
template <int DT, int BT>
__global__ __launch_bounds__(kBlock) void k_rev_fsmall(
    int64_t nga, const int32_t *__restrict__ cent, const int64_t *__restrict__ coff, int64_t nch,
    const typename Elt<DT>::P *__restrict__ P, int64_t ldp, typename Elt<DT>::H *__restrict__ Gc,
    int64_t ntiles, const int2 *__restrict__ terms, const int64_t *__restrict__ toff,
    const int64_t *__restrict__ prow0, int S, int nslot, int K, int plo, int nph,
    const typename Elt<DT>::H *__restrict__ Hc, int64_t ldh,
    const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg, typename Elt<DT>::H *__restrict__ part,
    int64_t nb, int glanes) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if ((int64_t)blockIdx.x < nga) {
        if (glanes)
            rev_g_lanes_body<DT, BT>(blockIdx.x, nga, cent, coff, nch, P, ldp, Gs, ldg, Gc, nb,
                                     reinterpret_cast<typename Elt<DT>::H *>(lds_raw));
        else
            rev_g_bt_body<DT, int32_t, BT>(blockIdx.x, nga, cent, coff, nch, P, ldp, Gs, ldg, Gc, nb, ntiles);
    } else
        fstripe_lds_body<DT, BT>(blockIdx.x - nga, (int64_t)gridDim.x - nga, terms, toff, prow0, S,
                                 nslot, K, plo, nph, Hc, ldh, Gs, ldg, part, nb,
                                 reinterpret_cast<typename Elt<DT>::H *>(lds_raw));
}

// k_rev_fstripe for batches of 64 or more: one workgroup per (phase, stripe s, slot u,
// 64-instance tile), wave w summing wave-stripe 4s + w's term list with lanes over instances
// (uniform terms, whole-row loads; list_sum), and the four totals added in wave order through LDS.
template <int DT>
__global__ __launch_bounds__(kFuseWaves * 64) void k_rev_fstripe_w(
    const int2 *__restrict__ terms, const int64_t *__restrict__ toff, const int64_t *__restrict__ prow0,
    int S, int nslot, int K, int plo, int nph, const typename Elt<DT>::H *__restrict__ Hc,
    int64_t ldh, const typename Elt<DT>::H *__restrict__ Gs, int64_t ldg,
    typename Elt<DT>::H *__restrict__ part, int64_t nb, int64_t ntiles) {
    using E = Elt<DT>;
    using HT = typename E::H;
    __shared__ HT tw[kFuseWaves][64];
    const int lane = (int)(threadIdx.x & 63);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t V = (int64_t)S * kFuseWaves;
    const int64_t nwork = (int64_t)nph * S * nslot * ntiles;
    for (int64_t g = blockIdx.x; g < nwork; g += gridDim.x) {
        const int64_t tile = g % ntiles, rest = g / ntiles;
        const int u = (int)(rest % nslot);
        const int64_t ps = rest / nslot;
        const int64_t s = ps % S;
        const int ph = (int)(ps / S);
        const int nrow = K - plo - ph;
        if (u >= nrow) continue;   // uniform over the workgroup
        const int64_t b = tile * 64 + lane;
        const int64_t bl = b < nb ? b : nb - 1;
        const int64_t li = ((int64_t)ph * nslot + u) * V + s * kFuseWaves + w;
        const HT tot = list_sum<DT>(terms, toff[li], toff[li + 1], Hc, ldh, Gs, ldg, bl);
        tw[w][lane] = tot;
        __syncthreads();
        if (w == 0) {
            HT sum = tw[0][lane];
#pragma unroll
            for (int x = 1; x < kFuseWaves; x++) sum = E::add_h(sum, tw[x][lane]);
            if (b < nb) part[(prow0[ph] + s * nrow + u) * nb + b] = sum;
        }
        __syncthreads();
    }
}

}  // namespace eggman

// Device-side generation of the plan's transition array.
//
// One thread per state of a level: unrank the state's colex position (plan_rank.hpp), then
// write its m-1 packed transitions in the reference's order.  The result is bit-identical to
// the reference's _build_dp_plan (src/dp.jl:248-304) and to the host builder, and is produced
// in HBM directly -- at K = 40 that is 21 GiB the host never has to build or upload.
#pragma once
#include "common.hpp"
#include "plan_rank.hpp"

namespace eggman {

__constant__ Binom d_binom = make_binom();

template <typename W>
__global__ __launch_bounds__(256) void k_plan_gen(W *__restrict__ trans_lvl, int64_t nst, int K,
                                                  int m, int lo, int64_t child_first, int bits) {
    const int n = K - lo + 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nst; r += stride) {
        int x[kMaxK];
        colex_unrank(d_binom.c, n, m, (uint64_t)r, x);
        state_transitions<W>(d_binom.c, K, m, lo, x, child_first, bits, trans_lvl + r * (m - 1));
    }
}

}  // namespace eggman

// Host-side plan layout, builder and validation.
#pragma once
#include <stdint.h>
#include <vector>

#include "common.hpp"
#include "plan_rank.hpp"

namespace eggman {

// Everything about a degree-K plan that follows from K alone.  Index L below is the
// reference's 1-based level number stored at [L-1]: level L holds the subsets of size 2(L-1).
struct PlanLayout {
    int K = 0, index_bytes = 4, bits = 9, nlevels = 0;
    int64_t npairs = 0, nstates = 0, ntrans = 0, max_level = 0;
    std::vector<int64_t> levels;   // 1-based first state of each level, plus nstates+1
    std::vector<int64_t> tstart;   // 0-based first transition of each level, plus ntrans
    std::vector<int> size;         // subset size of each level
    std::vector<int> lo;           // first index of the level's ground set
    std::vector<int64_t> nst;      // states in each level
};

const Binom &host_binom();
int pair_bits_for(int index_bytes);
int plan_layout(int K, int index_bytes, PlanLayout *lay);
int build_plan_host(const PlanLayout &lay, void *starts, void *trans, void *levels, int nthreads);
int validate_plan(const PlanLayout &lay, const void *starts, const void *trans, const void *levels,
                  int nthreads);

}  // namespace eggman

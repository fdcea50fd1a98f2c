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

// Index structures of the reverse sweep, derived from a forward transition array.  For the
// transitions of level L (parents in L, children in L-1):
//   child-major  cent/coff: for every state c (global, 0-based), its parent transitions
//                cent[coff[c] .. coff[c+1]) as (parent rank in its level << bits) | pair,
//                in ascending parent order;
//   pair-major   pent: the level's transitions sorted by (pair, parent), packed
//                (parent rank << 32) | child rank, cut into segments of at most `seg_len`
//                entries that never straddle a pair; qoff/qseg list each pair's segments in
//                ascending level order, which fixes the order of the final reduction.
struct RevSeg {
    int64_t start;   // first entry in pent
    int32_t len;     // entries
    int32_t level;   // 1-based level of the parents
};

struct RevPlanHost {
    std::vector<int64_t> coff;     // nstates + 1
    std::vector<uint8_t> cent;     // ntrans words of index_bytes each
    std::vector<uint64_t> pent;    // ntrans
    std::vector<RevSeg> segs;
    std::vector<int64_t> lvl_seg;  // first segment of each level (1-based L at [L-1]), + total
    std::vector<int64_t> qoff;     // npairs + 1
    std::vector<int32_t> qseg;     // segments of each pair, ascending level
};

int build_rev_plan(const PlanLayout &lay, const void *trans, int seg_len, RevPlanHost *rp);

const Binom &host_binom();
int pair_bits_for(int index_bytes);
int plan_layout(int K, int index_bytes, PlanLayout *lay);
int build_plan_host(const PlanLayout &lay, void *starts, void *trans, void *levels, int nthreads);
int validate_plan(const PlanLayout &lay, const void *starts, const void *trans, const void *levels,
                  int nthreads);

}  // namespace eggman

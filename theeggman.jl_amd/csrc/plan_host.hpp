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

// Parent-major dP tasks: the parents of one level with the same p = min(S), at most `len` of
// them, whose contributions H[child] * G[parent] all land in row p of dP (pairs (p, j), j > p).
// A task accumulates that row for its parents and writes it to rows [row0, row0 + K - p) of
// the partial buffer.
struct ParTask {
    int64_t start;   // first entry in prank
    int64_t row0;    // first partial row of this task
    int32_t len;     // parents
    int32_t p;       // 1-based row of dP
    int32_t level;   // 1-based level of the parents
    int32_t pad;
};

// The partial rows of dP row p are summed in two fixed-order stages: stage one sums runs of at
// most `fan_in` consecutive tasks of the row (ascending level, then ascending parent), one
// stage-one row per (run, column u); stage two sums a pair's stage-one rows in run order.
struct RedRow {
    int32_t tb, te;  // tasks rtask[tb .. te)
    int32_t u;       // column j - p - 1
    int32_t pad;
};

struct RevPlanHost {
    std::vector<int64_t> coff;     // nstates + 1
    std::vector<uint8_t> cent;     // ntrans words of index_bytes each
    std::vector<uint64_t> pent;    // ntrans
    std::vector<RevSeg> segs;
    std::vector<int64_t> lvl_seg;  // first segment of each level (1-based L at [L-1]), + total
    std::vector<int64_t> qoff;     // npairs + 1
    std::vector<int32_t> qseg;     // segments of each pair, ascending level
    // parent-major dP (kernels_rev.hpp: k_rev_dp_par)
    std::vector<int32_t> prank;    // per level, its parents' ranks grouped by p (state-indexed)
    std::vector<ParTask> tasks;
    std::vector<int64_t> lvl_task; // first task of each level, + total
    std::vector<int64_t> roff;     // tasks of each row p (index p, 1..K), ascending level
    std::vector<int32_t> rtask;
    std::vector<int32_t> qrow;     // 1-based row p of each 0-based pair index
    std::vector<int32_t> qcol;     // offset j - p - 1 of each pair inside its row
    int64_t part_rows = 0;
    std::vector<RedRow> red;       // stage-one rows
    std::vector<int32_t> srow0;    // first stage-one row of each row p (index p, 1..K)
    std::vector<int32_t> sruns;    // runs of each row p
};

int build_rev_plan(const PlanLayout &lay, const void *trans, int seg_len, int task_len,
                   int fan_in, RevPlanHost *rp);

// Where the gradient keeps the forward values of levels 1 .. K/2 (level L's values are read by
// the reverse step at L + 1; the top level's value is the result).  A kept level stays resident
// in the kept region at row `off[L]`; the other levels form segments (maximal runs of levels
// that are not kept) that share one segment region, each starting at row 0, and are recomputed
// from the kept level below when the reverse sweep reaches them.  Rows are state rows of one
// pass's instances.
struct GradLevels {
    std::vector<char> kept;       // index L
    std::vector<int64_t> off;     // index L
    int64_t kept_rows = 0, seg_rows = 0;
};

// Every level kept.
void grad_levels_full(const PlanLayout &lay, GradLevels *g);
// Level 1 and a set of checkpoint levels kept, chosen to minimise kept rows + the largest
// segment (exact, by a shortest-path search for every candidate segment bound).
void grad_levels_checkpointed(const PlanLayout &lay, GradLevels *g);

const Binom &host_binom();
int pair_bits_for(int index_bytes);
int plan_layout(int K, int index_bytes, PlanLayout *lay);
int build_plan_host(const PlanLayout &lay, void *starts, void *trans, void *levels, int nthreads);
int validate_plan(const PlanLayout &lay, const void *starts, const void *trans, const void *levels,
                  int nthreads);

}  // namespace eggman

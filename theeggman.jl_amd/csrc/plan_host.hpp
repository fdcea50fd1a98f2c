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

// Fused child-major reverse step (kernels_rev.hpp: k_rev_fused), Int32 layout.  Level L holds
// all m-subsets of {lo..K} (plan_rank.hpp), so the parents of a child c of level L-1 are exactly
// c + {p, j} with lo <= p < min(c) and j > p, j not in c: for each p ("phase") there are
// n_p = K - p - (m - 2) of them, and every child with min(c) > p has all n_p.  A child's parent
// list sorted by pair, (p, j) ascending, therefore holds its phases one after another, phase p
// starting at the same offset sum_{p' < p} n_p' in every list.
//
// The order every fused kernel sums dP in, a function of the plan only:
//   * children are cut into blocks of kFuseT consecutive ranks; block k belongs to wave-stripe
//     k mod V, V = S * kFuseWaves, and wave-stripe v to workgroup stripe v / kFuseWaves;
//   * total(p, v)[u] = the terms H[c] * G[parent] of slot u = j - p - 1 over the children c of
//     stripe v's blocks with min(c) > p, blocks ascending and ranks ascending within a block,
//     added one after another from zero;
//   * the stripe's partial row (p, s)[u] = total(p, 4s) + total(p, 4s+1) + ... in wave order;
//   * dP[(p, j)] = the partial rows of row p, levels ascending, stripes ascending, summed in
//     runs of `fan_in` and then over the runs (as the parent-major tasks are).
// And G[c] = the terms P[(p, j)] * G[parent] of c's sorted list, in list order, from zero.
// For small batches the terms of each (phase, slot, wave-stripe) are also listed in that order
// (`terms`), so one thread can sum them without visiting the other children.
constexpr int kFuseT = 8;        // consecutive child ranks per block
constexpr int kFuseWaves = 4;    // wave-stripes per workgroup stripe

struct FuseLevel {
    int S = 0;                   // workgroup stripes
    int plo = 0, phi = 0;        // phases p = plo .. phi
    int nslot = 0;               // K - plo: slots of the widest phase
    int csize = 0;               // subset size of the children
    int64_t nch = 0, nblk = 0;   // children, blocks
    int64_t cfirst = 0;          // 0-based first child state
    int64_t ent0 = 0;            // first entry of the children's sorted lists
    int64_t prow = 0;            // index in prow0 of phase plo
    int64_t toff = 0;            // index in toff of (phase plo, slot 0, wave-stripe 0)
};

struct FusePlanHost {
    std::vector<int32_t> fent;    // ntrans: each child's parent entries sorted by pair
    std::vector<int32_t> fco;     // nstates: start of a child's list, relative to its level's first
    std::vector<uint32_t> cmask;  // nstates: bit e-1 set for each element e
    std::vector<FuseLevel> lv;    // index L (the parents' level), 2 .. K/2 + 1
    std::vector<int64_t> prow0;   // first partial row of each (level, phase)
    // the terms of each (level, phase, slot u < nslot, wave-stripe v), in summation order, as
    // (child rank, parent rank) within their levels: terms[toff[i] .. toff[i + 1]), i = lv.toff +
    // (phase * nslot + u) * V + v
    std::vector<int32_t> terms;   // pairs of int32
    std::vector<int64_t> toff;
    int64_t part_rows = 0;
    std::vector<ParTask> tasks;   // one per (level, phase, stripe): row0, p, level
    std::vector<int32_t> rtask;   // tasks of each row p, levels then stripes ascending
    std::vector<int64_t> roff;    // index p, 1..K
    std::vector<RedRow> red;
    std::vector<int32_t> srow0, sruns;
};

// Workgroup stripes of a level: about 16 blocks per wave-stripe, at most one XCD's worth of
// resident workgroups (32 CUs x 2).  Fixed per degree, so the dP order is too.
int fuse_stripes(int K, int plo, int64_t nblk);

int build_fuse_plan(const PlanLayout &lay, const RevPlanHost &rp, int fan_in, FusePlanHost *fp);

// Checks the tables of build_fuse_plan against the plan's transitions (see eggman_fuse_validate).
int validate_fuse_plan(const PlanLayout &lay, const int32_t *trans, const RevPlanHost &rp,
                       const FusePlanHost &fp, int64_t *entries, int64_t *terms);

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

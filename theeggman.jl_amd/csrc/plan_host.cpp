// Host side of the plan: closed-form counts and layout, the host builder, and plan validation.
//
// The values written here are those of the reference's _build_dp_plan (src/dp.jl:248-304):
// states numbered by (subset size, mask value), transitions per state in ascending partner
// order, packed (child << bits) | pair with 1-based child and pair, starts/levels 1-based.
// The construction is different: each level's reachable subsets are all m-subsets of one
// ground set (see plan_rank.hpp), so states are produced in order by stepping through colex
// order and children are ranked arithmetically -- no set, no sort, no dictionary, and every
// level can be split across threads.
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "plan_host.hpp"

namespace eggman {

static constexpr Binom kHostBinom = make_binom();
const Binom &host_binom() { return kHostBinom; }

int pair_bits_for(int index_bytes) { return index_bytes == 4 ? 9 : 12; }   // src/dp.jl:170-171

int plan_layout(int K, int index_bytes, PlanLayout *lay) {
    if (K < 2 || K > kMaxK || (K & 1))
        return fail(EGGMAN_EINVAL, "degree K=%d must be even and in [2, %d]", K, kMaxK);
    if (index_bytes != 4 && index_bytes != 8)
        return fail(EGGMAN_EINVAL, "index_bytes must be 4 or 8, got %d", index_bytes);
    const uint64_t(*C)[kMaxK + 1] = kHostBinom.c;
    lay->K = K;
    lay->index_bytes = index_bytes;
    lay->bits = pair_bits_for(index_bytes);
    lay->npairs = (int64_t)K * (K - 1) / 2;
    lay->nlevels = K / 2 + 2;
    lay->levels.assign(lay->nlevels, 0);
    lay->tstart.assign(lay->nlevels, 0);
    lay->size.assign(lay->nlevels, 0);
    lay->lo.assign(lay->nlevels, 0);
    lay->nst.assign(lay->nlevels, 0);
    int64_t s = 1, t = 0;
    for (int L = 1; L <= K / 2 + 1; L++) {          // level L holds subsets of size 2(L-1)
        const int m = 2 * (L - 1);
        const int lo = ground_lo(K, m);
        const int n = K - lo + 1;
        const int64_t cnt = (int64_t)C[n][m];
        lay->levels[L - 1] = s;
        lay->tstart[L - 1] = t;
        lay->size[L - 1] = m;
        lay->lo[L - 1] = lo;
        lay->nst[L - 1] = cnt;
        s += cnt;
        t += cnt * (m > 0 ? m - 1 : 0);
    }
    lay->levels[lay->nlevels - 1] = s;
    lay->tstart[lay->nlevels - 1] = t;
    lay->nstates = s - 1;
    lay->ntrans = t;
    lay->max_level = 0;
    for (int L = 1; L <= K / 2 + 1; L++) lay->max_level = std::max(lay->max_level, lay->nst[L - 1]);

    // the refusals of _build_dp_plan (src/dp.jl:224-234)
    const int64_t imax = index_bytes == 4 ? INT32_MAX : INT64_MAX;
    if (lay->npairs > ((int64_t)1 << lay->bits) - 1)
        return fail(EGGMAN_EPLAN, "degree %d needs more than %d bits per pair index in Int%d", K,
                    lay->bits, index_bytes * 8);
    if (lay->nstates > (imax >> lay->bits))
        return fail(EGGMAN_EPLAN, "degree %d has too many states to pack into an Int%d transition",
                    K, index_bytes * 8);
    if (lay->ntrans > imax)
        return fail(EGGMAN_EPLAN, "degree %d has too many transitions to index with Int%d", K,
                    index_bytes * 8);
    return EGGMAN_OK;
}

// Fill the transitions of level L (1-based) for ranks [r0, r1) into `trans` (whole-plan array)
// and the matching `starts` entries.
template <typename W>
static void build_level_range(const PlanLayout &lay, int L, int64_t r0, int64_t r1, W *trans,
                              W *starts) {
    const uint64_t(*C)[kMaxK + 1] = kHostBinom.c;
    const int m = lay.size[L - 1];
    const int lo = lay.lo[L - 1];
    const int n = lay.K - lo + 1;
    const int k = m - 1;
    const int64_t first = lay.levels[L - 1];
    const int64_t child_first = lay.levels[L - 2];
    int x[kMaxK];
    colex_unrank(C, n, m, (uint64_t)r0, x);
    for (int64_t r = r0; r < r1; r++) {
        const int64_t t0 = lay.tstart[L - 1] + r * k;
        state_transitions<W>(C, lay.K, m, lo, x, child_first, lay.bits, trans + t0);
        starts[first + r] = (W)(t0 + k + 1);     // starts[s+1] in 1-based terms
        if (r + 1 < r1) colex_next(m, x);
    }
}

template <typename W>
static int build_plan_t(const PlanLayout &lay, W *starts, W *trans, W *levels, int nthreads) {
    for (int L = 0; L < lay.nlevels; L++) levels[L] = (W)lay.levels[L];
    starts[0] = 1;
    starts[1] = 1;                                   // the empty set has no transitions
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    struct Job { int L; int64_t r0, r1; };
    std::vector<Job> jobs;
    const int64_t chunk = 1 << 15;
    for (int L = 2; L <= lay.K / 2 + 1; L++)
        for (int64_t r = 0; r < lay.nst[L - 1]; r += chunk)
            jobs.push_back({L, r, std::min(lay.nst[L - 1], r + chunk)});
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        for (;;) {
            const size_t j = next.fetch_add(1);
            if (j >= jobs.size()) break;
            build_level_range<W>(lay, jobs[j].L, jobs[j].r0, jobs[j].r1, trans, starts);
        }
    };
    const int nt = (int)std::min<size_t>((size_t)nthreads, std::max<size_t>(1, jobs.size()));
    if (nt <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int i = 0; i < nt; i++) pool.emplace_back(worker);
        for (auto &th : pool) th.join();
    }
    return EGGMAN_OK;
}

int build_plan_host(const PlanLayout &lay, void *starts, void *trans, void *levels, int nthreads) {
    if (lay.index_bytes == 4)
        return build_plan_t<int32_t>(lay, (int32_t *)starts, (int32_t *)trans, (int32_t *)levels,
                                     nthreads);
    return build_plan_t<int64_t>(lay, (int64_t *)starts, (int64_t *)trans, (int64_t *)levels,
                                 nthreads);
}

// Two-stage reduction tables of partial rows.  row_tasks[p] lists the tasks of dP row p in their
// summation order; stage one sums runs of at most `fan_in` consecutive tasks, one stage-one row
// per (run, column u); stage two sums a pair's stage-one rows in run order.
static void build_reduction(int K, const std::vector<std::vector<int32_t>> &row_tasks, int fan_in,
                            std::vector<int64_t> *roff, std::vector<int32_t> *rtask,
                            std::vector<RedRow> *red, std::vector<int32_t> *srow0,
                            std::vector<int32_t> *sruns) {
    roff->assign(K + 2, 0);
    rtask->clear();
    for (int p = 1; p <= K; p++) {
        (*roff)[p] = (int64_t)rtask->size();
        rtask->insert(rtask->end(), row_tasks[p].begin(), row_tasks[p].end());
    }
    (*roff)[K + 1] = (int64_t)rtask->size();
    red->clear();
    srow0->assign(K + 2, 0);
    sruns->assign(K + 2, 0);
    for (int p = 1; p <= K; p++) {
        (*srow0)[p] = (int32_t)red->size();
        const int64_t tb = (*roff)[p], te = (*roff)[p + 1];
        for (int64_t a = tb; a < te; a += fan_in) {
            (*sruns)[p]++;
            for (int u = 0; u < K - p; u++)
                red->push_back(RedRow{(int32_t)a, (int32_t)std::min<int64_t>(te, a + fan_in), u, 0});
        }
    }
}

template <typename W>
static int build_rev_t(const PlanLayout &lay, const W *trans, int seg_len, int task_len,
                       int fan_in, RevPlanHost *rp) {
    const int64_t mask = ((int64_t)1 << lay.bits) - 1;
    rp->coff.assign(lay.nstates + 1, 0);
    rp->cent.assign((size_t)lay.ntrans * sizeof(W), 0);
    rp->pent.assign(lay.ntrans, 0);
    rp->segs.clear();
    rp->lvl_seg.assign(lay.nlevels, 0);
    W *cent = (W *)rp->cent.data();
    // parents per child (each child belongs to the level just below its parents)
    for (int64_t t = 0; t < lay.ntrans; t++) rp->coff[((int64_t)trans[t] >> lay.bits) - 1 + 1]++;
    for (int64_t s = 0; s < lay.nstates; s++) rp->coff[s + 1] += rp->coff[s];
    std::vector<int64_t> fill(rp->coff.begin(), rp->coff.end() - 1);
    std::vector<std::vector<int32_t>> per_pair(lay.npairs);
    for (int L = 2; L <= lay.K / 2 + 1; L++) {
        const int k = lay.size[L - 1] - 1;
        const int64_t cfirst = lay.levels[L - 2];
        const int64_t t0 = lay.tstart[L - 1], t1 = lay.tstart[L];
        // child-major, parents visited in ascending order
        for (int64_t t = t0; t < t1; t++) {
            const int64_t w = (int64_t)trans[t];
            const int64_t srank = (t - t0) / k;
            const int64_t c = (w >> lay.bits) - 1;
            cent[fill[c]++] = (W)((srank << lay.bits) | (w & mask));
        }
        // pair-major: counting sort of the level's transitions by pair, stable in parent order
        std::vector<int64_t> cnt(lay.npairs + 1, 0);
        for (int64_t t = t0; t < t1; t++) cnt[((int64_t)trans[t] & mask) - 1 + 1]++;
        for (int64_t q = 0; q < lay.npairs; q++) cnt[q + 1] += cnt[q];
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t t = t0; t < t1; t++) {
            const int64_t w = (int64_t)trans[t];
            const int64_t q = (w & mask) - 1;
            const uint64_t srank = (uint64_t)((t - t0) / k);
            const uint64_t crank = (uint64_t)((w >> lay.bits) - cfirst);
            rp->pent[t0 + pos[q]++] = (srank << 32) | crank;
        }
        rp->lvl_seg[L - 1] = (int64_t)rp->segs.size();
        for (int64_t q = 0; q < lay.npairs; q++) {
            for (int64_t a = cnt[q]; a < cnt[q + 1]; a += seg_len) {
                RevSeg sg;
                sg.start = t0 + a;
                sg.len = (int32_t)std::min<int64_t>(seg_len, cnt[q + 1] - a);
                sg.level = L;
                per_pair[q].push_back((int32_t)rp->segs.size());
                rp->segs.push_back(sg);
            }
        }
    }
    rp->lvl_seg[lay.nlevels - 1] = (int64_t)rp->segs.size();
    rp->qoff.assign(lay.npairs + 1, 0);
    rp->qseg.clear();
    for (int64_t q = 0; q < lay.npairs; q++) {
        rp->qoff[q] = (int64_t)rp->qseg.size();
        rp->qseg.insert(rp->qseg.end(), per_pair[q].begin(), per_pair[q].end());
    }
    rp->qoff[lay.npairs] = (int64_t)rp->qseg.size();

    // parent-major dP tasks
    const int K = lay.K;
    rp->qrow.assign(lay.npairs, 0);
    rp->qcol.assign(lay.npairs, 0);
    for (int i = 1; i < K; i++)
        for (int j = i + 1; j <= K; j++) {
            const int64_t q = pair_index(i, j, K) - 1;
            rp->qrow[q] = i;
            rp->qcol[q] = j - i - 1;
        }
    rp->prank.assign(lay.nstates, 0);
    rp->tasks.clear();
    rp->lvl_task.assign(lay.nlevels, 0);
    std::vector<std::vector<int32_t>> row_tasks(K + 1);
    int64_t rows = 0;
    for (int L = 2; L <= K / 2 + 1; L++) {
        rp->lvl_task[L - 1] = (int64_t)rp->tasks.size();
        const int k = lay.size[L - 1] - 1;
        const int64_t n = lay.nst[L - 1], base = lay.levels[L - 1] - 1, t0 = lay.tstart[L - 1];
        std::vector<std::vector<int32_t>> byp(K + 1);
        for (int64_t r = 0; r < n; r++) {
            const int64_t q = ((int64_t)trans[t0 + r * k] & mask) - 1;
            // the LDS row of a task holds row p only, and a batch of a parent's transitions
            // updates its slots together: the pairs must share the row and strictly ascend
            for (int v = 1; v < k; v++) {
                const int64_t qv = ((int64_t)trans[t0 + r * k + v] & mask) - 1;
                const int64_t qu = ((int64_t)trans[t0 + r * k + v - 1] & mask) - 1;
                if (rp->qrow[qv] != rp->qrow[q] || qv <= qu) {
                    rp->tasks.clear();
                    rp->lvl_task.clear();
                    return EGGMAN_OK;
                }
            }
            byp[rp->qrow[q]].push_back((int32_t)r);
        }
        int64_t pos = base;
        for (int p = 1; p <= K; p++) {
            const auto &v = byp[p];
            for (size_t a = 0; a < v.size(); a += task_len) {
                ParTask tk;
                tk.start = pos + (int64_t)a;
                tk.len = (int32_t)std::min<size_t>(task_len, v.size() - a);
                tk.p = p;
                tk.level = L;
                tk.row0 = rows;
                tk.pad = 0;
                rows += K - p;
                row_tasks[p].push_back((int32_t)rp->tasks.size());
                rp->tasks.push_back(tk);
            }
            for (size_t a = 0; a < v.size(); a++) rp->prank[pos + (int64_t)a] = v[a];
            pos += (int64_t)v.size();
        }
    }
    rp->lvl_task[lay.nlevels - 1] = (int64_t)rp->tasks.size();
    rp->part_rows = rows;
    build_reduction(K, row_tasks, fan_in, &rp->roff, &rp->rtask, &rp->red, &rp->srow0, &rp->sruns);
    return EGGMAN_OK;
}

int build_rev_plan(const PlanLayout &lay, const void *trans, int seg_len, int task_len,
                   int fan_in, RevPlanHost *rp) {
    if (seg_len < 1 || task_len < 1 || fan_in < 1)
        return fail(EGGMAN_EINVAL, "segment, task and fan-in lengths must be positive");
    if (lay.index_bytes == 4)
        return build_rev_t<int32_t>(lay, (const int32_t *)trans, seg_len, task_len, fan_in, rp);
    return build_rev_t<int64_t>(lay, (const int64_t *)trans, seg_len, task_len, fan_in, rp);
}

void grad_levels_full(const PlanLayout &lay, GradLevels *g) {
    const int n = lay.K / 2;   // levels 1 .. n hold values the reverse sweep reads
    g->kept.assign(n + 1, 1);
    g->off.assign(n + 1, 0);
    g->kept_rows = 0;
    g->seg_rows = 0;
    for (int L = 1; L <= n; L++) {
        g->off[L] = g->kept_rows;
        g->kept_rows += lay.nst[L - 1];
    }
}

void grad_levels_checkpointed(const PlanLayout &lay, GradLevels *g) {
    const int n = lay.K / 2;
    std::vector<int64_t> S(n + 1, 0);   // prefix sums of level rows
    for (int L = 1; L <= n; L++) S[L] = S[L - 1] + lay.nst[L - 1];
    auto run = [&](int a, int b) { return a > b ? (int64_t)0 : S[b] - S[a - 1]; };
    std::vector<int64_t> bounds{0};
    for (int a = 2; a <= n; a++)
        for (int b = a; b <= n; b++) bounds.push_back(run(a, b));
    std::sort(bounds.begin(), bounds.end());
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
    const int64_t INF = INT64_MAX / 4;
    int64_t best = INF;
    std::vector<char> best_kept;
    std::vector<int64_t> f(n + 1), prev(n + 1);
    for (int64_t T : bounds) {
        if (T >= best) break;
        // f[j]: fewest kept rows for levels 1..j with j kept and every segment below j <= T
        f[1] = lay.nst[0];
        prev[1] = 0;
        for (int j = 2; j <= n; j++) {
            f[j] = INF;
            for (int i = 1; i < j; i++)
                if (f[i] < INF && run(i + 1, j - 1) <= T && f[i] + lay.nst[j - 1] < f[j]) {
                    f[j] = f[i] + lay.nst[j - 1];
                    prev[j] = i;
                }
        }
        int last = 0;
        int64_t kept = INF;
        for (int i = 1; i <= n; i++)
            if (f[i] < kept && run(i + 1, n) <= T) {
                kept = f[i];
                last = i;
            }
        if (kept >= INF) continue;
        std::vector<char> k(n + 1, 0);
        for (int i = last; i >= 1; i = (int)prev[i]) k[i] = 1;
        int64_t seg = 0, cur = 0;
        for (int L = 2; L <= n; L++) {
            cur = k[L] ? 0 : cur + lay.nst[L - 1];
            seg = std::max(seg, cur);
        }
        if (kept + seg < best) {
            best = kept + seg;
            best_kept = k;
        }
    }
    g->kept = best_kept;
    g->off.assign(n + 1, 0);
    g->kept_rows = 0;
    g->seg_rows = 0;
    int64_t cur = 0;
    for (int L = 1; L <= n; L++) {
        if (g->kept[L]) {
            g->off[L] = g->kept_rows;
            g->kept_rows += lay.nst[L - 1];
            cur = 0;
        } else {
            g->off[L] = cur;
            cur += lay.nst[L - 1];
            g->seg_rows = std::max(g->seg_rows, cur);
        }
    }
}

// Check that caller-supplied arrays are a plan the kernels can walk safely: the closed-form
// sizes and level boundaries, a uniform out-degree per level (what lets the device drop
// `starts`), and every transition pointing into the level below with a valid pair index.
template <typename W>
static int validate_t(const PlanLayout &lay, const W *starts, const W *trans, const W *levels,
                      int nthreads) {
    for (int L = 0; L < lay.nlevels; L++)
        if ((int64_t)levels[L] != lay.levels[L])
            return fail(EGGMAN_EPLAN, "levels[%d] = %lld, expected %lld", L + 1,
                        (long long)levels[L], (long long)lay.levels[L]);
    if ((int64_t)starts[0] != 1 || (int64_t)starts[lay.nstates] != lay.ntrans + 1)
        return fail(EGGMAN_EPLAN, "starts does not span the transitions");
    for (int L = 1; L <= lay.K / 2 + 1; L++) {
        const int k = std::max(0, lay.size[L - 1] - 1);
        for (int64_t r = 0; r < lay.nst[L - 1]; r++) {
            const int64_t s = lay.levels[L - 1] + r;     // 1-based state
            if ((int64_t)starts[s - 1] != lay.tstart[L - 1] + r * k + 1 ||
                (int64_t)starts[s] != lay.tstart[L - 1] + (r + 1) * k + 1)
                return fail(EGGMAN_EPLAN, "state %lld does not have the %d transitions of its level",
                            (long long)s, k);
        }
    }
    const int64_t mask = ((int64_t)1 << lay.bits) - 1;
    std::atomic<int> bad(0);
    auto check = [&](int L, int64_t t0, int64_t t1) {
        const int64_t cf = lay.levels[L - 2], cl = lay.levels[L - 1];   // child level [cf, cl)
        // a state of this level has min(S) in [lo, K - m + 1], so its pairs lie in these rows
        const int m = lay.size[L - 1];
        const int64_t plo = lay.lo[L - 1], phi = std::min(lay.K - m + 1, lay.K - 1);
        const int64_t qlo = pair_index(plo, plo + 1, lay.K), qhi = pair_index(phi, lay.K, lay.K);
        for (int64_t t = t0; t < t1; t++) {
            const int64_t w = (int64_t)trans[t];
            const int64_t q = w & mask, c = w >> lay.bits;
            if (q < qlo || q > qhi || c < cf || c >= cl) { bad.store(1); return; }
        }
    };
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (int L = 2; L <= lay.K / 2 + 1; L++) {
        const int64_t a = lay.tstart[L - 1], b = lay.tstart[L];
        const int64_t per = std::max<int64_t>(1 << 20, (b - a + nthreads - 1) / nthreads);
        for (int64_t t = a; t < b; t += per) {
            pool.emplace_back(check, L, t, std::min(b, t + per));
            if ((int)pool.size() >= nthreads) {
                for (auto &th : pool) th.join();
                pool.clear();
            }
        }
    }
    for (auto &th : pool) th.join();
    if (bad.load()) return fail(EGGMAN_EPLAN, "a transition points outside the level below");
    return EGGMAN_OK;
}

int validate_plan(const PlanLayout &lay, const void *starts, const void *trans, const void *levels,
                  int nthreads) {
    if (lay.index_bytes == 4)
        return validate_t<int32_t>(lay, (const int32_t *)starts, (const int32_t *)trans,
                                   (const int32_t *)levels, nthreads);
    return validate_t<int64_t>(lay, (const int64_t *)starts, (const int64_t *)trans,
                               (const int64_t *)levels, nthreads);
}

int fuse_stripes(int K, int plo, int64_t nblk) {
    (void)K;
    (void)plo;
    // the fused kernel runs two workgroups per CU (its registers allow two waves per SIMD), so an
    // XCD (32 CUs) holds 64 stripes of a tile at once and they sweep the level together
    const int64_t want = (nblk + 4 * kFuseWaves - 1) / (4 * kFuseWaves);
    return (int)std::max<int64_t>(1, std::min<int64_t>(64, want));
}

int build_fuse_plan(const PlanLayout &lay, const RevPlanHost &rp, int fan_in, FusePlanHost *fp) {
    if (lay.index_bytes != 4) return fail(EGGMAN_EINVAL, "the fused reverse step needs the Int32 layout");
    if (fan_in < 1) return fail(EGGMAN_EINVAL, "fan-in must be positive");
    const uint64_t(*C)[kMaxK + 1] = kHostBinom.c;
    const int K = lay.K;
    const int64_t mask = ((int64_t)1 << lay.bits) - 1;
    const int32_t *cent = (const int32_t *)rp.cent.data();
    // each child's parent entries sorted by pair; a child has at most one parent per pair
    fp->fent.assign(cent, cent + lay.ntrans);
    fp->fco.assign(lay.nstates, 0);
    for (int L = 1; L <= K / 2; L++) {
        const int64_t c0 = lay.levels[L - 1] - 1, c1 = c0 + lay.nst[L - 1];
        for (int64_t c = c0; c < c1; c++) {
            std::sort(fp->fent.begin() + rp.coff[c], fp->fent.begin() + rp.coff[c + 1],
                      [&](int32_t a, int32_t b) { return (a & mask) < (b & mask); });
            fp->fco[c] = (int32_t)(rp.coff[c] - rp.coff[c0]);
        }
    }
    // element masks
    fp->cmask.assign(lay.nstates, 0);
    for (int L = 2; L <= K / 2 + 1; L++) {
        const int m = lay.size[L - 1], lo = lay.lo[L - 1];
        const int64_t s0 = lay.levels[L - 1] - 1;
        int x[kMaxK];
        colex_unrank(C, K - lo + 1, m, 0, x);
        for (int64_t r = 0; r < lay.nst[L - 1]; r++) {
            uint32_t bm = 0;
            for (int i = 0; i < m; i++) bm |= 1u << (lo + x[i] - 1);
            fp->cmask[s0 + r] = bm;
            if (r + 1 < lay.nst[L - 1]) colex_next(m, x);
        }
    }
    // levels, phases and partial rows
    fp->lv.assign(K / 2 + 2, FuseLevel{});
    fp->prow0.clear();
    fp->tasks.clear();
    std::vector<std::vector<int32_t>> row_tasks(K + 1);
    int64_t rows = 0;
    for (int L = 2; L <= K / 2 + 1; L++) {
        FuseLevel &f = fp->lv[L];
        f.plo = lay.lo[L - 1];
        f.phi = 2 * f.plo - 1;
        f.nslot = K - f.plo;
        f.csize = lay.size[L - 2];
        f.nch = lay.nst[L - 2];
        f.nblk = (f.nch + kFuseT - 1) / kFuseT;
        f.cfirst = lay.levels[L - 2] - 1;
        f.ent0 = rp.coff[f.cfirst];
        f.S = fuse_stripes(K, f.plo, f.nblk);
        f.prow = (int64_t)fp->prow0.size();
        for (int p = f.plo; p <= f.phi; p++) {
            const int nrow = K - p;
            if (nrow - f.csize < 1) return fail(EGGMAN_EINVAL, "phase %d of level %d has no parents", p, L);
            fp->prow0.push_back(rows);
            for (int s = 0; s < f.S; s++) {
                ParTask tk;
                tk.start = 0;
                tk.row0 = rows + (int64_t)s * nrow;
                tk.len = 0;
                tk.p = p;
                tk.level = L;
                tk.pad = 0;
                row_tasks[p].push_back((int32_t)fp->tasks.size());
                fp->tasks.push_back(tk);
            }
            rows += (int64_t)f.S * nrow;
        }
    }
    fp->part_rows = rows;
    build_reduction(K, row_tasks, fan_in, &fp->roff, &fp->rtask, &fp->red, &fp->srow0, &fp->sruns);

    // the terms of every (level, phase, slot, wave-stripe) in summation order: count, then fill,
    // both visiting children in rank order
    fp->toff.clear();
    fp->terms.assign((size_t)2 * lay.ntrans, 0);
    int64_t pos = 0;
    for (int L = 2; L <= K / 2 + 1; L++) {
        FuseLevel &f = fp->lv[L];
        const int nph = f.phi - f.plo + 1;
        const int64_t V = (int64_t)f.S * kFuseWaves;
        const int64_t nlist = (int64_t)nph * f.nslot * V;
        f.toff = (int64_t)fp->toff.size();
        std::vector<int64_t> cnt(nlist + 1, 0);
        auto visit = [&](auto &&emit) {
            for (int64_t c = 0; c < f.nch; c++) {
                const uint32_t cm = fp->cmask[f.cfirst + c];
                const int64_t v = (c / kFuseT) % V;
                const int32_t *lst = fp->fent.data() + f.ent0 + fp->fco[f.cfirst + c];
                int off = 0;
                for (int p = f.plo; p <= f.phi; p++) {
                    const int np = K - p - f.csize;
                    if ((cm & ((1u << p) - 1u)) != 0) break;     // min(c) <= p: no later phase either
                    for (int u = 0; u < np; u++) {
                        const int32_t e = lst[off + u];
                        const int q = (int)(e & mask);
                        // slot of pair (p, j): j - p - 1 = q - 1 - (0-based index of (p, p + 1))
                        const int slot = q - 1 - (int)(pair_index(p, p + 1, K) - 1);
                        emit(((int64_t)(p - f.plo) * f.nslot + slot) * V + v, (int32_t)c,
                             (int32_t)(e >> lay.bits));
                    }
                    off += np;
                }
            }
        };
        visit([&](int64_t li, int32_t, int32_t) { cnt[li + 1]++; });
        for (int64_t i = 0; i < nlist; i++) cnt[i + 1] += cnt[i];
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        visit([&](int64_t li, int32_t c, int32_t sr) {
            const int64_t at = pos + fill[li]++;
            fp->terms[(size_t)2 * at] = c;
            fp->terms[(size_t)2 * at + 1] = sr;
        });
        for (int64_t i = 0; i < nlist; i++) fp->toff.push_back(pos + cnt[i]);
        pos += cnt[nlist];
    }
    fp->toff.push_back(pos);
    if (pos != lay.ntrans) return fail(EGGMAN_EINVAL, "fused term lists hold %lld of %lld transitions",
                                       (long long)pos, (long long)lay.ntrans);
    return EGGMAN_OK;
}

int validate_fuse_plan(const PlanLayout &lay, const int32_t *trans, const RevPlanHost &rp,
                       const FusePlanHost &fp, int64_t *entries, int64_t *terms) {
    const int K = lay.K;
    const int32_t mask = (1 << lay.bits) - 1;
    std::vector<int> qi(lay.npairs + 1), qj(lay.npairs + 1);
    for (int i = 1; i < K; i++)
        for (int j = i + 1; j <= K; j++) {
            qi[pair_index(i, j, K)] = i;
            qj[pair_index(i, j, K)] = j;
        }
    // element masks against the plan: a state's mask is its child's plus {p, j}
    for (int L = 2; L <= K / 2 + 1; L++) {
        const int k = lay.size[L - 1] - 1;
        const int64_t s0 = lay.levels[L - 1] - 1;
        for (int64_t r = 0; r < lay.nst[L - 1]; r++)
            for (int v = 0; v < k; v++) {
                const int32_t w = trans[lay.tstart[L - 1] + r * k + v];
                const int64_t c = (int64_t)(w >> lay.bits) - 1;
                const int q = w & mask;
                if (fp.cmask[s0 + r] != (fp.cmask[c] | (1u << (qi[q] - 1)) | (1u << (qj[q] - 1))))
                    return fail(EGGMAN_EPLAN, "element mask of state %lld disagrees with its transition %d",
                                (long long)(s0 + r), v);
            }
    }
    int64_t ne = 0, nt = 0, rows = 0;
    for (int L = 2; L <= K / 2 + 1; L++) {
        const FuseLevel &f = fp.lv[L];
        const int64_t pfirst = lay.levels[L - 1] - 1;
        if (f.cfirst != lay.levels[L - 2] - 1 || f.nch != lay.nst[L - 2] || f.plo != lay.lo[L - 1] ||
            f.S < 1 || f.nblk != (f.nch + kFuseT - 1) / kFuseT)
            return fail(EGGMAN_EPLAN, "level %d: fused layout disagrees with the plan", L);
        // each child's sorted list: phases p = plo .. min(c) - 1, each the n_p parents c + {p, j}
        for (int64_t c = 0; c < f.nch; c++) {
            const uint32_t cm = fp.cmask[f.cfirst + c];
            const int32_t *lst = fp.fent.data() + f.ent0 + fp.fco[f.cfirst + c];
            const int64_t len = rp.coff[f.cfirst + c + 1] - rp.coff[f.cfirst + c];
            int64_t off = 0;
            for (int p = f.plo; p <= f.phi; p++) {
                if ((cm & ((1u << p) - 1u)) != 0) break;
                const int np = K - p - f.csize;
                int lastj = p;
                for (int u = 0; u < np; u++, ne++) {
                    const int32_t e = lst[off + u];
                    const int q = e & mask;
                    if (off + u >= len || qi[q] != p || qj[q] <= lastj || ((cm >> (qj[q] - 1)) & 1u) ||
                        fp.cmask[pfirst + (e >> lay.bits)] != (cm | (1u << (p - 1)) | (1u << (qj[q] - 1))))
                        return fail(EGGMAN_EPLAN, "level %d child %lld: phase %d entry %d is wrong", L,
                                    (long long)c, p, u);
                    lastj = qj[q];
                }
                off += np;
            }
            if (off != len)
                return fail(EGGMAN_EPLAN, "level %d child %lld: list of %lld entries, phases cover %lld", L,
                            (long long)c, (long long)len, (long long)off);
        }
        // term lists in (block, rank) order, each term a transition of its pair and stripe
        const int64_t V = (int64_t)f.S * kFuseWaves;
        for (int p = f.plo; p <= f.phi; p++) {
            if (fp.prow0[f.prow + (p - f.plo)] != rows)
                return fail(EGGMAN_EPLAN, "level %d phase %d: partial rows start at %lld, expected %lld", L,
                            p, (long long)fp.prow0[f.prow + (p - f.plo)], (long long)rows);
            rows += (int64_t)f.S * (K - p);
            for (int u = 0; u < f.nslot; u++)
                for (int64_t v = 0; v < V; v++) {
                    const int64_t li = f.toff + ((int64_t)(p - f.plo) * f.nslot + u) * V + v;
                    int64_t lastc = -1;
                    for (int64_t t = fp.toff[li]; t < fp.toff[li + 1]; t++, nt++) {
                        const int32_t c = fp.terms[2 * t], sr = fp.terms[2 * t + 1];
                        const int j = p + 1 + u;
                        const uint32_t cm = (c >= 0 && c < f.nch) ? fp.cmask[f.cfirst + c] : 0xffffffffu;
                        if (c <= lastc || c >= f.nch || (c / kFuseT) % V != v || (cm & ((1u << p) - 1u)) ||
                            j > K || ((cm >> (j - 1)) & 1u) || sr < 0 || sr >= lay.nst[L - 1] ||
                            fp.cmask[pfirst + sr] != (cm | (1u << (p - 1)) | (1u << (j - 1))))
                            return fail(EGGMAN_EPLAN, "level %d phase %d slot %d stripe %lld: term %lld is wrong",
                                        L, p, u, (long long)v, (long long)t);
                        lastc = c;
                    }
                }
        }
    }
    if (ne != lay.ntrans || nt != lay.ntrans || rows != fp.part_rows)
        return fail(EGGMAN_EPLAN, "fused tables hold %lld entries and %lld terms for %lld transitions, "
                    "%lld partial rows for %lld", (long long)ne, (long long)nt, (long long)lay.ntrans,
                    (long long)rows, (long long)fp.part_rows);
    // every task's rows lie in the partial buffer
    for (const ParTask &tk : fp.tasks)
        if (tk.row0 < 0 || tk.row0 + (K - tk.p) > fp.part_rows)
            return fail(EGGMAN_EPLAN, "a partial-row task lies outside the buffer");
    if (entries) *entries = ne;
    if (terms) *terms = nt;
    return EGGMAN_OK;
}

}  // namespace eggman

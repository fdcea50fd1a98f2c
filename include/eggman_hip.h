/*
 * eggman_hip.h -- C ABI of the MI355X (gfx950) engine for TheEggman.jl's subset-DP hafnian path.
 *
 * The library replaces the device hook `_haf_dp_backend(backend, P, plan; max_batch)`
 * (src/dp.jl:514-531, implemented for KernelAbstractions at
 * ext/TheEggmanKernelAbstractionsExt.jl:90-146) together with the device caches around it
 * (src/dp.jl:533-587).  Everything upstream -- validation, strategy choice, the gather into P,
 * and the plan -- stays with the caller.
 *
 * Conventions
 *   - Every function returns EGGMAN_OK (0) or a negative status; eggman_last_error() then
 *     describes the failure (thread-local).
 *   - Plan arrays keep the reference's 1-based values and packing: a transition is
 *     (child << bits) | pair with bits = 9 for 4-byte and 12 for 8-byte words
 *     (src/dp.jl:150-180).
 *   - P is laid out [pair][batch] with the batch index fastest: element (b, q) of Julia's
 *     column-major Matrix{T}(B, npairs) from _gather_pairs_batch (src/dp.jl:464-512) is at
 *     q*B + b (0-based).  Complex values are interleaved (re, im).
 *   - Host-pointer entry points are synchronous on return; device-pointer entry points are
 *     asynchronous on the given HIP stream (NULL = the library's own stream for that device).
 *   - All calls are thread-safe: one mutex serialises use of the library's device buffers, as
 *     the reference's _DEVICE_PLAN_LOCK does (ext/TheEggmanKernelAbstractionsExt.jl:117-119).
 */
#ifndef EGGMAN_HIP_H
#define EGGMAN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define EGGMAN_OK 0
#define EGGMAN_EINVAL (-1)       /* bad argument (maps to ArgumentError / DimensionMismatch) */
#define EGGMAN_ENOMEM (-2)       /* host or device allocation failed, or would not fit */
#define EGGMAN_EPLAN (-3)        /* plan cannot be represented, or is inconsistent */
#define EGGMAN_EHIP (-4)         /* a HIP runtime call failed */
#define EGGMAN_EUNSUPPORTED (-5) /* valid request this build does not handle */
#define EGGMAN_EOVERFLOW (-6)    /* EGGMAN_I128 request whose exact result may not fit in 127 bits
                                    (maps to OverflowError) */

/* element types (the T of Matrix{T}) */
#define EGGMAN_F32 1  /* Float32 */
#define EGGMAN_F64 2  /* Float64 */
#define EGGMAN_C32 3  /* ComplexF32, interleaved */
#define EGGMAN_C64 4  /* ComplexF64, interleaved */
#define EGGMAN_I128 5 /* exact: P holds int64 entries; results are 128-bit two's-complement
                         integers (lo, hi uint64 limbs).  Every entry point refuses
                         (EGGMAN_EOVERFLOW, before any walk) an instance whose hafnian is not
                         provably below 2^127 in magnitude: the bound is the smaller of
                         (K-1)!! * m^(K/2) and sqrt(prod_i max(r_i, 1)), with m = max(1, max |p|)
                         and r_i = sum_{j != i} |p_ij| (haf^2 <= perm <= prod r_i for |A|); it
                         also bounds every (K-2)-minor, so gradients are covered.  Device-pointer
                         entry points check on the device and synchronise the stream for it. */

/* Version string of the library. */
const char *eggman_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *eggman_last_error(void);

/* Number of visible HIP devices (0 when none; never an error). */
int eggman_device_count(void);

/* ---------------------------------------------------------------- plan (host side) */

/* Closed-form plan size for degree K: replaces _dp_counts (src/dp.jl:192-208). */
int eggman_dp_counts(int K, int64_t *nstates, int64_t *ntrans);

/*
 * Build the degree-K plan on the host into caller arrays of int32 (index_bytes 4) or int64 (8):
 * starts[nstates+1], trans[ntrans], levels[K/2+2].  Same values, order and packing as
 * _build_dp_plan (src/dp.jl:223-304), including its refusals (K > 64, packing overflow) --
 * computed by combinatorial ranking instead of a set and a dictionary, on `nthreads` host
 * threads (0 = all).
 */
int eggman_plan_build(int K, int index_bytes, void *starts, void *trans, void *levels,
                      int nthreads);

/* ---------------------------------------------------------------- plan (device side) */

/*
 * Upload a caller-built plan (borrowed; read only during the call) to the current device and
 * return a handle.  Replaces _device_plan (ext/TheEggmanKernelAbstractionsExt.jl:31-42): the
 * device copy is cached per (device, K, index_bytes), so uploading the same degree again returns
 * the cached handle without copying.  Rejects (EGGMAN_EPLAN) arrays that are not a valid plan
 * layout (sizes, levels, uniform out-degree per level).
 */
int eggman_plan_upload(int K, int index_bytes, const void *starts, int64_t nstarts,
                       const void *trans, int64_t ntrans, const void *levels, int64_t nlevels,
                       uint64_t *handle);

/*
 * Generate the degree-K plan directly in device memory (no host build, no upload) and return a
 * handle; bit-identical to the uploaded plan.  Cached like eggman_plan_upload.  New: the
 * reference has no device-side plan generation.
 */
int eggman_plan_generate(int K, int index_bytes, uint64_t *handle);

/*
 * Check caller-supplied plan arrays without touching a device: the sizes and level boundaries
 * the closed form gives, one out-degree per level, and every transition pointing into the level
 * below with a pair index in 1..npairs.  eggman_plan_upload runs the same check; a plan that
 * fails it is never walked.  Mirrors the invariants of test/runtests.jl:241-267.
 */
int eggman_plan_validate(int K, int index_bytes, const void *starts, int64_t nstarts,
                         const void *trans, int64_t ntrans, const void *levels, int64_t nlevels);

/* Degree, index width, state and transition counts of a plan handle. */
int eggman_plan_info(uint64_t handle, int *K, int *index_bytes, int64_t *nstates,
                     int64_t *ntrans);

/* Copy a device plan's transitions back to the host (ntrans words), for verification. */
int eggman_plan_download(uint64_t handle, void *trans_host, int64_t ntrans);

/* ---------------------------------------------------------------- forward walk */

/*
 * Hafnians of the B instances in P_host ([pair][batch], npairs = K(K-1)/2 rows) into out_host
 * (B elements; 16 bytes each for EGGMAN_I128).  Replaces
 * _haf_dp_backend(backend, P, plan; max_batch) (ext/TheEggmanKernelAbstractionsExt.jl:90-146):
 * batches are processed in chunks of max_batch instances (<= 0: the whole batch); the result is
 * bit-identical for every chunking, and every state's sum is accumulated in plan order without
 * fused multiply-add, exactly as _dp_states! (src/dp.jl:344-363) does.  B == 0 is a no-op.
 */
int eggman_dp_forward(uint64_t plan, int dtype, const void *P_host, int64_t B, int64_t npairs,
                      int64_t max_batch, void *out_host);

/*
 * Device-resident form of eggman_dp_forward: P_dev is [pair][ldp] in device memory (instance b
 * of pair q at q*ldp + b, ldp >= B), out_dev receives B results.  Asynchronous on `stream`
 * (a hipStream_t; NULL = the library's stream).  No host transfer happens.  New: the
 * reference only exposes the host form.  A call of four or more passes runs them two at a time,
 * on `stream` and on a side stream of the library's that waits for `stream` where the call's
 * passes begin and that `stream` waits for before the call's last launches: the call stays
 * ordered on `stream` like a single launch sequence.
 */
int eggman_dp_forward_device(uint64_t plan, int dtype, const void *P_dev, int64_t ldp,
                             int64_t B, void *out_dev, void *stream);

/*
 * eggman_dp_forward_device without any allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The caller passes P and out as for
 * eggman_dp_forward_device; the state levels live in one static device array of 128 MiB per GPU
 * (held from the library's first use of that GPU until the process ends), and the call makes no
 * hipMalloc/hipFree, captures no graph and, after its first call on a device (which looks up the
 * array's address), creates nothing.  Pass sizes are fitted to the array (the values are the same
 * bits as eggman_dp_forward_device); a degree whose two largest levels of one instance do not fit
 * is refused with EGGMAN_ENOMEM (e.g. complex N >= 36), and EGGMAN_I128 with EGGMAN_EUNSUPPORTED.
 * Every call on one GPU shares the array and leaves its contents behind: the library orders its
 * own calls, but work that overlaps outside that order on one GPU (replays of graphs captured from
 * these calls on different streams) races.  Asynchronous on `stream` (NULL = the library's).
 */
int eggman_dp_forward_device_noalloc(uint64_t plan, int dtype, const void *P_dev, int64_t ldp,
                                     int64_t B, void *out_dev, void *stream);

/*
 * The boson-sampling batch shape, hafnian_repeated(A, rpts; backend) (src/hafnian.jl:1317-1359):
 * one matrix A (N x N, column-major as Julia stores it; int64 entries for EGGMAN_I128) and B index
 * lists idx_host[b][0..K-1] (int32, 1-based rows of A, repeats allowed, K = the plan's degree).
 * A and the lists are copied to the device and P[q][b] = A[idx_b[i], idx_b[j]] (pair q = (i < j))
 * is gathered there, replacing the host loop of _gather_pairs_batch (src/dp.jl:479-494) and the
 * B x npairs host-to-device copy.  Results as eggman_dp_forward, bit-identical to it.
 */
int eggman_dp_forward_repeated(uint64_t plan, int dtype, const void *A_host, int64_t N,
                               const int32_t *idx_host, int64_t B, int64_t max_batch,
                               void *out_host);

/*
 * Hafnians of instances of different degrees in one device pass.  Group g has counts[g]
 * instances sharing plan handle plans[g]; P_host is the concatenation of the groups'
 * [pair][counts[g]] blocks (group 0 first) and out_host receives the groups' results in the same
 * order.  Groups of 256 or more instances, which fill the GPU by themselves, each walk with their
 * own tuned launch on a stream of their own, concurrently with one another and with the shared
 * walk of the smaller groups, in which every launch advances each of them by one level.  All
 * plans must live on one device and use one index width.  The call is processed in chunks that
 * take the same share of every group: enough to keep the staging within the free device memory,
 * with at most max_batch instances of any one group in a chunk (<= 0: no limit).  Results are bit-identical to
 * eggman_dp_forward on each group alone, for every chunking.  New: the reference requires one
 * size per batch (src/hafnian.jl:1277-1280).
 */
int eggman_dp_forward_mixed(int ngroups, const uint64_t *plans, const int64_t *counts, int dtype,
                            const void *P_host, int64_t max_batch, void *out_host);

/* Device-resident form of eggman_dp_forward_mixed, asynchronous on `stream`. */
int eggman_dp_forward_mixed_device(int ngroups, const uint64_t *plans, const int64_t *counts,
                                   int dtype, const void *P_dev, void *out_dev, void *stream);

/* ---------------------------------------------------------------- reverse sweep */

/*
 * Hafnians and their gradients with respect to the pair values: out_host as for
 * eggman_dp_forward, dP_host ([pair][batch], same layout and element size as the results)
 * receives d haf_b / d P[q, b] for every pair q.  New: the reference has no derivative.  The
 * caller maps dP back to the matrix through its gather (src/dp.jl:464-512): with distinct
 * indices, d haf / d a_ij = dP[pair_index(i, j)] for the symmetric pair variable a_ij, which
 * equals the hafnian of A with rows and columns i and j removed.  The sums are deterministic
 * (fixed order, no atomics).  Instances are processed in chunks of at most max_batch (<= 0:
 * as many as fit in device memory).
 */
int eggman_dp_forward_grad(uint64_t plan, int dtype, const void *P_host, int64_t B,
                           int64_t npairs, int64_t max_batch, void *out_host, void *dP_host);

/* Device-resident form of eggman_dp_forward_grad: dP_dev rows are lddp elements apart.
 * Asynchronous on `stream` (NULL = the library's stream); passes may run two at a time on
 * `stream` and a side stream, joined back into `stream` as in eggman_dp_forward_device. */
int eggman_dp_forward_grad_device(uint64_t plan, int dtype, const void *P_dev, int64_t ldp,
                                  int64_t B, void *out_dev, void *dP_dev, int64_t lddp,
                                  void *stream);

/* ---------------------------------------------------------------- autotuner */

/*
 * Time the launch choices for (plan, dtype, B) -- lanes per state, LDS staging of the pair rows,
 * the level kernel (64-instance rows, 16-instance tiles or 16-byte loads), waves per workgroup,
 * workgroups per CU, and the instances per pass over the levels -- on synthetic pair values,
 * `reps` runs each, and keep the fastest for this (device, degree, layout, dtype, B rounded up
 * to a power of two).  Later calls of the forward entry points use it; a batch size with no entry
 * of its own uses the nearest tuned bucket of the same degree, layout and dtype (the largest
 * below it, else the smallest above); a degree or dtype with no entry at all uses defaults that
 * follow the tuned FP64 choices (the fused LDS walk where it fits, else from 64 instances the
 * 16-byte-load kernel in passes of about 2^26 state values).  Every choice gives bit-identical
 * results.  A text report
 * (one line per candidate) is written to `report`.
 * New: the reference has no tuning.
 */
int eggman_autotune(uint64_t plan, int dtype, int64_t B, int reps, char *report,
                    int64_t report_len);

/* Tuning flags */
#define EGGMAN_TUNE_LDS 1   /* stage the level's pair rows in LDS (64 lanes per state) */
#define EGGMAN_TUNE_GRAPH 2 /* replay the level sequence from a HIP graph captured on first use,
                               reused while the same buffers are passed */
#define EGGMAN_TUNE_FUSED 4 /* one workgroup per instance walks every level with all state
                               values in LDS (degrees whose state array fits) */
#define EGGMAN_TUNE_WIDE 8  /* 16 waves per workgroup in the LDS kernel instead of 8 */
#define EGGMAN_TUNE_CKPT 16 /* gradient: keep checkpoint levels only and recompute the others
                               (chosen by itself when keeping every level would not fit a
                               full pass in device memory) */
#define EGGMAN_TUNE_NARROW 32 /* LDS kernel with 16-instance tiles and 4 states per wave: a
                                 child row of a tile is one cache line, so L2 holds 4x more */
#define EGGMAN_TUNE_VEC 64    /* LDS kernel with 16-byte loads: 16 / sizeof(T) instances per
                                 lane, 8 lanes per state, 8 states per wave (level rows padded
                                 to whole 128-byte tile rows) */
#define EGGMAN_TUNE_DEEP 128  /* with EGGMAN_TUNE_VEC: child loads in batches of 6 transitions
                                 instead of 4 */
#define EGGMAN_TUNE_PAIRS 256 /* with EGGMAN_TUNE_VEC: an 8-lane slot takes two parents {p} u T,
                                 {p+1} u T that share their children, loading them once (a
                                 quarter fewer child and plan-word loads); the per-level item
                                 list is built on first use */

/*
 * Set / read a tuning entry directly (e.g. to persist tuned choices across processes).  bt:
 * lanes per state (1..64, power of two); chunk: instances per pass over the levels; flags:
 * EGGMAN_TUNE_*; occ: workgroups per CU for the LDS kernel (0 = as many as LDS allows).
 */
int eggman_tune_set(int K, int index_bytes, int dtype, int64_t B, int bt, int64_t chunk,
                    int flags, int occ);
int eggman_tune_get(int K, int index_bytes, int dtype, int64_t B, int *bt, int64_t *chunk,
                    int *flags, int *occ);

/*
 * The range check of the exact walk, host only: log2 of the bound described at EGGMAN_I128 for
 * each of the B instances of P (int64, [pair][ldp], npairs = K(K-1)/2 rows) into log2_bound
 * (B doubles; may be NULL).  Returns EGGMAN_EOVERFLOW if an instance's bound reaches 2^127,
 * which every EGGMAN_I128 entry point refuses.  New: the reference has no exact path.
 */
int eggman_exact_bound(int K, const int64_t *P, int64_t ldp, int64_t B, double *log2_bound);

/*
 * Measurement helper (no counterpart in the reference): the rate, in 128-byte lines per second,
 * at which the current device serves random reads of whole lines -- each read the way the
 * 16-byte-load level kernels read a child row, 8 lanes x 16 bytes -- from a zeroed working set of
 * working_set_bytes (sized like a pass's level buffers, so served by L2, MALL or HBM as the
 * walk's child reads are).  The practical ceiling of the level kernels' gathers, reported by the
 * benchmark beside the device-copy rate.  Allocates and frees its working set.
 */
int eggman_gather_rate(int64_t working_set_bytes, double *lines_per_s);

/* ---------------------------------------------------------------- caches */

/* Device bytes held by cached plans and work buffers: the library's share of gpu_cache_bytes
 * (src/dp.jl:552-560). */
int64_t eggman_cache_bytes(void);

/* Release every cached plan and work buffer: empty_gpu_cache! (src/dp.jl:567-573).  Existing
 * handles become invalid. */
int eggman_cache_clear(void);

/* Device memory one batch needs besides the shared plan, B*(nstates+npairs)*sizeof(T):
 * dp_batch_bytes (src/dp.jl:584-587). */
int64_t eggman_batch_bytes(int K, int dtype, int64_t B);

/*
 * Which forward levels the gradient keeps resident for degree K (host-only query).  Level L
 * (1 .. K/2) holds the values the reverse step at L+1 reads.  checkpointed = 0: every level
 * (kept_rows = nstates - 1, seg_rows = 0).  checkpointed = 1: level 1 and checkpoint levels,
 * chosen to minimise kept_rows + seg_rows, where the other levels are recomputed segment by
 * segment in a region of seg_rows state rows.  kept[L] = 1 for a kept level (kept[0] unused,
 * nkept >= K/2 + 1).  Rows count state values per instance.  New: the reference has no
 * derivative; the reverse sweep recomputes levels from checkpoints above a memory threshold
 * (EGGMAN_TUNE_CKPT forces it).
 */
int eggman_grad_levels(int K, int checkpointed, int64_t *kept_rows, int64_t *seg_rows,
                       uint8_t *kept, int64_t nkept);

/*
 * Builds, on the host, the index tables of the gradient's fused reverse step for degree K
 * (Int32 layout, 4 <= K <= 32) from the host plan and checks them against it (host-only; no
 * GPU needed): every child's parent list, sorted by pair, holds for each phase p < min(child)
 * exactly the parents child + {p, j}, j ascending, j not in the child; every stripe term list
 * holds its (phase, slot, wave-stripe)'s terms in (block, rank) order; the partial-row table
 * covers every (level, phase, stripe).  Counts the list entries and terms it checked.  Returns
 * EGGMAN_OK, or EGGMAN_EPLAN naming the first inconsistency.  New: the reference has no
 * derivative; this is the self-check of the tables behind eggman_dp_forward_grad.
 */
int eggman_fuse_validate(int K, int64_t *entries_checked, int64_t *terms_checked);

#ifdef __cplusplus
}
#endif

#endif /* EGGMAN_HIP_H */

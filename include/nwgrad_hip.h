/* nwgrad_hip.h -- the C ABI of the MI355X (gfx950) alignment hot path.
 *
 * A plain-C boundary (no HIP or torch types in any signature): device-resident batches of
 * encoded sequence pairs, the hard / score-only / soft passes over them, and the
 * reductions of their per-pair gradients.  The header-only C++ engine
 * (nwgrad_amd/cpp/nwgrad_amd/) and the Python module sit above it; nothing below it runs
 * on the CPU except argument checks and the batch planner.
 *
 * Reference interface replaced (paths relative to the reference's src/nwgrad/cpp/nwgrad/):
 * the reference's kernels are noexcept functions taking one plain-data job struct each,
 * called through LevelKernels function pointers (simd_levels.hpp:1291-1393): ViterbiJob
 * (:188-219), ScoreJob (:253-267), InterJobT (:497-530), InterSoftJobT (:542-562),
 * SoftPairJobT (:581-627).  Their callers are BatchEngine's group and pair runners
 * (batch_engine.hpp:2446-2456, 997-1012; aligner.hpp:3427-3465, 6946-6956).  This ABI
 * moves the boundary up one level, to the batch, so that pairs and results stay in HBM
 * between calls: each entry below names the BatchEngine / SeqPairBatchT member whose work
 * it does.
 *
 * Conventions.
 *  - Every call returns NWG_OK (0) or an NWG_E* status; nwg_last_error() describes the
 *    last failure of the context.  The C++ engine turns statuses into the reference's
 *    exceptions (batch: ValueError for NWG_EINVAL and NWG_ERANGE, RuntimeError for
 *    NWG_ESTATE, IndexError for NWG_EINDEX).
 *  - Host pointers unless a name ends in _dev.  Calls that take host output buffers
 *    synchronise the context's stream before returning; *_dev calls are asynchronous on it.
 *  - One caller per context at a time (the reference's batch has the same rule,
 *    batch_engine.hpp:2005-2014).
 */
#ifndef NWGRAD_HIP_H
#define NWGRAD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NWG_OK 0
#define NWG_EINVAL 1   /* invalid argument (std::invalid_argument) */
#define NWG_ESTATE 2   /* results requested before they exist (std::logic_error) */
#define NWG_ERANGE 3   /* soft range failure (std::domain_error) */
#define NWG_EINDEX 4   /* index out of range (std::out_of_range) */
#define NWG_EDEVICE 5  /* HIP runtime failure */
#define NWG_ENOMEM 6   /* device or host allocation failure */

typedef struct nwg_ctx nwg_ctx;
typedef struct nwg_batch nwg_batch;

/* Problem type of a batch, fixed at creation (SeqPairBatchT ctor, seq_pair_batch.hpp:73-78). */
typedef struct {
    int32_t precision; /* 0: float32 Viterbi/soft (SeqPairBatch), 1: float64 (SeqPairBatchDouble) */
    int32_t gap_model; /* 0 linear, 1 affine */
    int32_t mode;      /* 0 global, 1 local, 2 glocal */
    int32_t grad_mode; /* 0 none (score only), 1 hard, 2 soft */
} nwg_problem;

/* Context on one device.  Replaces the reference's thread-pool setup
 * (parallel.hpp:165-243): the GPU is the pool. */
int nwg_init(int device, nwg_ctx **out);
void nwg_destroy(nwg_ctx *ctx);
const char *nwg_last_error(const nwg_ctx *ctx);
/* Work of this context goes to `hip_stream` (a hipStream_t; NULL: the null stream). */
int nwg_set_stream(nwg_ctx *ctx, void *hip_stream);
/* Give the context a stream of its own (non-blocking, default priority; destroyed with the
 * context), so that batches of different contexts run concurrently.  No counterpart in the
 * reference, whose calls are synchronous; the serving loop runs one context per slot. */
int nwg_ctx_own_stream(nwg_ctx *ctx);

/* Upload N pairs (codes in [0, n_alpha), offsets of N+1 entries) and plan them into
 * tiles.  Replaces BatchEngine's constructor and planner (batch_engine.hpp:284-350,
 * 2610-2676): codes stay device-resident for the life of the batch. */
int nwg_batch_create(nwg_ctx *ctx, const nwg_problem *prob, int n_alpha, int64_t n_pairs,
                     const uint8_t *codes_a, const int64_t *off_a, const uint8_t *codes_b,
                     const int64_t *off_b, nwg_batch **out);
void nwg_batch_free(nwg_batch *b);
/* Replace the batch's pairs with N new ones of the same problem type and alphabet, keeping
 * its device buffers (grown when too small, never shrunk) and its params.  Uploads go to the
 * context's stream; the call returns once the new plan is on the device, and leaves
 * results invalid.  The batch must be idle (nwg_batch_query).  No counterpart in the
 * reference (a batch there is built once); the serving loop reloads its slots so that a
 * launch allocates and frees no device memory once the slots have grown. */
int nwg_batch_reload(nwg_batch *b, int64_t n_pairs, const uint8_t *codes_a, const int64_t *off_a,
                     const uint8_t *codes_b, const int64_t *off_b);
/* *done = 1 when all work enqueued for the batch on its context's stream has finished, else 0
 * (hipStreamQuery; never blocks).  No counterpart in the reference (synchronous calls). */
int nwg_batch_query(nwg_batch *b, int *done);

/* Params: matrix n_alpha*n_alpha row-major, gaps4 = go_a, ge_a, go_b, ge_b.  Cast to the
 * batch's precision on upload as static_cast<T>(double).  Invalidates cached results.
 * Replaces BatchEngine::set_params (batch_engine.hpp:597-606). */
int nwg_batch_set_params(nwg_batch *b, const double *matrix, const double *gaps4);
/* K parameter sets: `matrices` K blocks of n_alpha*n_alpha, `gaps` K rows of go_a, ge_a,
 * go_b, ge_b.  K > 1 makes the batch tracked (default tracks: every residue set 0) and the
 * per-pair gradient rows K * (n_alpha^2 + 4) doubles: K matrices, then K gap rows.  Hard and
 * score-only batches.  Replaces BatchEngine::set_params with an AlignParams of K sets
 * (batch_engine.hpp:597-606; align_params.hpp:55-97). */
int nwg_batch_set_params_k(nwg_batch *b, int K, const double *matrices, const double *gaps);
/* Tracks (per-row parameter sets): `tracks` one set per residue of A, pair p at
 * off_a[p] .. off_a[p+1]; `gap_tracks` one per DP row 0..m, pair p at off_a[p] + p (m + 1
 * entries).  NULL: the reference's defaults (set 0; a gap track following the track).  Sets
 * must lie in [0, K).  Row i scores with matrix tracks[i-1] and pays gaps in B from that set,
 * gaps in A from set gap_tracks[i]; each gradient count lands in its row's set.  Hard and
 * score-only batches.  Replaces BatchEngine's tracks / gap_tracks arguments
 * (batch_engine.hpp:284-350, 1467-1500; aligner.hpp:1800-2008). */
int nwg_batch_set_tracks(nwg_batch *b, const uint16_t *tracks, const uint16_t *gap_tracks);
/* The batch's parameter-set count and whether it is tracked. */
int nwg_batch_num_sets(nwg_batch *b, int *K, int *tracked);
/* The same from a device buffer of n_alpha*n_alpha + 4 doubles (matrix, then go_a, ge_a,
 * go_b, ge_b), copied on the context's stream without synchronising: the parameter update of
 * a training step stays on the device (and inside a captured graph).  Soft batches only; no
 * counterpart in the reference, whose parameters live on the host. */
int nwg_batch_set_params_dev(nwg_batch *b, const double *theta_dev);

/* fill 0 (default): the batch's tiles in one launch, most work first -- short hard / score
 * tiles one per block in dispatch order, soft and long-pair work pulled from device queues by
 * a persistent grid; fill 1: plain per-bucket launches.  Results do not depend on it
 * (the reference's `fill=` knob, py_exports.cpp:1180-1207). */
int nwg_batch_set_fill(nwg_batch *b, int fill);

/* Columns a lane of the hard and score-only tiles: 0 (default) chosen per batch -- the
 * lane-width table's matching rule, else the built-in rule (8 when the tiles at 16 number
 * fewer than two per SIMD and every pair fits a wave at 8, else 16) -- or 8 or 16 forced (16
 * when a pair is wider than 512 columns).  The lane-group width follows: the power of two of
 * ceil(len(B) / columns) lanes, so 64 / width pairs a wave.  Results do not depend on it.
 * The reference's counterpart is its runtime gates and tuning table (simd_levels.hpp:704-751,
 * _tuning.py), which choose its inter-pair group width per ISA.  nwg_batch_lanes gives the
 * columns the batch's tiles are cut at (soft batches: the soft pass's 8). */
int nwg_batch_set_lanes(nwg_batch *b, int columns);
int nwg_batch_lanes(nwg_batch *b, int *columns);
/* The process-wide lane-width table: n_rules rules of three int32 (n_max, tiles16_max,
 * columns), first match wins -- a batch whose median len(B) is at most n_max and whose tiles
 * at 16 columns a lane number at most tiles16_max / 16 per CU is cut at `columns` (8 or 16).
 * n_rules 0 clears it.  Batches planned afterwards follow it (the tuner in
 * nwgrad_amd/tuning.py measures and writes it). */
int nwg_set_lane_table(int n_rules, const int32_t *rules);

/* Soft settings (py_exports.cpp soft_temperature / soft_precision).  precision 0: float32
 * scaled pass (float32 batches only), 1: float64 scaled pass. */
int nwg_batch_set_soft(nwg_batch *b, double temperature, int precision);

/* The full pass: score, and gradient unless grad_mode is none.  Per-pair results stay on
 * the device; *total_out (may be NULL) receives the sum of scores in pair order.
 * Replaces BatchEngine::score_and_grad / soft_score_and_grad (batch_engine.hpp:615-629,
 * batch_soft_defer.inl:167-240). */
int nwg_batch_run(nwg_batch *b, double *total_out);
/* Enqueue the pass on the context's stream and return (no total, no synchronisation). */
int nwg_batch_run_async(nwg_batch *b);
/* After nwg_batch_run_async of a soft batch: synchronise, rerun in double the pairs the
 * float32 pass refused, and fail as nwg_batch_run would if any remain (the reference's
 * float32 -> double -> error policy, aligner.hpp:6554-6586). */
int nwg_batch_soft_finish(nwg_batch *b);

/* Banded re-alignment around each pair's guide from the last pass (the training-loop step after
 * a set_params): the pass of nwg_batch_run with only the cells of each row's band
 * [max(1, g[i] - bw), max(g[i] + bw, g[i+1] - 1 + bw)] (Global's row m to n; bw = min(bandwidth,
 * n)), the others -inf -- the reference's GuideBanded fill, bit for bit.  The pass writes new
 * guides.  Hard batches; NWG_ESTATE before any pass, NWG_EINVAL for bandwidth < 0 (a width of 0
 * follows the guide's own cells, as streaming align() does with guides and band 0).  Replaces
 * BatchEngine::banded_grad (batch_engine.hpp:633-653, 2853-2982; aligner.hpp:2621-2730,
 * 5991-6098). */
int nwg_batch_run_banded(nwg_batch *b, int bandwidth, double *total_out);

/* Guides of the last hard pass: per pair the column after each row of A (m + 1 entries, pair p's
 * at off_a[p] + p, sum(m) + N in all), interpolated where a Local path leaves rows uncovered --
 * the reference's guide_j (aligner.hpp:3155-3160, 2734-2749; store_guide_, batch_engine.hpp:
 * 2045-2055).  They survive set_params.  Hard batches; NWG_ESTATE before any pass. */
int nwg_batch_guides(nwg_batch *b, int32_t *out);
int nwg_batch_pair_guide(nwg_batch *b, int64_t i, int32_t *out);
/* Guides from the caller (the same layout), for a banded pass around given alignments; checked
 * as the reference checks a supplied guide (validate_guide, aligner.hpp:2622-2642: every entry
 * in [0, n], never decreasing).  Replaces the guides of streaming align(band, aligned_a,
 * aligned_b) (batch_engine.hpp:748-1273). */
int nwg_batch_set_guides(nwg_batch *b, const int32_t *guides);

/* Stored paths (keep_paths): keep != 0 makes the following hard passes store each pair's path
 * as run-length ops, (type << 30) | length with type 0 both advance, 1 a gap in B, 2 a gap in A
 * (batch_engine.hpp:1415-1418, 2056-2092).  nwg_batch_paths copies them out: `ops` holds
 * sum(m + n) entries, pair p's buffer of m + n at off_a[p] + off_b[p] with its n_ops[p] ops
 * at the buffer's end, ends[2p], ends[2p+1] the cell the path ends at (the reference's
 * alignment_end).  set_params and nwg_batch_drop_paths invalidate them (NWG_ESTATE after). */
int nwg_batch_set_keep_paths(nwg_batch *b, int keep);
int nwg_batch_paths(nwg_batch *b, uint32_t *ops, int32_t *n_ops, int32_t *ends);
int nwg_batch_drop_paths(nwg_batch *b);

/* Copies of the cached results (BatchEngine::scores / grads_into,
 * batch_engine.hpp:689-695, 722-746).  mats: N*n_alpha*n_alpha, gaps: N*4. */
int nwg_batch_scores(nwg_batch *b, double *out);
int nwg_batch_grads(nwg_batch *b, double *mats, double *gaps);

/* sum_i w_i * grad_i in 4096-pair blocks (each summed in pair order, then the block sums
 * in block order): the reference's blocked_sum_ (batch_engine.hpp:702-718, 1788-1827),
 * bit for bit.  w NULL: compute_grad.  out: n_alpha*n_alpha + 4 doubles. */
int nwg_batch_weighted_grad(nwg_batch *b, const double *w, int64_t n, double *out);
/* The same on device buffers, asynchronous on the context's stream (for autograd). */
int nwg_batch_weighted_grad_dev(nwg_batch *b, const double *w_dev, double *out_dev);
/* Pair i's cached score and (grad non-NULL) gradient row, K * (n_alpha^2 + 4) doubles: K
 * matrices, then K rows of go_a, ge_a, go_b, ge_b.  NWG_EINDEX outside [0, N).  Replaces the
 * per-pair accessors behind SeqPairBatch.__getitem__ views (py_exports.cpp:895-905;
 * seq_pair.hpp score() / grad()). */
int nwg_batch_pair_result(nwg_batch *b, int64_t i, double *score, double *grad);
/* Device pointer to the per-pair scores (N doubles), valid until the batch is freed. */
int nwg_batch_scores_dev(nwg_batch *b, const double **out_dev);
/* Device-to-device copy of the per-pair scores into dst_dev (N doubles), on the context's
 * stream: a caller-owned copy that the next pass does not overwrite. */
int nwg_batch_scores_copy_dev(nwg_batch *b, double *dst_dev);

/* Introspection for tests and the bench: number of tiles and the lane-group width used. */
int nwg_batch_plan_info(nwg_batch *b, int64_t *n_tiles, int64_t *n_cells);
/* Device-time of the last pass's kernels in milliseconds (HIP events on the context's
 * stream around the kernels only). */
int nwg_batch_last_kernel_ms(nwg_batch *b, double *ms);
/* Pairs of the last soft pass that the float32 pass refused (lost range or mass beyond
 * 2^-20) and that were rerun in double -- the reference's kF32Refused fallback
 * (aligner.hpp:6554-6572, batch_engine.hpp:1285-1306). */
int nwg_batch_soft_reruns(nwg_batch *b, int64_t *n);
/* The float32 refusals of a soft batch's passes, as its asynchronous (or graph-captured)
 * passes leave them: pair-passes refused, the first pass (1-based) that refused any (0: none)
 * and passes run.  Until nwg_batch_soft_finish reruns them, a refused pair's gradient row is
 * left out of nwg_batch_weighted_grad_dev, so a training step never applies it; this record
 * says whether that happened.  reset: clear the refusal counts (the pass counter keeps
 * counting).  No counterpart in the reference, whose calls rerun refusals before returning
 * (aligner.hpp:6554-6586; batch_engine.hpp:1285-1306 kF32Refused). */
int nwg_batch_soft_refusals(nwg_batch *b, int64_t *pair_passes, int64_t *first_pass, int64_t *passes, int reset);
/* Shape of the last soft launch (diagnostics; no counterpart in the reference, whose worker
 * pool has no grid): blocks of the persistent launch (0 for plain launches), blocks that
 * drain the wide-tile queue first, and resident blocks per CU from the occupancy query. */
int nwg_batch_launch_info(nwg_batch *b, int64_t *grid, int64_t *wide_slots, int64_t *per_cu);
/* Device memory the soft pass may take for its stored forward rows, per launch (0: the
 * default, 96 GiB).  Half of it at most goes to the wide pairs' pool (never less than the
 * largest wide tile); a smaller pool means fewer wide pairs in flight, never a different
 * result.  No counterpart in the
 * reference, whose CPU pass keeps one pair's rows per worker thread (soft_kernel_impl.inl
 * row storage); it bounds the HBM a serving process spends on soft batches. */
int nwg_batch_set_soft_scratch(nwg_batch *b, int64_t bytes);
/* Traceback of the hard passes: 0 auto, 1 pointers, 2 scores, 3 hirschberg, 4 hirschberg_pmax
 * (the reference's TracebackMode, aligner.hpp:30-111, fixed at construction by
 * SeqPairBatch(traceback=), py_exports.cpp:796-831), and the Hirschberg base case in rows
 * (SeqPairBatch.hb_cutoff, seq_pair_batch.hpp:136-153).  Under 3 / 4 a hard pass takes every
 * pair of more than hb_cutoff rows through the linear-memory recursion (aligner.hpp:4118-4560),
 * Global affine untracked pairs only; the others run as Pointers, which the other values all
 * give (auto's Grid returns Pointers' path, aligner.hpp:89-100). */
int nwg_batch_set_traceback(nwg_batch *b, int traceback, int hb_cutoff);
/* The last hard pass's recursion: pairs that split, levels, base-case blocks, the device
 * scratch it held in bytes, and the cells its half sweeps and base cases filled (about twice
 * the pairs' cells).  No reference counterpart (diagnostics of the GPU path). */
int nwg_batch_hb_info(nwg_batch *b, int64_t *pairs, int64_t *levels, int64_t *base_blocks, int64_t *scratch_bytes,
                      int64_t *filled_cells);
/* Device memory the hard fill of pairs up to 1,024 columns may take for its pointer codes
 * (0: the default, 16 GiB).  Fill 0 runs one plain launch, a code slot per tile, when that
 * fits, else one persistent launch whose waves (one slot each) fit it; fill 1 splits each
 * bucket into launches that fit it (never fewer than 1,024 tiles a launch).  Results never
 * depend on it.  No counterpart in the reference, whose CPU walk keeps one pair's codes per
 * worker thread. */
int nwg_batch_set_table_budget(nwg_batch *b, int64_t bytes);
/* The hard / score-only fill of pairs up to 1,024 columns under the current fill mode and
 * table budget (diagnostics, no counterpart in the reference): its launches, whether fill 0
 * runs persistent, and the grid of the last fill if it ran persistent (0 if plain). */
int nwg_batch_hard_launches(nwg_batch *b, int64_t *launches, int64_t *persistent, int64_t *last_grid);
/* The last persistent soft launch's wide tiles (diagnostics, no counterpart in the reference):
 * how many ran pipelined, as panel items across waves (0: none), how many took a region of
 * the circular wide pool (the pipelined ones, or every wide tile when block slots for the
 * whole ones would not fit), the pool's rows (0: none), and the queue lag, in tiles, of a
 * pipelined tile's backward items behind its forward ones. */
int nwg_batch_wide_info(nwg_batch *b, int64_t *tiles, int64_t *pooled, int64_t *pool_rows, int64_t *lag);
/* How nwg_batch_run has formed this batch's totals so far (diagnostics, no counterpart in the
 * reference): from the device's sum, where every order of the additions is exact (all
 * scores multiples of 2^q with the sum of magnitudes below 2^(53+q)), or by summing the
 * scores in pair order on the host.  Both give the pair-order total. */
int nwg_batch_total_info(nwg_batch *b, int64_t *from_device, int64_t *in_pair_order);

/* A hard or score-only pass that allocates nothing: no counterpart in the reference.  It is
 * for callers that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The caller passes device buffers: the
 * pairs (codes_*_dev in [0, n_alpha), offsets off_*_dev of N+1 entries), scores_dev (N
 * doubles) and, hard, grads_dev (N rows of n_alpha^2 + 4 doubles, the batch path's layout);
 * off_a / off_b are host copies of the offsets, for the planner.  Every other buffer (the
 * plan, the matrix, the pointer codes) lives in a 64 MiB device array of the library, one
 * per device: at most 2^20 pairs, len(B) at most 1,024, untracked params; a plan or tile
 * that does not fit is refused (NWG_EINVAL).  Results equal nwg_batch_run's for the same
 * pairs.  Calls on one device share the array: calls that overlap on a device race.  The
 * call synchronises the context's stream before returning. */
int nwg_hard_noalloc(nwg_ctx *ctx, const nwg_problem *prob, int n_alpha, int64_t n_pairs,
                     const uint8_t *codes_a_dev, const int64_t *off_a_dev, const uint8_t *codes_b_dev,
                     const int64_t *off_b_dev, const int64_t *off_a, const int64_t *off_b,
                     const double *matrix, const double *gaps4, double *scores_dev, double *grads_dev);

#ifdef __cplusplus
}
#endif
#endif /* NWGRAD_HIP_H */

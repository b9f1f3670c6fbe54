/*
 * gcmi.h -- C ABI of libgcmi.so: the tiled push-relabel grid cut and its
 * tile-list primitives for AMD Instinct MI355X (gfx950).
 *
 * Conventions
 *   - Every function returns an int status: GCMI_OK (0) or a negative code;
 *     gcmi_last_error() gives a thread-local message for the last failure.
 *   - "d_" pointers are device pointers (from gcmi_malloc or any HIP
 *     allocation on the current device); "h_" pointers are host pointers,
 *     borrowed for the duration of the call only.
 *   - `stream` arguments are hipStream_t passed as void*; NULL means the
 *     library's own stream for the current device.  The primitives
 *     (prefix sum, compaction, tile lists, memset) only enqueue their work
 *     on that stream and leave lengths and totals in device memory; the
 *     gcmi_memcpy_* calls (on the library's stream) and calls that return a
 *     value to the host synchronise.
 *   - Handles (gcmi_cut, gcmi_batch, gcmi_server) own their device memory and
 *     one HIP stream each.  A handle is not thread-safe; distinct handles are
 *     independent.
 *
 * Reference interfaces replaced (file:line in rheiland/cl-util):
 *   PrefixSum.scan              PrefixSum.pyx:83-129, prefixsum.cl:81-212
 *   StreamCompact.compact       StreamCompact.pyx:66-82, streamcompact.cl:1-14
 *   IncrementalTileList kernels incremental_tile_list.cl:13-154,
 *                               host logic IncrementalTileList.pyx:100-189
 *   GraphCut                    GraphCut.pyx:80-418, graphcut.cl:20-1016
 */
#ifndef GCMI_H
#define GCMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define GCMI_OK 0
#define GCMI_EINVAL (-1)        /* bad argument                              */
#define GCMI_ENOMEM (-2)        /* device or host allocation failed          */
#define GCMI_ENODEVICE (-3)     /* no usable HIP device                      */
#define GCMI_ENOTCONVERGED (-4) /* cut hit its round cap before terminating  */
#define GCMI_EHIP (-5)          /* a HIP runtime call failed                 */
#define GCMI_EBUDGET (-6)       /* request exceeds the configured HBM budget */
#define GCMI_EBUSY (-7)         /* no free slot / request still in flight    */

/* Operators and logicals of IncrementalTileList (IncrementalTileList.pyx:26-35). */
#define GCMI_OP_EQUAL 0
#define GCMI_OP_GT 1
#define GCMI_OP_LT 2
#define GCMI_OP_GTE 3
#define GCMI_OP_LTE 4
#define GCMI_LOGICAL_AND 0
#define GCMI_LOGICAL_OR 1

/* Thread-local description of the last error (empty string if none). */
const char *gcmi_last_error(void);
/* Library version string. */
const char *gcmi_version(void);
/* Number of HIP devices visible (0 without a GPU; returns GCMI_OK). */
int gcmi_device_count(int *out);

/* ------------------------------------------------- device memory plumbing
 * The reference allocates through pyopencl (cl.Buffer, Buffer2D.py:8-34);
 * these are the equivalents the Python layer uses. */
int gcmi_malloc(void **d_out, size_t bytes);
int gcmi_free(void *d_ptr);
/* Page-locked host memory (hipHostMalloc): copies from it to the device run
 * at the link's rate and need no staging.  No counterpart in the reference
 * (pyopencl copies from ordinary host arrays); the load generator of the
 * serving benchmark keeps its colour images in it. */
int gcmi_host_alloc(void **h_out, size_t bytes);
int gcmi_host_free(void *h_ptr);
int gcmi_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int gcmi_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
int gcmi_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
int gcmi_memset_i32(int32_t *d_dst, int32_t value, size_t count, void *stream);
int gcmi_stream_sync(void *stream);
/* Waits for all work on the current device (hipDeviceSynchronize). */
int gcmi_device_sync(void);

/* ------------------------------------------------------------- primitives */

/* PrefixSum.scan(dArray, dTotal, length) -- PrefixSum.pyx:83-129.
 * In-place exclusive scan of d_array[0..n); the sum of all n inputs is
 * written to d_total[0] (int32 arithmetic, wrapping as the reference's
 * OpenCL ints do).  n may be 0 (total = 0). */
int gcmi_prefix_sum(int32_t *d_array, int32_t *d_total, int n, void *stream);
/* The same, then waits for the scan: *ns_out = its device time in ns, from HIP events around its launches (the reference adds each
 * launch's OpenCL profile time to PrefixSum.elapsed, PrefixSum.pyx:100-101,
 * 116-117, 128-129). */
int gcmi_prefix_sum_timed(int32_t *d_array, int32_t *d_total, int n,
                          void *stream, int64_t *ns_out);

/* StreamCompact.compact(dFlags, dList, dLength, length) --
 * StreamCompact.pyx:66-82.  d_list[0..k) = ascending indices i < n with
 * d_flags[i] != 0; d_length[0] = k.  d_list needs room for n entries.
 * d_flags is not modified (the reference's in-place scatter race,
 * streamcompact.cl:13, does not exist here). */
int gcmi_stream_compact(const int32_t *d_flags, int32_t *d_list,
                        int32_t *d_length, int n, void *stream);

/* init_incremental (incremental_tile_list.cl:27-38): d_tiles[i] = value. */
int gcmi_tiles_fill(int32_t *d_tiles, int n, int32_t value, void *stream);

/* flag (incremental_tile_list.cl:40-64): d_flags[i] = (d_tiles[i] OP operand). */
int gcmi_tiles_flag(const int32_t *d_tiles, int32_t *d_flags, int n, int op,
                    int32_t operand, void *stream);

/* flag_logical (incremental_tile_list.cl:66-108):
 * d_flags[i] = (t1[i] OP1 a) LOGICAL (t2[i] OP2 b). */
int gcmi_tiles_flag_logical(const int32_t *d_t1, const int32_t *d_t2,
                            int32_t *d_flags, int n, int op1, int op2,
                            int32_t a, int32_t b, int logical, void *stream);

/* increment_logical (incremental_tile_list.cl:110-154), used by
 * IncrementalTileList.incorporate (IncrementalTileList.pyx:138-155):
 * where (t1[i] OP1 a) LOGICAL (t2[i] OP2 b), t1[i] = iteration. */
int gcmi_tiles_increment_logical(int32_t *d_t1, const int32_t *d_t2, int n,
                                 int op1, int op2, int32_t a, int32_t b,
                                 int logical, int32_t iteration, void *stream);

/* IncrementalTileList.build(op, operand) (IncrementalTileList.pyx:119-136):
 * flag + compact fused: d_list = ascending i with d_tiles[i] OP operand,
 * d_length[0] = count. */
int gcmi_tiles_build(const int32_t *d_tiles, int n, int op, int32_t operand,
                     int32_t *d_list, int32_t *d_length, void *stream);

/* No reference counterpart.  gcmi_tiles_build for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the scratch of its multi-block
 * pass is a fixed array of the library (about 64 KiB per device, held until
 * the process ends) instead of memory the library allocates; it calls no
 * hipMalloc/hipFree.  n <= GCMI_NOALLOC_MAX_TILES (GCMI_EINVAL above).
 * NULL stream: HIP's null stream.  Every call on one device shares that
 * device's array, so calls that overlap on one device (concurrent streams
 * or threads) race; calls ordered on one stream do not. */
#define GCMI_NOALLOC_MAX_TILES (1 << 26)
int gcmi_tiles_build_noalloc(const int32_t *d_tiles, int n, int op,
                             int32_t operand, int32_t *d_list,
                             int32_t *d_length, void *stream);

/* IncrementalTileList.buildLogical (IncrementalTileList.pyx:157-176), fused. */
int gcmi_tiles_build_logical(const int32_t *d_t1, const int32_t *d_t2, int n,
                             int op1, int op2, int32_t a, int32_t b,
                             int logical, int32_t *d_list, int32_t *d_length,
                             void *stream);

/* IncrementalTileList's visualisation (IncrementalTileList.pyx:191-208,
 * filter incremental_tile_list.cl:156-175): d_out[y * width + x] =
 * 0xFF0000FF if the stamp of the pixel's tile, d_tiles[(y / tile_h) *
 * tiles_w + x / tile_w] of a tiles_w x tiles_h grid, equals `iteration`,
 * else 0xFF000000 (also for pixels beyond the grid's last whole tile). */
int gcmi_tiles_filter(const int32_t *d_tiles, int tiles_w, int tiles_h,
                      int tile_w, int tile_h, int width, int height,
                      int32_t iteration, uint32_t *d_out, void *stream);

/* ---------------------------------------------------------------- GrowCut
 * GrowCut (GrowCut.py:26-153, growcut.cl:28-179): a cellular automaton
 * over a colour image on 16x16 tiles with an IncrementalTileList.  Each
 * pixel has a label (0..255) and a strength (float32); an evolve step
 * updates the tiles stamped with the current iteration synchronously: a
 * pixel's neighbours up, down, left, right attack in turn with g(|c_p -
 * c_q|) * strength_q, and an attack above the defence takes over the
 * neighbour's label and strength (colours c / 255 per channel, alpha
 * included for 4-channel images; |.| = sqrt(((dr^2 + dg^2) + db^2) + da^2);
 * g(X) with x = X / sqrt(3): LINEAR 1 - x, POW2 1 - x^2, POW3 1 - x^3,
 * POW1_5 1 - x sqrt(x), SQRT 1 - sqrt(x); float32, each operation rounded
 * once, in this order).  A tile that changes, and a tile across a changed
 * edge pixel, is stamped with the step's iteration. */
#define GCMI_GROWCUT_WEIGHT_LINEAR 0
#define GCMI_GROWCUT_WEIGHT_POW2 1
#define GCMI_GROWCUT_WEIGHT_POW3 2
#define GCMI_GROWCUT_WEIGHT_POW1_5 3
#define GCMI_GROWCUT_WEIGHT_SQRT 4
typedef struct gcmi_growcut gcmi_growcut;
/* An automaton over a [height][width][channels] uint8 image (copied);
 * labels and strengths start at 0 and every tile's stamp at -1. */
int gcmi_growcut_create(int width, int height, const uint8_t *h_rgb,
                        int channels, int weight, gcmi_growcut **out);
int gcmi_growcut_destroy(gcmi_growcut *gc);
/* GrowCut.label (GrowCut.py:110-123): pixels h_points[0..n) (y * width + x)
 * get `label` and strength 1; their tiles, and the tiles across an edge a
 * point lies on, the current iteration. */
int gcmi_growcut_label(gcmi_growcut *gc, int n, const int32_t *h_points,
                       int label);
/* GrowCut.evolve (GrowCut.py:126-153): up to `iterations` steps, the list's
 * length read every few steps; *complete_out = 1 once a step had no tile
 * to update (nothing changes any more); *steps_out = steps run. */
int gcmi_growcut_evolve(gcmi_growcut *gc, int iterations, int *complete_out,
                        int *steps_out);
/* Current labels (width*height bytes), strengths (float32), tile stamps
 * (tiles across x down int32) and iteration; any pointer may be NULL. */
int gcmi_growcut_read(gcmi_growcut *gc, uint8_t *h_labels, float *h_strength,
                      int32_t *h_tiles, int32_t *iteration_out);

/* ------------------------------------------------------------- grid cut
 * GraphCut (GraphCut.pyx:80-418).  The reference builds its pairwise weights
 * from an RGBA image on the device and takes one float "excess" image; this
 * interface takes int32 capacities instead (all arrays [height][width],
 * row-major, C-contiguous):
 *   source[p] >= 0 (s->p), sink[p] >= 0 (p->t),
 *   right[p] = w(p, p+1) >= 0 (last column ignored),
 *   down[p]  = w(p, p+width) >= 0 (last row ignored);
 *   max(source, sink) + 4*max(right, down) must be < 2^31.
 * Any width/height >= 1 is accepted; the grid is padded internally to a
 * multiple of the tile with zero-capacity pixels. */
typedef struct gcmi_cut gcmi_cut;

/* tile: 16, 32 or 64 (edge length of a square tile). */
int gcmi_cut_create(int width, int height, int tile, gcmi_cut **out);
int gcmi_cut_destroy(gcmi_cut *cut);
/* Upload capacities from host memory (copied; pointers not retained). */
int gcmi_cut_upload(gcmi_cut *cut, const int32_t *h_source,
                    const int32_t *h_sink, const int32_t *h_right,
                    const int32_t *h_down);
/* Same from device memory (row-major int32 [height][width] each). */
int gcmi_cut_upload_device(gcmi_cut *cut, const int32_t *d_source,
                           const int32_t *d_sink, const int32_t *d_right,
                           const int32_t *d_down);
/* Capacities built on the device from an 8-bit colour image (the
 * reference builds its pairwise weights on the device from an RGBA image:
 * weight / load_tiles, graphcut.cl:1018-1131; GraphCut.pyx:80-154).  For a
 * pixel colour I (channels 0..2; a 4th channel is ignored) and the squared
 * colour difference d2 of two 4-neighbours, in float64:
 *   source = rint(unary_scale * min(|I - mu_bg|^2 / unary_div, unary_cap))
 *   sink   = rint(unary_scale * min(|I - mu_fg|^2 / unary_div, unary_cap))
 *   GCMI_PAIR_EXP: w = rint(pair_scale * exp(-d2 / (2 beta))), beta = the
 *     mean d2 over all horizontal and vertical pairs of the image (1 if 0)
 *   GCMI_PAIR_INV: w = rint(pair_scale / (inv_beta * sqrt(d2) + inv_eps)),
 *     the reference's weight with pair_scale = lamda (GraphCut.pyx:80,
 *     graphcut.cl:1030, BETA 1/10, EPSILON 0.05)
 * rint rounds half to even.  Bounds: unary_scale * unary_cap <= 2^26 and
 * the largest weight (pair_scale, or pair_scale / inv_eps) <= 2^24. */
#define GCMI_PAIR_EXP 0
#define GCMI_PAIR_INV 1
typedef struct gcmi_rgb_model {
  double mu_bg[3];
  double mu_fg[3];
  double unary_scale, unary_div, unary_cap;
  int32_t pair_kind;
  int32_t reserved;
  double pair_scale, inv_beta, inv_eps;
} gcmi_rgb_model;
/* The synthetic workloads' model: mu_bg (60,80,100), mu_fg (180,150,90),
 * unaries 10 * min(d / 100, 100), GCMI_PAIR_EXP with pair_scale 50
 * (inv_beta 0.1, inv_eps 0.05 for GCMI_PAIR_INV). */
void gcmi_rgb_model_default(gcmi_rgb_model *model);
/* Builds and uploads the capacities from an [height][width][channels]
 * uint8 image (channels 3 or 4) in host memory (copied). */
int gcmi_cut_upload_rgb(gcmi_cut *cut, const uint8_t *h_rgb, int channels,
                        const gcmi_rgb_model *model);
/* ------------------------------------------------------------ colour GMMs
 * GMM (GMM.pyx:18-420, gmm.cl:33-432): a mixture of k <= 8 Gaussians with
 * diagonal covariances over a pixel's colour channels 0..2 (values 0..255),
 * fitted by EM on the device and scored per pixel.  The score of colour x
 * is its negative log-likelihood, computed in float64 as
 *   e_i = c_i - ((d_0^2 h_i0 + d_1^2 h_i1) + d_2^2 h_i2),  d_j = x_j - mean_ij,
 *   c_i = log(weight_i) - 0.5 ((log(2 pi var_i0) + log(2 pi var_i1))
 *                              + log(2 pi var_i2)),  h_ij = 0.5 / var_ij,
 *   score = -(m + log(sum_i exp(e_i - m))),  m = max_i e_i,
 * each operation rounded once, in this order. */
#define GCMI_GMM_MAX_COMPONENTS 8
typedef struct gcmi_gmm {
  int32_t k; /* components, 1..GCMI_GMM_MAX_COMPONENTS */
  int32_t reserved;
  double weight[GCMI_GMM_MAX_COMPONENTS];  /* >= 0, sum > 0 */
  double mean[GCMI_GMM_MAX_COMPONENTS][3];
  double var[GCMI_GMM_MAX_COMPONENTS][3]; /* diagonal covariance, > 0 */
} gcmi_gmm;
/* GMM.score (GMM.pyx:130-163, score_buf / score_img2d gmm.cl:373-432):
 * d_out[i] = the score of pixel i of n pixels of channels (3 or 4) bytes
 * each at d_rgb, rounded to float32.  Enqueued on `stream`. */
int gcmi_gmm_score(const gcmi_gmm *model, const uint8_t *d_rgb, int channels,
                   long long n, float *d_out, void *stream);
/* GMM.fit (GMM.pyx:165-366; initA, em1, em2, eval gmm.cl:33-340): EM over
 * n >= k samples (channels bytes each at d_rgb), in float64.  Start: the
 * model in *model_io if init_from_model, else weights 1/k, variances
 * init_covar and means the samples at i * (n - 1) / (k - 1).  Each
 * iteration: responsibilities r_ij of every sample (E), then weight_i =
 * sum_j r_ij / n, mean_i = sum_j r_ij x_j / sum_j r_ij, var_i = sum_j
 * r_ij x_j^2 / sum_j r_ij - mean_i^2 + min_covar (M; a component with no
 * responsibility keeps its mean and variance).  Stops after `iters`
 * iterations or when the mean log-likelihood of the E step changed by less
 * than tol (tol <= 0: never).  *model_io gets the model, *loglik_out (may
 * be NULL) the last E step's mean log-likelihood, *iters_out (may be NULL)
 * the iterations run.  Synchronises. */
int gcmi_gmm_fit(const uint8_t *d_rgb, int channels, long long n, int k, int iters,
                 double tol, double init_covar, double min_covar, int init_from_model,
                 gcmi_gmm *model_io, double *loglik_out, int *iters_out);
/* Capacities built on the device from an 8-bit colour image with GMM
 * unaries (the reference's tools feed GMM scores to the cut as unaries,
 * test_GraphCut.py:44-47): source = q(score under bg), sink = q(score under
 * fg), q(s) = rint(unary_scale * min(max(s, 0), unary_cap)) (half to even);
 * pairwise weights as gcmi_cut_upload_rgb builds them with `pairwise` (its
 * unary fields are ignored).  unary_scale * unary_cap <= 2^26. */
int gcmi_cut_upload_rgb_gmm(gcmi_cut *cut, const uint8_t *h_rgb, int channels,
                            const gcmi_rgb_model *pairwise, const gcmi_gmm *fg,
                            const gcmi_gmm *bg, double unary_scale,
                            double unary_cap);
/* Copies the stored capacities (as uploaded, built or edited; each array
 * [height][width] int32) to host memory. */
int gcmi_cut_capacities(gcmi_cut *cut, int32_t *h_source, int32_t *h_sink,
                        int32_t *h_right, int32_t *h_down);
/* GraphCut.cut (GraphCut.pyx:348-418): run push-relabel to termination.
 *   sweeps      in-tile push/relabel sweeps per tile residency (<=0: default)
 *   bfs         push rounds between global relabels (<=0: default)
 *   max_rounds  cap on push rounds (<=0: no cap); GCMI_ENOTCONVERGED if hit
 *   flow_out    max-flow value (int64), may be NULL
 * After it returns GCMI_OK, gcmi_cut_labels gives the minimal source set. */
int gcmi_cut_run(gcmi_cut *cut, int sweeps, int bfs, int64_t max_rounds,
                 int64_t *flow_out);
/* The reference's step hooks (GraphCut.pyx:187-346), for callers that
 * drive the solve themselves.  They solve the cut's current capacities:
 * the first hook after an upload, edit or run sets the solve up (from
 * scratch, or warm after edits), and an upload, edit, paint, forget or run
 * drops a stepped solve in progress.
 *   gcmi_cut_push       GraphCut.push: one push round (sweeps in-tile
 *                       sweeps, <= 0: default) over the tiles holding
 *                       active pixels.
 *   gcmi_cut_relabel    GraphCut.relabel: a relabel of the tiles whose arcs
 *                       or excess changed since their distances were
 *                       computed (every tile if this solve has not relabelled
 *                       yet and started from scratch), run until its queue
 *                       drains; *active_out (may be NULL) = tiles left
 *                       holding active pixels.
 *   gcmi_cut_start_bfs  GraphCut.startBfs: the same over every tile.
 *   gcmi_cut_is_completed  GraphCut.isCompleted: *done_out = 1 right after
 *                       a relabel that left no active tile -- the solve is
 *                       then finished (minimal source set, labels, flow in
 *                       *flow_out, may be NULL) -- or when no solve is in
 *                       progress and the last one is complete; else 0.
 * The result is the same exact cut as gcmi_cut_run's. */
int gcmi_cut_push(gcmi_cut *cut, int sweeps);
int gcmi_cut_relabel(gcmi_cut *cut, int32_t *active_out);
int gcmi_cut_start_bfs(gcmi_cut *cut, int32_t *active_out);
int gcmi_cut_is_completed(gcmi_cut *cut, int *done_out, int64_t *flow_out);
/* Per-tile flags of the cut's current state, the nearest counterparts of
 * the reference's four tile lists (GraphCut.pyx:142-149): d_out[ty * tiles
 * across + tx] (int32, device memory, one per tile of the padded grid) is
 * nonzero for a tile that
 *   GCMI_TILES_ACTIVE   holds active pixels (tilelistLoad: tiles pushed next)
 *   GCMI_TILES_BORDER   has inflow pending from a neighbour (tilelistBorder)
 *   GCMI_TILES_CHANGED  changed since its distances were computed (the
 *                       next relabel's tiles; tilelistBfs)
 *   GCMI_TILES_EXCESS   holds positive excess (where the minimal source set
 *                       grows from; tilelistEdges has no exact counterpart) */
#define GCMI_TILES_ACTIVE 0
#define GCMI_TILES_BORDER 1
#define GCMI_TILES_CHANGED 2
#define GCMI_TILES_EXCESS 3
int gcmi_cut_tile_flags(gcmi_cut *cut, int which, int32_t *d_out);
/* Warm-start edit (no reference counterpart; the brush tools of the
 * reference re-upload the whole excess image, test_GraphCut.py:74-75): new
 * source/sink capacities for n distinct pixels (h_index = y*width+x).  The
 * next gcmi_cut_run reflows from the current residual graph instead of
 * starting over; its flow and labels equal a from-scratch cut of the edited
 * capacities.  Before the first run an edit only rewrites the capacities. */
int gcmi_cut_edit(gcmi_cut *cut, int n, const int32_t *h_index,
                  const int32_t *h_source, const int32_t *h_sink);
/* Warm-start edit by a brush stroke, rasterised on the device (no reference
 * counterpart; the reference's brush paints the excess image on the host,
 * test_GraphCut.py:74-75).  h_xy holds npts (x, y) points (2*npts doubles,
 * npts <= 16); every pixel whose centre (x, y) lies within width/2 of the
 * polyline -- a union of capsules, one per segment, each test rounded in
 * float64 in the order documented in the library -- gets source/sink
 * capacities (source, sink).  Fewer than 2 points paint nothing. */
int gcmi_cut_paint(gcmi_cut *cut, int npts, const double *h_xy, double width,
                   int32_t source, int32_t sink);
/* Labels of the last cut: h_out[y*width+x] = 1 for the minimal source
 * (foreground) set -- pixels reachable from the source in the final residual
 * graph -- else 0.  (The reference exposes only dHeight, GraphCut.pyx:62.) */
int gcmi_cut_labels(gcmi_cut *cut, uint8_t *h_out);
/* Same into device memory (width*height bytes). */
int gcmi_cut_labels_device(gcmi_cut *cut, uint8_t *d_out);
/* Same packed to bits: bit p%32 of h_bits[p/32] is the label of pixel
 * p = y*width+x (ceil(width*height/32) words; no reference counterpart). */
int gcmi_cut_labels_packed(gcmi_cut *cut, uint32_t *h_bits);
/* Drops the solved state, so the next gcmi_cut_run starts from the stored
 * capacities (every edit included) instead of reflowing warm.  (The
 * reference's GraphCut.reset, GraphCut.pyx:420-421, does nothing.) */
int gcmi_cut_forget(gcmi_cut *cut);
/* Counters of the last run (the first n of them):
 *   [0] push rounds  [1] tile residencies  [2] global relabels
 *   [3] BFS tile passes  [4] flood tile passes  [5] host syncs
 *   [6] algorithmic bytes: 16 per image pixel read + 25 per padded pixel
 *       written by the first push round, 49 per pixel of a later residency,
 *       5 per pixel of a relabel's first pass, 9 of a later BFS pass, 2 of
 *       a flood pass, 2 per image pixel of the labels
 *   [7] tiles  [8] relabel first-pass tiles  [9] later BFS tile passes  */
int gcmi_cut_stats(gcmi_cut *cut, int64_t *out, int n);
/* Device-resident tile-blocked height image (int32, padded) for inspection:
 * copies padded_w*padded_h ints in row-major order into h_out. */
int gcmi_cut_heights(gcmi_cut *cut, int32_t *h_out);
/* Any int32 state field of the tile-blocked solver state, row-major over the
 * padded grid (padded_w*padded_h ints): the counterpart of GraphCut.execute
 * / filter_transpose (GraphCut.pyx:423-437, graphcut.cl:1134-1169), which
 * turn the reference's tile-transposed buffers into images for display. */
#define GCMI_FIELD_HEIGHT 0
#define GCMI_FIELD_EXCESS 1
#define GCMI_FIELD_CAP_UP 2
#define GCMI_FIELD_CAP_DOWN 3
#define GCMI_FIELD_CAP_LEFT 4
#define GCMI_FIELD_CAP_RIGHT 5
int gcmi_cut_field(gcmi_cut *cut, int which, int32_t *h_out);
/* Padded geometry: out[0]=padded width, out[1]=padded height, out[2]=tile. */
int gcmi_cut_geometry(gcmi_cut *cut, int *out3);
/* Per-kernel HIP-event timing of later runs (no reference counterpart; the
 * reference sums OpenCL event profiles in PrefixSum.elapsed,
 * PrefixSum.pyx:100-101).  Kernel classes, in order:
 *   0 push_first (round 1, builds the state from the inputs)  1 push
 *   2 fold  3 bfs_init  4 bfs_round  5 flood  6 labels  7 compact */
#define GCMI_KERNEL_CLASSES 8
int gcmi_cut_profile(gcmi_cut *cut, int enable);
/* Solver options (no reference counterpart).
 * GCMI_OPT_LOCAL_RELABEL (default 0): in runs that reflow edits, and after
 *   the first relabel of any run, the global relabel recomputes the tiles
 *   whose arcs or excess changed instead of every tile, for at most a few
 *   relabels per run; the result is the same exact cut either way. */
#define GCMI_OPT_LOCAL_RELABEL 1
/* GCMI_OPT_DEVICE_ROUNDS (default 0, at most 64): value > 0 makes runs
 *   device-driven: two push/relabel iterations (three when edits of more
 *   than two images are reflowed together) with `value` BFS rounds per
 *   relabel, then the minimal-source-
 *   set flood with `value` rounds, every launch of a fixed shape that exits
 *   at once when the run is already done, and one host synchronisation at
 *   the end.  A run the sequence does not finish goes on host-driven from
 *   where it stopped.  0: host-driven (the host reads counts between
 *   phases).  The result is the same exact cut either way.
 * GCMI_OPT_GRAPH (default 0): 1 records the device-driven sequence of a run
 *   as one hipGraph and replays it for later runs of the same images and
 *   parameters (with 8 device rounds unless set).
 * Both are ignored while profiling. */
#define GCMI_OPT_DEVICE_ROUNDS 2
#define GCMI_OPT_GRAPH 3
int gcmi_cut_set_option(gcmi_cut *cut, int option, int value);
/* Summed milliseconds and launch counts per class for the last run. */
int gcmi_cut_kernel_times(gcmi_cut *cut, double *ms_out, int64_t *launches_out,
                          int n);

/* ---------------------------------------------------------------- batches
 * Many independent images (any sizes) solved together: every kernel launch
 * of a run covers the tiles of all images, so the launches, host syncs and
 * tail rounds of `count` cuts are shared.  No reference counterpart (the
 * reference cuts one image per GraphCut object).  Image i keeps its own
 * capacities, flow and labels; a batch is a pool of images in one set of
 * device arrays. */
typedef struct gcmi_batch gcmi_batch;

int gcmi_batch_create(int count, const int *widths, const int *heights,
                      int tile, gcmi_batch **out);
int gcmi_batch_destroy(gcmi_batch *batch);
int gcmi_batch_upload(gcmi_batch *batch, int i, const int32_t *h_source,
                      const int32_t *h_sink, const int32_t *h_right,
                      const int32_t *h_down);
int gcmi_batch_upload_device(gcmi_batch *batch, int i,
                             const int32_t *d_source, const int32_t *d_sink,
                             const int32_t *d_right, const int32_t *d_down);
/* Capacities of image i built from a colour image (as gcmi_cut_upload_rgb). */
int gcmi_batch_upload_rgb(gcmi_batch *batch, int i, const uint8_t *h_rgb,
                          int channels, const gcmi_rgb_model *model);
/* Capacities of image i with GMM unaries (as gcmi_cut_upload_rgb_gmm). */
int gcmi_batch_upload_rgb_gmm(gcmi_batch *batch, int i, const uint8_t *h_rgb,
                              int channels, const gcmi_rgb_model *pairwise,
                              const gcmi_gmm *fg, const gcmi_gmm *bg,
                              double unary_scale, double unary_cap);
/* Warm-start edit of image i (as gcmi_cut_edit). */
int gcmi_batch_edit(gcmi_batch *batch, int i, int n, const int32_t *h_index,
                    const int32_t *h_source, const int32_t *h_sink);
/* Brush stroke on image i (as gcmi_cut_paint). */
int gcmi_batch_paint(gcmi_batch *batch, int i, int npts, const double *h_xy,
                     double width, int32_t source, int32_t sink);
/* Solves all images; flows_out (count int64, may be NULL) gets each flow. */
int gcmi_batch_run(gcmi_batch *batch, int sweeps, int bfs, int64_t max_rounds,
                   int64_t *flows_out);
/* Solves images which[0..n) only (flows_out gets n values). */
int gcmi_batch_run_subset(gcmi_batch *batch, int n, const int *which,
                          int sweeps, int bfs, int64_t max_rounds,
                          int64_t *flows_out);
int gcmi_batch_labels(gcmi_batch *batch, int i, uint8_t *h_out);
int gcmi_batch_stats(gcmi_batch *batch, int64_t *out, int n);
int gcmi_batch_profile(gcmi_batch *batch, int enable);
/* GCMI_OPT_* options, as gcmi_cut_set_option. */
int gcmi_batch_set_option(gcmi_batch *batch, int option, int value);
int gcmi_batch_kernel_times(gcmi_batch *batch, double *ms_out,
                            int64_t *launches_out, int n);

/* ---------------------------------------------------------------- serving
 * Resident sessions and a continuously batching worker.  No reference
 * counterpart: the reference serves one interactive user
 * (test_GraphCut.py:74-75 re-uploads the excess image and re-cuts after
 * each brush stroke).  Each session is one image of a pool sized by an HBM
 * budget; its residual graph, heights and tile state stay in device memory
 * between requests.  A worker thread takes every queued request it can (one
 * per session, at most max_batch, oldest first), applies them and solves
 * those sessions in one pool run; requests that arrive meanwhile go into
 * the next run.  Strokes reflow warm from the session's residual graph
 * (as gcmi_cut_edit).  The handle is thread-safe: any thread may submit,
 * wait and read masks.  Modes: */
#define GCMI_SERVE_WARM 0  /* batched; strokes reflow warm                */
#define GCMI_SERVE_PLAIN 1 /* one request at a time, cut from scratch     */
#define GCMI_SERVE_COLD 2  /* batched; every session cut from scratch     */
typedef struct gcmi_server gcmi_server;

/* hbm_budget: device bytes the session pools may use (GCMI_EBUDGET if too
 * small for one tile); max_sessions: session slots; max_batch: sessions
 * per solve.  The warm and cold modes serve from 3 lanes, the plain mode
 * from 1 (gcmi_server_create_lanes). */
int gcmi_server_create(int tile, size_t hbm_budget, int max_sessions,
                       int max_batch, int mode, gcmi_server **out);
/* The same with `lanes` (1..4; 0: the mode's default) lanes: each lane is
 * a pool with its own stream and worker, holding its share of the sessions
 * (a session opens in the lane with the most free room) and of the budget;
 * the lanes' solves run concurrently on the device.  The plain mode always
 * has one lane. */
int gcmi_server_create_lanes(int tile, size_t hbm_budget, int max_sessions,
                             int max_batch, int mode, int lanes,
                             gcmi_server **out);
/* Stops the worker and frees everything; queued requests are dropped.  No
 * other call on this server may be in progress or follow. */
int gcmi_server_destroy(gcmi_server *server);
/* sweeps / bfs of every later solve (<= 0: defaults, as gcmi_cut_run). */
int gcmi_server_set_params(gcmi_server *server, int sweeps, int bfs);
/* GCMI_OPT_* options of every later solve, as gcmi_cut_set_option. */
int gcmi_server_set_option(gcmi_server *server, int option, int value);
/* Admits a width x height session if a slot and pool room are free
 * (GCMI_EBUDGET otherwise); *session_out = its id. */
int gcmi_server_open(gcmi_server *server, int width, int height,
                     int *session_out);
/* Waits for the session's queued requests, then frees its room. */
int gcmi_server_close(gcmi_server *server, int session);
/* Queues a new image (capacities as gcmi_cut_upload, copied at once);
 * *req_out = request id. */
int gcmi_server_submit_image(gcmi_server *server, int session,
                             const int32_t *h_source, const int32_t *h_sink,
                             const int32_t *h_right, const int32_t *h_down,
                             int64_t *req_out);
/* The colour model of later gcmi_server_submit_rgb requests (default:
 * gcmi_rgb_model_default). */
int gcmi_server_set_rgb_model(gcmi_server *server, const gcmi_rgb_model *model);
/* Queues a new image given as an [height][width][channels] uint8 colour
 * image (channels 3 or 4, copied at once); its capacities are built on the
 * device with the server's colour model. */
int gcmi_server_submit_rgb(gcmi_server *server, int session, const uint8_t *h_rgb,
                           int channels, int64_t *req_out);
/* Queues a brush stroke: new source/sink capacities for n distinct pixels
 * (checked as gcmi_cut_edit; needs a submitted image). */
int gcmi_server_submit_stroke(gcmi_server *server, int session, int n,
                              const int32_t *h_index, const int32_t *h_source,
                              const int32_t *h_sink, int64_t *req_out);
/* Queues a brush stroke given as a polyline (as gcmi_cut_paint). */
int gcmi_server_submit_polyline(gcmi_server *server, int session, int npts,
                                const double *h_xy, double width,
                                int32_t source, int32_t sink,
                                int64_t *req_out);

/* One request of a replayed trace. */
#define GCMI_REQ_IMAGE 0    /* p0..p3: source, sink, right, down          */
#define GCMI_REQ_STROKE 1   /* n pixels; p0..p2: index, source, sink      */
#define GCMI_REQ_POLYLINE 2 /* n points; p0: xy; width, source, sink     */
#define GCMI_REQ_RGB 3      /* n channels; p0: colour image               */
typedef struct gcmi_request {
  int64_t t_ns; /* submit at this steady-clock (CLOCK_MONOTONIC) time, in
                   ns; it is the request's arrival time.  <= 0: at once */
  int32_t session;
  int32_t kind;
  int32_t n;
  int32_t source, sink;
  int32_t reserved;
  double width;
  const void *p0, *p1, *p2, *p3;
} gcmi_request;
/* Open-loop load generator: submits reqs[0..count) in order from the
 * calling thread, each at its time (sleeping until then; a late submission
 * keeps the scheduled arrival time, so the lag counts as latency).
 * req_out (count int64, may be NULL) gets the request ids.  Stops at the
 * first refused request and returns its status. */
int gcmi_server_replay(gcmi_server *server, int count, const gcmi_request *reqs,
                       int64_t *req_out);

/* Waits up to timeout_ms (< 0: no limit) for a request.  info_out (7 int64,
 * may be NULL): status (0 queued, 1 done, 2 failed), flow, arrival, start
 * and completion times (ns, steady clock), sessions solved with it, and the
 * time it entered the queue (after a new image's staging copy).
 * GCMI_EBUSY if still queued at the timeout. */
int gcmi_server_wait(gcmi_server *server, int64_t req, int timeout_ms,
                     int64_t *info_out);
/* Bit-packed mask of the session's last completed request (bit p%32 of
 * word p/32 = label of pixel p, ceil(width*height/32) words); *req_out
 * (may be NULL) = that request's id.  GCMI_EBUSY while being solved. */
int gcmi_server_mask(gcmi_server *server, int session, uint32_t *h_bits,
                     int64_t *req_out);
/* [0] solves  [1] requests served  [2] queued  [3] open sessions
 * [4] pool tiles in use  [5] pool tile capacity  [6] device bytes
 * [7] [8] [9] worker ns applying requests, solving, listing mask copies
 * (summed over lanes)  [10] lanes  [11] worker ns waiting for the device
 * (inside the solves' synchronisations; the rest of [8] is host work) */
int gcmi_server_stats(gcmi_server *server, int64_t *out, int n);
/* Drops the record of a completed request. */
int gcmi_server_forget(gcmi_server *server, int64_t req);

#ifdef __cplusplus
}
#endif
#endif /* GCMI_H */

// Host side of an image pool: device arrays shared by all images of the
// pool, per-image tile/pixel ranges, and the solve driver that runs the
// push-relabel rounds over any subset of the pool's images at once.
//
// A plain cut (GraphCut) is a pool of one image; a batch is a pool of many;
// the serving loop keeps every resident session in one pool and solves the
// sessions that have requests together.
#pragma once

#include <vector>

#include "gcmi_kernels.h"

namespace gcmi {

constexpr int kDefaultSweeps = 4;  // in-tile sweeps per residency
constexpr int kDefaultBfs = 1;     // push rounds per global relabel
constexpr int kRoundBatch = 2;     // BFS / flood rounds before the first host check
constexpr int kRoundBatchMax = 8;  // ... and before each later one
constexpr int kRelabelGridCap = 65536;  // blocks of a relabel's first pass
constexpr int kLocalRelabels = 6;       // local relabels per solve, then full ones
constexpr int kDeviceRoundsMax = 64;     // BFS / flood rounds per relabel, device-driven
constexpr int kDeviceRoundsDefault = 8;

enum {
  CNT_ACT = 0,
  CNT_POS = 1,
  CNT_SEL = 2,
  CNT_FRESH = 3,
  CNT_Q = 4,  // counts of the three rotating tile queues
  CNT_DIRTY = 7,
  CNT_EP = 8,     // two alternating queue-epoch slots
  CNT_DONE = 10,  // a complete relabel found no active tile
  CNT_GSEL = 11,  // selected tiles, or 0 once the solve is done
  CNT_FIRST = 12,  // tiles of an incremental flood's first pass
  CNT_BLOCKS = 13,  // blocks of a gated compaction that have finished (0 between launches)
  CNT_N = 14
};

enum {
  STAT_ROUNDS = 0,
  STAT_RESIDENCIES = 1,
  STAT_GLOBAL_RELABELS = 2,
  STAT_BFS_PASSES = 3,
  STAT_FLOOD_PASSES = 4,
  STAT_SYNCS = 5,
  STAT_ALG_BYTES = 6,
  STAT_TILES = 7,
  STAT_RELABEL_TILES = 8,  // tiles of the relabels' first passes
  STAT_BFS_ROUND_PASSES = 9,  // tile passes of the later BFS rounds
  STAT_N = 10
};

// Kernel classes timed by the optional HIP-event profiler.
enum {
  K_PUSH_FIRST = 0,
  K_PUSH = 1,
  K_FOLD = 2,
  K_BFS_INIT = 3,
  K_BFS_ROUND = 4,
  K_FLOOD = 5,
  K_LABELS = 6,
  K_COMPACT = 7,
  K_N = 8
};

// Per-launch HIP events on the pool's stream, summed per kernel class
// after the solve's final synchronisation.
struct Prof {
  bool on = false;
  int used = 0;
  int cap = 0;
  hipEvent_t* ev = nullptr;  // 2 per launch
  int* cls = nullptr;
  double ms[K_N] = {0};
  int64_t launches[K_N] = {0};
  void reset();
  void begin(hipStream_t s, int k);
  void end(hipStream_t s);
  void collect();
  void release();
};

struct Pool;

// A push launch: a later round, or round 1 of the fresh images, storing
// its heights or not.  Round 1 may leave them unstored when a full relabel
// of those tiles comes next: it rewrites every distance of each tile it
// lists and, for a listed neighbour, reads its mask only.
enum PushKind { PUSH_LATER = 0, PUSH_FIRST = 1, PUSH_FIRST_NO_H = 2 };

// Kernel launchers for one tile size.
struct PoolOps {
  int tile;
  int threads;
  void (*push)(Pool*, int grid, int sweeps, int kind);
  void (*fold)(Pool*, int grid, const int32_t* list, int cnt);
  void (*bfs_init)(Pool*, int grid, const int32_t* list, int cnt, bool all, const Queue& q);
  void (*bfs_round)(Pool*, int grid, const Queue& q);
  void (*flood)(Pool*, int grid, const Queue& q, int mode, int32_t* lmark);
  void (*labels)(Pool*, int grid, const int32_t* list, int cnt, int nflows);
  void (*posf)(Pool*, int grid, unsigned zmask);
  void (*flood_set)(Pool*, int nsel, int32_t* lmark, const int32_t* done);
  void (*flood_start_full)(Pool*, int nsel, unsigned zmask, const int32_t* done);
  void (*edit)(Pool*, int img, int n, const int32_t* idx, const int32_t* cs, const int32_t* ct);
  void (*paint)(Pool*, const PaintJob* d_jobs, int njobs, int nrows);
  void (*untile)(Pool*, int img, const int32_t* src, int32_t* out);
};

const PoolOps* pool_ops(int tile);

// Row-major pixel ranges start on a multiple of kPixAlign elements: the
// rows of an image whose width is a multiple of 4 load as 16-byte vectors,
// and an image's bit-packed mask starts at bit 0 of a word of its own.
constexpr long long kPixAlign = 32;
inline long long pix_len(int W, int H) {
  return ((long long)W * H + kPixAlign - 1) / kPixAlign * kPixAlign;
}

constexpr size_t kStatusTally = 64;
constexpr size_t kStatusFlows = 128;
static_assert(CNT_N * 4 <= kStatusTally && kStatusTally + TALLY_N * 8 <= kStatusFlows,
              "status block layout");

struct PoolImage {
  bool live = false;
  int W = 0, H = 0, tw = 0, th = 0;
  int tile0 = 0, ntiles = 0;
  long long pix0 = 0;
  int32_t max_pair = 0;  // largest pairwise weight uploaded
  bool uploaded = false;  // capacities present
  bool fresh = true;      // no solved state yet (or capacities re-uploaded)
  bool edited = false;    // edited since the last solve: next solve is warm
  bool solved = false;    // labels / flow valid
  int64_t flow = 0;
};

// A first-fit allocator of [start, start + len) ranges.
struct RangeAlloc {
  std::vector<std::pair<long long, long long>> free_;  // (start, len), sorted
  long long cap = 0;
  void init(long long c);
  long long alloc(long long len);  // -1 if no room
  void release(long long start, long long len);
  long long largest() const;
};

enum { STAGE_PREP = 0, STAGE_EDIT = 1, STAGE_JOBS = 2, STAGE_RGBJOBS = 3, STAGE_N = 4 };

struct Pool {
  const PoolOps* ops = nullptr;
  int T = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int cap_tiles = 0;
  long long cap_pix = 0;
  int cap_imgs = 0;
  Geo geo{};
  CutDev dev{};
  // device arrays (beyond those in dev)
  ImgDesc* d_img = nullptr;
  int32_t* tile_img = nullptr;
  int4* tgeo = nullptr;  // per tile: two int4 of geometry (tile_geo)
  int32_t *sel = nullptr, *fresh = nullptr, *plist = nullptr, *alist = nullptr;
  int32_t* q = nullptr;  // three rotating tile queues of cap_tiles entries
  int32_t* dlist = nullptr;  // dirty tiles of the current relabel
  // In solves with edited images, and after a solve's first relabel:
  // relabel the tiles that changed (every tile of a fresh image has)
  // instead of every tile.  Clean tiles keep their stored distances, valid
  // but possibly stale; the solve switches to full relabels after
  // kLocalRelabels.  Exact either way: termination is decided by a relabel
  // over a valid labelling.
  bool local_relabel = false;
  bool trace = false;  // GCMI_TRACE=1: per-phase lines on stderr, one round per host check
  int ep_par = 0;         // epoch slot the next queue round reads
  bool flood_inc = false;  // the current flood recomputes changed tiles only
  bool flood_ran = false;  // a flood has added to the final-excess sums in this solve
  bool first_no_h = false;  // this solve's round 1 left its heights unstored
  // pinned staging buffers with a host-to-device copy queued on the stream
  // since its last synchronisation (a buffer is rewritten only after it)
  bool stage_pending[STAGE_N] = {};
  bool ep_reset = false;  // restart stamps and epochs before the next solve
  // Device-driven solves (dd_rounds > 0): every launch of a fixed shape,
  // dd_rounds BFS / flood rounds per relabel, gates on the device instead of
  // host checks, one synchronisation at the end (the host loop finishes a
  // solve the sequence did not).  graph: record that sequence for an image
  // list once as a hipGraph and replay it while the list and parameters
  // stay the same.
  int dd_rounds = 0;
  bool graph = false;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<int> graph_key;
  int graph_flood_cur = 0, graph_ep_par = 0;
  // the lists the kernels of the current solve read: null when the
  // selection is the whole pool in tile order (positions are tiles)
  int32_t *lsel = nullptr, *lfresh = nullptr;
  // the solve's status block: counts, then tallies (dev.tally) and the
  // images' flows, at byte offsets 0, kStatusTally and kStatusFlows, so
  // one copy brings all three to the host at a solve's end
  uint8_t* d_status = nullptr;
  uint8_t* h_status = nullptr;
  // the table d_prep holds (the last one copied), and whether a solve is
  // being recorded (its copies become graph nodes)
  std::vector<uint8_t> last_prep;
  bool capturing = false;
  int32_t* counts = nullptr;
  int32_t* scratch = nullptr;
  Prep* d_prep = nullptr;
  long long* d_flows = nullptr;
  int32_t* d_edit = nullptr;  // staging for edits: idx, cs, ct per edit
  int32_t* d_imgfresh = nullptr;           // per image: cut from scratch in this solve
  unsigned long long* d_rgbsum = nullptr;  // pair-difference sums of a colour upload (padded slots)
  // colour uploads: the job table of a launch's images (allocated on first
  // use)
  void* h_rgbjobs = nullptr;
  void* d_rgbjobs = nullptr;
  uint8_t* d_rgb = nullptr;                // staging for host colour images
  long long rgb_cap = 0;                   // bytes
  PaintJob *h_jobs = nullptr, *d_jobs = nullptr;  // staging for paint launches
  int jobs_cap = 0;
  long long edit_cap = 0;     // pixels
  std::vector<uint64_t> mark;  // host bitmap for the distinct-index check
  // pinned host buffers
  int32_t* h_counts = nullptr;
  Prep* h_prep = nullptr;
  long long* h_flows = nullptr;
  unsigned long long* h_tally = nullptr;
  int32_t* h_edit = nullptr;
  // host bookkeeping
  std::vector<PoolImage> imgs;
  RangeAlloc tiles_alloc, pix_alloc;
  size_t device_bytes = 0;
  int grid_cap = 2048;
  int64_t stats[STAT_N] = {0};
  int64_t ns_wait = 0;  // host time spent waiting for the stream (all solves so far)
  Prof prof;
  // a solve driven step by step (pool_step_*): its image (-1: none), whether
  // it started from scratch, its in-tile sweeps, the last step, the active
  // tile count of the last relabel and the relabels so far
  int step_img = -1;
  bool step_fresh = false;
  int step_sweeps = kDefaultSweeps;
  int step_last = 0;
  int32_t step_nact = -1;
  int step_relabels = 0;
};

enum { STEP_BEGIN = 0, STEP_PUSH = 1, STEP_RELABEL = 2 };

// Bytes of device memory a pool of the given capacity allocates.
size_t pool_bytes(int tile, int cap_tiles, long long cap_pix, int cap_imgs);
int pool_init(Pool* p, int tile, int cap_tiles, long long cap_pix, int cap_imgs,
              hipStream_t stream);
void pool_free(Pool* p);
// Adds an image; returns its id in *out.  GCMI_EBUDGET if no room.
int pool_add_image(Pool* p, int W, int H, int* out);
int pool_remove_image(Pool* p, int img);
int pool_upload(Pool* p, int img, const int32_t* cs, const int32_t* ct, const int32_t* right,
                const int32_t* down, bool device_src);
// Copies an image's stored capacities (row-major) to host memory.
int pool_capacities(Pool* p, int img, int32_t* cs, int32_t* ct, int32_t* right, int32_t* down);
// Checks a colour model; *max_pair (may be null) gets its largest weight.
int check_rgb_model(const gcmi_rgb_model* m, int32_t* max_pair);
// Builds an image's capacities on the device from a colour image in device
// memory ([H][W][channels] uint8), on the pool's stream (no synchronisation).
// Colour uploads of several distinct images in one set of launches (device
// bytes, 4-byte aligned).
struct RgbReq {
  int img;
  const uint8_t* d_rgb;
  int channels;
  const gcmi_rgb_model* model;
};
int pool_upload_rgb_many(Pool* p, int count, const RgbReq* reqs);
int pool_upload_rgb(Pool* p, int img, const uint8_t* d_rgb, int channels,
                    const gcmi_rgb_model* m);
// The same from host memory (copied; synchronises).
int pool_upload_rgb_host(Pool* p, int img, const uint8_t* h_rgb, int channels,
                         const gcmi_rgb_model* m);
// Copies a host colour image of image img into the pool's colour staging
// (grown on demand) on the pool's stream; *d_out = its device address.
int pool_stage_rgb(Pool* p, int img, const uint8_t* h_rgb, int channels, const uint8_t** d_out);
// Capacities of image img from a host colour image: pairwise weights as
// the colour model `pair` builds them, unaries from two colour GMMs --
// source = the background model's quantised score, sink = the
// foreground's (gcmi_gmm.hip; copied, synchronises).
int pool_upload_rgb_gmm_host(Pool* p, int img, const uint8_t* h_rgb, int channels,
                             const gcmi_rgb_model* pair, const gcmi_gmm* fg, const gcmi_gmm* bg,
                             double scale, double cap);
// Warm-start edit of n distinct pixels (row-major indices local to the
// image) from host arrays.
int pool_edit(Pool* p, int img, int n, const int32_t* idx, const int32_t* cs, const int32_t* ct);
// Several images' edits with one staging copy.
struct EditReq {
  int img;
  int n;
  const int32_t *idx, *cs, *ct;
};
int pool_edit_many(Pool* p, int count, const EditReq* e);
// Checks an edit of an image of npx pixels: indices distinct and in range,
// capacities >= 0 and max(cs, ct) + 4 * wmax < 2^31.  `mark` is a zeroed
// bitmap of at least npx bits (grown here if smaller) and is zeroed again
// on return.
int check_edit(long long npx, int n, const int32_t* idx, const int32_t* cs, const int32_t* ct,
               int64_t wmax, std::vector<uint64_t>& mark);
// Builds a stroke for a W x H image from npts (x, y) points (xy, 2 * npts
// doubles), checking its arguments; a stroke of fewer than 2 points or
// outside the image gets an empty box (bw = 0).
int make_polyline(int W, int H, int npts, const double* xy, double width, int32_t cs, int32_t ct,
                  int64_t wmax, Polyline* out);
// Warm-start edit by a brush stroke: the pixels within width/2 of the
// polyline get unaries (cs, ct).
int pool_paint(Pool* p, int img, int npts, const double* xy, double width, int32_t cs,
               int32_t ct);
// Several strokes (of different images) painted by one launch; all are
// checked before any is painted.
struct PaintReq {
  int img;
  int npts;
  const double* xy;
  double width;
  int32_t cs, ct;
};
int pool_paint_many(Pool* p, int count, const PaintReq* reqs);
// Solves the listed images together (fresh ones from their capacities,
// edited ones warm).  flows_out (may be NULL) receives one flow per image.
int pool_solve(Pool* p, const int* img_ids, int n, int sweeps, int bfs, int64_t max_rounds,
               int64_t* flows_out);
// One image's solve driven step by step, as the reference's GraphCut hooks
// (GraphCut.pyx:187-346).  The first step sets the solve up (from scratch,
// or warm after edits); uploads, edits, forget and pool_solve drop it.
//   push: one push round over the tiles holding active pixels;
//   relabel: a global relabel run until its queue drains -- over every
//     tile (full), or over the tiles whose arcs or excess changed (valid
//     once this solve has relabelled, or after a warm edit); *nact gets the
//     tiles left active;
//   completed: right after a relabel that left no active tile, finishes the
//     solve (minimal source set, labels, flow) and sets *done = 1.
int pool_step_push(Pool* p, int img, int sweeps);
int pool_step_relabel(Pool* p, int img, bool full, int32_t* nact);
int pool_step_completed(Pool* p, int img, int* done, int64_t* flow);
int pool_labels(Pool* p, int img, uint8_t* out, bool device_dst);
// Labels packed to bits (bit p % 32 of word p / 32 = label of pixel p).
int pool_labels_packed(Pool* p, int img, uint32_t* out);
// The device address of an image's bit-packed mask (ceil(W*H/32) words, bit
// p % 32 of word p / 32 = label of pixel p), kept up to date by every solve.
const uint32_t* packed_mask(const Pool* p, int img);
// Drops the solved state: the next solve starts from the capacities.
int pool_forget(Pool* p, int img);
// GCMI_OPT_* options of the C ABI.
int pool_set_option(Pool* p, int option, int value);
int pool_heights(Pool* p, int img, int32_t* out);
// A tile-blocked int32 state field (GCMI_FIELD_*) of an image, row-major
// over its padded grid, to host memory.
int pool_field(Pool* p, int img, int which, int32_t* out);
int validate_caps(const int32_t* cs, const int32_t* ct, const int32_t* right,
                  const int32_t* down, size_t n, int32_t* max_pair);
int tiles_for(int W, int H, int tile, int* tw, int* th);

}  // namespace gcmi

// Image pool: allocation, uploads, warm-start edits and the solve driver.
//
// Solve driver (GraphCut.cut, GraphCut.pyx:348-418), over the selected
// images at once:
//   prepare                  tile lists; zero the sums being recomputed
//   fresh images: round 1    k_push<FIRST> builds state from the capacities
//   edited images: round 1   k_push over the tiles the edits made active
//   repeat:
//     fold                   pending edge inflow -> state (dirty tiles)
//     global relabel         k_bfs_init over the dirty tiles (those whose
//                            arcs or excess changed since their distances
//                            were computed), then rounds over tile queues
//                            until a queue is empty;
//                            act = tiles with excess that can reach a sink
//     stop if act is empty
//     push round             k_push over the active tiles
//   minimal source set       flood from {e > 0}; labels; flow values
// Lists are built on the device.  The host reads list lengths once per
// relabel (the reference waits after every launch and reads a length per
// tile-list build, GraphCut.pyx:169-323).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>

#include "gcmi_pool.h"

namespace gcmi {

// ------------------------------------------------------------- profiler

void Prof::reset() {
  used = 0;
  for (int k = 0; k < K_N; ++k) {
    ms[k] = 0;
    launches[k] = 0;
  }
}

void Prof::begin(hipStream_t s, int k) {
  if (used == cap) {
    int ncap = cap ? cap * 2 : 256;
    hipEvent_t* nev = (hipEvent_t*)realloc(ev, sizeof(hipEvent_t) * 2 * ncap);
    if (nev) ev = nev;
    int* ncls = (int*)realloc(cls, sizeof(int) * ncap);
    if (ncls) cls = ncls;
    if (!nev || !ncls) {
      on = false;
      return;
    }
    for (int i = 2 * cap; i < 2 * ncap; ++i) (void)hipEventCreate(&ev[i]);
    cap = ncap;
  }
  cls[used] = k;
  (void)hipEventRecord(ev[2 * used], s);
}

void Prof::end(hipStream_t s) {
  (void)hipEventRecord(ev[2 * used + 1], s);
  ++used;
}

void Prof::collect() {
  for (int i = 0; i < used; ++i) {
    float t = 0;
    if (hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]) == hipSuccess) {
      ms[cls[i]] += t;
      launches[cls[i]]++;
    }
  }
  used = 0;
}

void Prof::release() {
  for (int i = 0; i < 2 * cap; ++i) (void)hipEventDestroy(ev[i]);
  free(ev);
  free(cls);
  ev = nullptr;
  cls = nullptr;
  cap = used = 0;
}

// ------------------------------------------------------------- launchers

__global__ void k_fill_i32(int32_t* p, long long n, int32_t v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Writes the geometry lines of image `img`'s tiles.
__global__ void k_tgeo_fill(int4* tgeo, ImgDesc d, int img) {
  for (int lt = blockIdx.x * blockDim.x + threadIdx.x; lt < d.ntiles;
       lt += gridDim.x * blockDim.x) {
    const size_t t = (size_t)d.tile0 + lt;
    const int tx = lt % d.tw, ty = lt / d.tw;
    tgeo[2 * t] = make_int4(img, (int)((unsigned)tx | ((unsigned)ty << 16)),
                            (int)((unsigned)d.tw | ((unsigned)d.th << 16)), d.W);
    tgeo[2 * t + 1] = make_int4(d.H, (int)(unsigned)(unsigned long long)d.pix0,
                                (int)(unsigned)((unsigned long long)d.pix0 >> 32), 0);
  }
}

__global__ void k_clear_listed(int32_t* flags, const int32_t* __restrict__ list,
                               const int32_t* __restrict__ count) {
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    flags[list ? list[i] : i] = 0;
}


// The tiles whose reach a flood must recompute: the changed tiles (fchg, the
// `n` tiles of `cur`) and, transitively, every tile that imported reach
// (imp) next to one -- its reach may have come through the change.  One
// block, a worklist of the tiles added last (in buf0 / buf1, each room for
// every tile); a tile enters once (the fchg swap).  Ends when a pass adds
// nothing, after at most as many passes as there are tiles.  All threads of
// the block call it.
__device__ void imp_closure(const Geo& g, CutDev& s, const int32_t* cur, int n, int32_t* buf0,
                            int32_t* buf1, int* nnext) {
  int32_t* next = buf0;
  while (n > 0) {
    if (threadIdx.x == 0) *nnext = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const TG q = tile_geo(g, cur[i]);
      const int nb[4] = {q.ty > 0 ? q.t - q.tw : -1, q.ty < q.th - 1 ? q.t + q.tw : -1,
                         q.tx > 0 ? q.t - 1 : -1, q.tx < q.tw - 1 ? q.t + 1 : -1};
      // the four neighbours' loads, then their claims, each group in flight
      // together: unconditional, a missing neighbour standing in as the
      // tile itself (fchg holds 0 or 1, so or-ing 0 only reads it)
      int im[4], was[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) im[k] = s.imp[nb[k] >= 0 ? nb[k] : q.t] != 0 && nb[k] >= 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        was[k] = atomicOr(&s.fchg[nb[k] >= 0 ? nb[k] : q.t], im[k] ? 1 : 0);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (im[k] && was[k] == 0) next[atomicAdd(nnext, 1)] = nb[k];
    }
    __syncthreads();
    n = *nnext;
    cur = next;
    next = next == buf0 ? buf1 : buf0;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(1024)
k_imp_closure(Geo g, CutDev s, const int32_t* __restrict__ init, const int32_t* __restrict__ count,
              int32_t* buf0, int32_t* buf1) {
  __shared__ int nnext;
  imp_closure(g, s, init, *count, buf0, buf1, &nnext);
}

// A small incremental flood's start (at most kSmallSel tiles) in one block:
// the images' final positive-excess sums (k_posf), the changed tiles of the
// images reflowed warm (none unless *done, when given) and their import
// closure -- one launch instead of three on the path of a single image's
// reflow.  zmask: the counts zeroed, as k_posf's.  buf0 receives the
// changed tiles; buf1 / buf2 hold the closure's worklists.
constexpr int kSmallSel = 16384;
__global__ void __launch_bounds__(1024)
k_flood_start_small(Geo g, CutDev s, const int32_t* __restrict__ sel, int n, int32_t* counts,
                    unsigned zmask, const int32_t* __restrict__ imgfresh,
                    const int32_t* __restrict__ done, int32_t* buf0, int32_t* buf1, int32_t* buf2) {
  __shared__ int nseed, nnext;
  zero_counts(counts, zmask);
  if (threadIdx.x == 0) nseed = 0;
  __syncthreads();
  const bool seed = !done || *done != 0;
  const unsigned long long below = (1ull << lane_id()) - 1;
  for (int base = 0; base < n; base += blockDim.x) {
    const int i = base + threadIdx.x;
    int img = -1, t = 0;
    long long v = 0;
    bool f = false;
    if (i < n) {
      t = sel ? sel[i] : i;
      img = g.tile_img[t];
      v = s.tpos[t];
      f = seed && s.fchg[t] != 0 && !imgfresh[img];
    }
    const int img0 = __shfl(img, 0);
    const unsigned long long row = SUM_POSF * kSumSlots + (base / blockDim.x) % kSumSlots;
    if (__all(img == img0 || img < 0)) {
      v = wave_sum64(v);
      if (lane_id() == 0 && img0 >= 0 && v)
        atomicAdd(&s.sums[(size_t)img0 * SUM_ROWS * kSumSlots + row], (unsigned long long)v);
    } else if (img >= 0 && v) {
      atomicAdd(&s.sums[(size_t)img * SUM_ROWS * kSumSlots + row], (unsigned long long)v);
    }
    const unsigned long long m = __ballot(f);
    int at = 0;
    if (lane_id() == 0 && m) at = atomicAdd(&nseed, __popcll(m));
    at = __shfl(at, 0);
    if (f) buf0[at + __popcll(m & below)] = t;
  }
  __syncthreads();
  imp_closure(g, s, buf0, nseed, buf1, buf2, &nnext);
}

// Local relabel set: the dirty tiles (arcs or excess changed; a push dirties
// both tiles of every arc it changes), appended to `out` in any order
// (*out_count zero before) and marked rflag[t] = the epoch the relabel's
// first pass runs with (*epoch; older marks never equal it, so none is
// cleared).  No ring of clean
// tiles is needed: a distance that rose in a dirty tile only loosens the
// constraint h(c) <= h(d) + 1 on a clean neighbour's stored distance, and
// one that fell is passed on by the rounds.
__global__ void k_dilate(Geo g, CutDev s, const int32_t* __restrict__ list,
                         const int32_t* __restrict__ count, int32_t* out, int32_t* out_count,
                         const int32_t* __restrict__ epoch) {
  const int n = *count;
  const int ep = *epoch;
  const unsigned long long below = (1ull << lane_id()) - 1;
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    int f = 0, t = 0;
    if (i < n) {
      const TG q = tile_geo(g, list ? list[i] : i);
      t = q.t;
      f = s.dirty[t];
      if (f) s.rflag[t] = ep;
    }
    const unsigned long long m = __ballot(f != 0);
    if (m == 0) continue;
    int at = 0;
    if (lane_id() == 0) at = atomicAdd(out_count, __popcll(m));
    at = __shfl(at, 0);
    if (f) out[at + __popcll(m & below)] = t;
  }
}

// Builds the solve's tile lists, resets the fresh images' tiles (pending
// slots and flags: the tile range may have held another image before),
// zeroes the sums it recomputes (all rows of a fresh image, the
// final-excess row of every image) and the tallies, sets the counts and the
// first queue epoch.  Grid: (images, blocks per image); an image's tiles
// are a contiguous range, so its lists and pending slots are filled as flat
// arrays.  pend_per_tile: the pending slots of a tile (4T).
__global__ void k_prepare(Geo g, CutDev s, const Prep* __restrict__ prep, int nsel_tiles,
                          int nfresh_tiles, int32_t* sel, int32_t* fresh, int32_t* counts,
                          int32_t* imgfresh, int pend_per_tile) {
  const Prep p = prep[blockIdx.x];
  const bool fr = p.fresh_off >= 0;
  if (blockIdx.y == 0 && threadIdx.x == 0) imgfresh[p.img] = fr;
  const ImgDesc d = g.img[p.img];
  const int stride = gridDim.y * blockDim.x;
  for (int j = blockIdx.y * blockDim.x + threadIdx.x; j < d.ntiles; j += stride) {
    const int t = d.tile0 + j;
    sel[p.sel_off + j] = t;
    if (fr) {
      fresh[p.fresh_off + j] = t;
      s.pflag[t] = 0;
      s.act[t] = 0;
      s.haspos[t] = 0;
    }
  }
  if (fr) {
    int4* pe = reinterpret_cast<int4*>(s.pend + (size_t)d.tile0 * pend_per_tile);
    const long long nv = (long long)d.ntiles * pend_per_tile / 4;
    for (long long k = blockIdx.y * blockDim.x + threadIdx.x; k < nv; k += stride)
      pe[k] = make_int4(0, 0, 0, 0);
  }
  if (blockIdx.y == 0)
    for (int j = threadIdx.x; j < SUM_ROWS * kSumSlots; j += blockDim.x)
      if (fr || j / kSumSlots == SUM_POSF) s.sums[(size_t)p.img * SUM_ROWS * kSumSlots + j] = 0;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    counts[CNT_SEL] = nsel_tiles;
    counts[CNT_FRESH] = nfresh_tiles;
    counts[CNT_GSEL] = nsel_tiles;
    counts[CNT_DONE] = 0;
    counts[CNT_ACT] = 0;
    counts[CNT_POS] = 0;
    // the first relabel's queue and local-set counts (a device-driven run
    // appends to them without a fill)
    counts[CNT_Q] = counts[CNT_Q + 1] = counts[CNT_Q + 2] = counts[CNT_DIRTY] = 0;
    // the next epoch: above every stamp written so far (stamps start at 0)
    counts[CNT_EP] = max(max(counts[CNT_EP], counts[CNT_EP + 1]), 1);
    for (int k = 0; k < TALLY_N; ++k) s.tally[k] = 0;
  }
}

// ---------------------------------------------------- capacities from colour
// Squared colour difference of two pixels (channels 0..2).
__device__ __forceinline__ int rgb_d2(const uint8_t* a, const uint8_t* b) {
  const int dr = (int)a[0] - (int)b[0], dg = (int)a[1] - (int)b[1], db = (int)a[2] - (int)b[2];
  return dr * dr + dg * dg + db * db;
}

// Consecutive row-major pixels per thread of the colour kernels (their loads
// are independent, so a thread has them all in flight at once).
constexpr int kRgbPx = 4;
constexpr int kRgbNT = 256;

// Column of pixel i of a row-major image W wide (32-bit arithmetic when the
// image allows it).
__device__ __forceinline__ int rgb_col(long long i, int W, bool small) {
  return small ? (int)((unsigned)i % (unsigned)W) : (int)(i % W);
}

// A chunk of the colour kernels is kRgbNT * kRgbPx consecutive pixels; a
// block covers a few consecutive chunks of one image and stages the bytes of
// each in LDS: its pixels and the next one (right neighbours), and the
// pixels one row below them.  The bytes of the next chunk are loaded into
// registers while the block works on the current one.  The capacities
// (float64 work, few waves per SIMD) take one chunk per block, so that a
// 1024^2 image has four blocks per CU rather than one working its chunks in
// turn (k_rgb_caps on 4096^2 / 1024^2: 97.6 / 19.7 us with 4 chunks, 82.2 /
// 11.0 us with 1).  The pair sums read their words straight from the image
// and take 2 chunks' pixels per block (4096^2 / 1024^2: 27.0 / 6.0 us; 27.1
// / 9.2 us with 4).  Words per staged range (up to 4 channels, plus the
// range's offset inside its first word).
constexpr int kRgbBlockPx = kRgbNT * kRgbPx;
constexpr int kRgbSumChunks = 2, kRgbCapsChunks = 1;
constexpr int kRgbStageWords = (kRgbBlockPx + 1) + 1;
constexpr int kRgbStageIters = (kRgbStageWords + kRgbNT - 1) / kRgbNT;

// One thread's share of a staged byte range: words threadIdx.x + j * kRgbNT.
struct RgbWords {
  uint32_t v[kRgbStageIters];
  int nw;  // words in the range
};

// Loads bytes [b0, b1) of a colour image of nb bytes (4-byte aligned),
// clipped to the image: whole words where the image holds them, a last
// partial word byte by byte.  Byte i lands at word (i - (b0 & ~3)) / 4.
// (the image pointer comes from the job table, so the compiler cannot tell
// that it is global memory: flat loads count against the LDS counter too,
// and every wait for the staged LDS bytes would also wait for them)
using GlobalU8 = const __attribute__((address_space(1))) uint8_t;
using GlobalU32 = const __attribute__((address_space(1))) uint32_t;
// four words at a word-aligned address (one 16-byte load)
typedef uint32_t RgbW4 __attribute__((ext_vector_type(4), aligned(4)));
using GlobalW4 = const __attribute__((address_space(1))) RgbW4;
__device__ __forceinline__ void rgb_load(const uint8_t* __restrict__ rgb_flat, long long nb,
                                         long long b0, long long b1, RgbWords& w) {
  GlobalU8* rgb = (GlobalU8*)rgb_flat;
  const long long a0 = b0 & ~3LL;
  b1 = b1 < nb ? b1 : nb;
  w.nw = b1 > a0 ? (int)((b1 - a0 + 3) >> 2) : 0;
#pragma unroll
  for (int j = 0; j < kRgbStageIters; ++j) {
    const int k = threadIdx.x + j * kRgbNT;
    const long long o = a0 + 4LL * k;
    w.v[j] = 0;
    if (k < w.nw) {
      if (o + 4 <= nb) {
        w.v[j] = *(GlobalU32*)(rgb + o);
      } else {
        for (int i = 0; i < 4 && o + i < nb; ++i) w.v[j] |= (uint32_t)rgb[o + i] << (8 * i);
      }
    }
  }
}

__device__ __forceinline__ void rgb_store(const RgbWords& w, uint32_t* dst) {
#pragma unroll
  for (int j = 0; j < kRgbStageIters; ++j) {
    const int k = threadIdx.x + j * kRgbNT;
    if (k < w.nw) dst[k] = w.v[j];
  }
}

// The loads of the chunk at pixel p0: own holds pixels p0 .. p0 +
// kRgbBlockPx (one past, for right neighbours), dn the pixels a row below.
__device__ __forceinline__ void rgb_load_chunk(const uint8_t* __restrict__ rgb, long long n,
                                               int W, int ch, long long p0, RgbWords& own,
                                               RgbWords& dn) {
  const long long nb = n * ch;
  rgb_load(rgb, nb, p0 * ch, (p0 + kRgbBlockPx + 1) * ch, own);
  rgb_load(rgb, nb, (p0 + W) * ch, (p0 + W + kRgbBlockPx) * ch, dn);
}

// Walks a block's chunks of image n = W * H pixels: f(p0, O, own0, D, dn0)
// for each chunk at pixel p0 that starts inside the image, with its bytes
// in LDS (pixel i at O + (i * ch - own0) for p0 <= i <= p0 + kRgbBlockPx,
// the pixel a row below at D + ((i + W) * ch - dn0)).  All threads call it.
template <int CHUNKS, class F>
__device__ __forceinline__ void rgb_chunks(const uint8_t* __restrict__ rgb, long long n, int W,
                                           int ch, long long first, F&& f) {
  __shared__ uint32_t own[kRgbStageWords], dn[kRgbStageWords];
  const uint8_t* O = reinterpret_cast<const uint8_t*>(own);
  const uint8_t* D = reinterpret_cast<const uint8_t*>(dn);
  RgbWords wo, wd;
  rgb_load_chunk(rgb, n, W, ch, first, wo, wd);
  for (int c = 0; c < CHUNKS; ++c) {
    const long long p0 = first + (long long)c * kRgbBlockPx;
    if (p0 >= n) break;  // uniform
    rgb_store(wo, own);
    rgb_store(wd, dn);
    __syncthreads();
    const long long p1 = p0 + kRgbBlockPx;
    if (c + 1 < CHUNKS && p1 < n) rgb_load_chunk(rgb, n, W, ch, p1, wo, wd);
    f(p0, O, (p0 * ch) & ~3LL, D, ((p0 + W) * ch) & ~3LL);
    __syncthreads();  // the chunk's bytes are read before the next store
  }
}

// One colour image of a launch: its bytes, pool pixel offset, size and
// model, and its first block in each of the launch's two kernels (blocks of
// kRgbSumChunks / kRgbCapsChunks chunks, images in turn).
enum { RGB_SUM = 0, RGB_CAPS = 1 };
struct RgbJob {
  const uint8_t* rgb;
  long long pix0;
  int W, H, ch;
  int block0[2];
  gcmi_rgb_model m;
};
constexpr int kRgbJobsMax = 64;  // images per launch
// u64 words per pair-sum slot: each slot of a colour upload's sums on a
// 128-byte line of its own.  Atomics on slots that share a line are done one
// after the other (16,384 blocks adding to 64 packed slots take ~40 us, to
// 64 padded ones ~1 us).
constexpr int kRgbSumStride = 16;

// A launch's job table: the first blocks of its jobs in each kernel
// (ascending, from 0), then the jobs.
struct RgbJobs {
  int block0[2][kRgbJobsMax];
  RgbJob job[kRgbJobsMax];
};

// The job this block of kernel K (RGB_SUM, RGB_CAPS) belongs to: one ballot
// of wave 0 over the first blocks (njobs <= kWave); the result is
// wave-uniform.  All threads call it.
template <int K>
__device__ __forceinline__ int rgb_job(const RgbJobs* __restrict__ J, int njobs, int* s_j) {
  static_assert(kRgbJobsMax <= kWave, "one lane per job");
  if (threadIdx.x < kWave) {
    const int b0 = threadIdx.x < njobs ? J->block0[K][threadIdx.x] : 0x7fffffff;
    const unsigned long long m = __ballot(b0 <= (int)blockIdx.x);
    if (threadIdx.x == 0) *s_j = __popcll(m) - 1;
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*s_j);
}

// d2 of the pairs of four pixels whose bytes are whole in registers: the
// pixels' CH-byte colours from byte 0 of own (and the next pixel's after
// them: their right neighbours), the pixels a row below from byte 0 of dn.
template <int CH>
__device__ __forceinline__ uint32_t rgb_d2_group(const uint32_t* own, const uint32_t* dn) {
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int dr = 0, dd = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int b = CH * k + c;
      const int v = (own[b >> 2] >> (8 * (b & 3))) & 0xff;
      const int br = b + CH;
      const int r = (own[br >> 2] >> (8 * (br & 3))) & 0xff;
      const int d = (dn[b >> 2] >> (8 * (b & 3))) & 0xff;
      dr += (v - r) * (v - r);
      dd += (v - d) * (v - d);
    }
    acc += (uint32_t)(dr + dd);
  }
  return acc;
}

// Per image of the launch, the sum of d2 over its horizontal and vertical
// neighbour pairs (beta's numerator; an exact integer), spread over
// kSumSlots slots at sums[(job * kSumSlots + slot) * kRgbSumStride].  A
// block takes kRgbSumChunks * kRgbBlockPx consecutive pixels, a thread
// groups of four pixels kRgbNT groups apart; the grid covers every image
// once.  A group inside the image whose pixels all have a right neighbour
// and a row below reads its bytes as words straight from the image (the
// group starts on a word: 4 * ch * group; its row below is realigned with
// alignbyte); the groups at row and image ends read byte by byte.
constexpr int kRgbSumGroups = kRgbSumChunks * kRgbBlockPx / (4 * kRgbNT);
static_assert(kRgbPx == 4 && kRgbSumGroups * 4 * kRgbNT == kRgbSumChunks * kRgbBlockPx,
              "four pixels per group");
__global__ void __launch_bounds__(kRgbNT)
k_rgb_pair_sum(const RgbJobs* __restrict__ J, int njobs, unsigned long long* sums) {
  __shared__ long long red[kRgbNT / kWave];
  __shared__ int s_j;
  const int j = rgb_job<RGB_SUM>(J, njobs, &s_j);
  const RgbJob* jobs = J->job;
  GlobalU8* rgb = (GlobalU8*)jobs[j].rgb;
  const int W = jobs[j].W, H = jobs[j].H, ch = jobs[j].ch;
  const long long n = (long long)W * H, nb = n * ch;
  const bool small = n <= 0x7fffffffLL;
  const long long first =
      (long long)(blockIdx.x - jobs[j].block0[RGB_SUM]) * kRgbSumChunks * kRgbBlockPx;
  // the words of every group first (all loads in flight), then the sums
  uint32_t own[kRgbSumGroups][5], dn[kRgbSumGroups][4];
  bool fast[kRgbSumGroups];
#pragma unroll
  for (int g = 0; g < kRgbSumGroups; ++g) {
    const long long i0 = first + 4LL * (g * kRgbNT + threadIdx.x);
    const int x0 = i0 < n ? rgb_col(i0, W, small) : 0;
    const long long ob = i0 * ch, db = (i0 + W) * ch, da = db & ~3LL;
    // own: 4 ch + 3 bytes (the group and its right neighbour's colour) in
    // 4 (ch 3) or 5 (ch 4) words; below: the 4 ch bytes from db, inside the
    // 4 words from da (ch 3: 12 bytes at db - da <= 3; ch 4: db = da)
    const int nwo = ch == 3 ? 4 : 5;
    fast[g] = i0 + 4 <= n && x0 + 4 < W && i0 + 4 + W <= n && ob + 4 * nwo <= nb &&
              da + 16 <= nb;
    if (fast[g]) {
      const RgbW4 a = *(GlobalW4*)(rgb + ob), b = *(GlobalW4*)(rgb + da);
      own[g][0] = a.x, own[g][1] = a.y, own[g][2] = a.z, own[g][3] = a.w;
      own[g][4] = nwo == 5 ? *(GlobalU32*)(rgb + ob + 16) : 0u;
      dn[g][0] = b.x, dn[g][1] = b.y, dn[g][2] = b.z, dn[g][3] = b.w;
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) own[g][k] = dn[g][k] = 0u;
      own[g][4] = 0u;
    }
  }
  uint32_t acc = 0;  // at most 4 kRgbSumGroups 2 3 255^2 < 2^32
#pragma unroll
  for (int g = 0; g < kRgbSumGroups; ++g) {
    const long long i0 = first + 4LL * (g * kRgbNT + threadIdx.x);
    if (fast[g]) {
      const int sh = (int)(((i0 + W) * ch) & 3);
      // (d[3] is read only for ch 4, where sh = 0)
      uint32_t d[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        d[k] = __builtin_amdgcn_alignbyte(k < 3 ? dn[g][k + 1] : 0u, dn[g][k], sh);
      acc += ch == 3 ? rgb_d2_group<3>(own[g], d) : rgb_d2_group<4>(own[g], d);
    } else if (i0 < n) {
      int x = rgb_col(i0, W, small);
      for (int k = 0; k < 4 && i0 + k < n; ++k) {
        const long long i = i0 + k;
        GlobalU8* q = rgb + i * ch;
        int v = 0;
        if (x + 1 < W)
          for (int c = 0; c < 3; ++c) v += ((int)q[c] - (int)q[ch + c]) * ((int)q[c] - (int)q[ch + c]);
        if (i + W < n) {
          GlobalU8* r = q + (long long)W * ch;
          for (int c = 0; c < 3; ++c) v += ((int)q[c] - (int)r[c]) * ((int)q[c] - (int)r[c]);
        }
        acc += (uint32_t)v;
        if (++x == W) x = 0;
      }
    }
  }
  const long long tot = block_sum64<kRgbNT>((long long)acc, red);
  if (threadIdx.x == 0 && tot)
    atomicAdd(&sums[((size_t)j * kSumSlots + blockIdx.x % kSumSlots) * kRgbSumStride],
              (unsigned long long)tot);
}

// 2 beta of the exponential pairwise model from an image's pair-sum slots
// (beta = 1 if the image has no pairs or they are all equal).  Called by
// every thread of a block; result in every thread.
__device__ double rgb_two_beta(int W, int H, const unsigned long long* __restrict__ pair_sum,
                               double* slot) {
#pragma clang fp contract(off)
  if (threadIdx.x < kWave) {
    unsigned long long v = 0;
    for (int j = threadIdx.x; j < kSumSlots; j += kWave) v += pair_sum[(size_t)j * kRgbSumStride];
    v = (unsigned long long)wave_sum64((long long)v);
    if (threadIdx.x == 0) {
      double beta = 1.0;
      const long long cnt = (long long)(W - 1) * H + (long long)W * (H - 1);
      if (cnt > 0) beta = (double)v / (double)cnt;
      if (!(beta > 0.0)) beta = 1.0;
      *slot = 2.0 * beta;
    }
  }
  __syncthreads();
  return *slot;
}

// The models round each float64 operation once, in the order they state,
// and end with rint.  Their divisions by a per-image constant c are first
// tried as a product with 1/c (rounded once): that quotient is within 3
// units of roundoff (u = 2^-53) of the correctly rounded one, so the value
// handed to rint moves by less than 5u of itself for the unary and less
// than (3 * 750 + 6)u = 2^-41.9 of itself for the exponential weight (exp's
// argument is at most about 750 in size where its result is not 0).  When
// that value lies further than 16 times the bound (2^-46 and 2^-38
// relative, plus the same absolute) from every half-integer, rint gives the
// same integer as with the correctly rounded division; otherwise the
// division is done.  The capacities are those of the divisions, bit for
// bit.
__device__ __forceinline__ bool rint_settled(double v, double rel) {
  const double k = rint(v);
  return 0.5 - fabs(v - k) > rel * (fabs(v) + 1.0);
}

// Weight of a pair with squared colour difference d2; inv_tb = 1 / two_beta,
// inv_tb_f and scale_f: 1 / two_beta and pair_scale as float.  The
// exponential weight is first tried in float32: with a = d2 * inv_tb_f
// (d2 < 2^24 is exact), the float value scale_f * exp(-a) is within (a + 1)
// 2^-22 of itself of the float64 one (the roundings of inv_tb_f, a, the
// hardware exp's argument and result, scale_f and the product); where it
// lies further than 4 (2a + 16) 2^-22 (w + 1) from every half-integer, rint
// of either is the same integer; otherwise the float64 steps decide.  For
// a > 80 the weight is below 2^24 e^-79.9 < 0.5 (the model caps pair_scale
// at 2^24): 0.
__device__ __forceinline__ int32_t rgb_pair_weight(const gcmi_rgb_model& m, int d2,
                                                   double two_beta, double inv_tb,
                                                   float inv_tb_f, float scale_f) {
#pragma clang fp contract(off)
  const double d = (double)d2;
  if (m.pair_kind == GCMI_PAIR_EXP) {
    const float a = (float)d2 * inv_tb_f;
    if (a > 80.0f) return 0;
    const float wf = scale_f * __expf(-a);
    const float k = rintf(wf);
    if (0.5f - fabsf(wf - k) > (2.0f * a + 16.0f) * 0x1p-20f * (wf + 1.0f)) return (int32_t)k;
    const double w = m.pair_scale * exp(-(d * inv_tb));
    if (rint_settled(w, 0x1p-38)) return (int32_t)rint(w);
    return (int32_t)rint(m.pair_scale * exp(-d / two_beta));
  }
  return (int32_t)rint(m.pair_scale / (m.inv_beta * sqrt(d) + m.inv_eps));
}

// Unary of a pixel of colour (r, g, b) against mean mu; inv_div = 1 /
// unary_div.
__device__ __forceinline__ int32_t rgb_unary(const gcmi_rgb_model& m, const double* mu, double r,
                                             double g, double b, double inv_div) {
#pragma clang fp contract(off)
  const double e0 = r - mu[0], e1 = g - mu[1], e2 = b - mu[2];
  const double d = (e0 * e0 + e1 * e1) + e2 * e2;
  const double t = m.unary_scale * fmin(d * inv_div, m.unary_cap);
  if (rint_settled(t, 0x1p-46)) return (int32_t)rint(t);
  return (int32_t)rint(m.unary_scale * fmin(d / m.unary_div, m.unary_cap));
}

// Row-major int32 capacities of the launch's images from their colour
// images (beta from each image's pair sum); kRgbPx pixels per thread and
// chunk, stored as one 16-byte vector per array when whole (the outputs
// start 16-byte aligned: pool pixel offsets are multiples of 32).
__global__ void __launch_bounds__(kRgbNT)
k_rgb_caps(const RgbJobs* __restrict__ J, int njobs, const unsigned long long* __restrict__ sums,
           int32_t* __restrict__ cs, int32_t* __restrict__ ct, int32_t* __restrict__ right,
           int32_t* __restrict__ down) {
#pragma clang fp contract(off)
  __shared__ int s_j;
  const int j = rgb_job<RGB_CAPS>(J, njobs, &s_j);
  const RgbJob* jobs = J->job;
  const uint8_t* rgb = jobs[j].rgb;
  const int W = jobs[j].W, H = jobs[j].H, ch = jobs[j].ch;
  const gcmi_rgb_model& m = jobs[j].m;
  __shared__ double s_two_beta;
  const double two_beta = m.pair_kind == GCMI_PAIR_EXP
      ? rgb_two_beta(W, H, sums + (size_t)j * kSumSlots * kRgbSumStride, &s_two_beta) : 0.0;
  const double inv_tb = m.pair_kind == GCMI_PAIR_EXP ? 1.0 / two_beta : 0.0;
  const float inv_tb_f = (float)inv_tb, scale_f = (float)m.pair_scale;
  const double inv_div = 1.0 / m.unary_div;
  const long long pix0 = jobs[j].pix0;
  cs += pix0;
  ct += pix0;
  right += pix0;
  down += pix0;
  const long long n = (long long)W * H;
  const bool small = n <= 0x7fffffffLL;
  const long long first =
      (long long)(blockIdx.x - jobs[j].block0[RGB_CAPS]) * kRgbCapsChunks * kRgbBlockPx;
  rgb_chunks<kRgbCapsChunks>(rgb, n, W, ch, first,
             [&](long long p0, const uint8_t* O, long long own0, const uint8_t* D, long long dn0) {
#pragma clang fp contract(off)
               const long long i0 = p0 + (long long)threadIdx.x * kRgbPx;
               int x = i0 < n ? rgb_col(i0, W, small) : 0;
               int32_t a[kRgbPx], b[kRgbPx], rr[kRgbPx], dd[kRgbPx];
#pragma unroll
               for (int k = 0; k < kRgbPx; ++k) {
                 const long long i = i0 + k;
                 a[k] = b[k] = rr[k] = dd[k] = 0;
                 if (i < n) {
                   const uint8_t* q = O + (i * ch - own0);
                   const double r = q[0], g = q[1], bl = q[2];
                   a[k] = rgb_unary(m, m.mu_bg, r, g, bl, inv_div);
                   b[k] = rgb_unary(m, m.mu_fg, r, g, bl, inv_div);
                   if (x + 1 < W)
                     rr[k] = rgb_pair_weight(m, rgb_d2(q, q + ch), two_beta, inv_tb, inv_tb_f,
                                             scale_f);
                   if (i + W < n)
                     dd[k] = rgb_pair_weight(m, rgb_d2(q, D + ((i + W) * ch - dn0)), two_beta,
                                             inv_tb, inv_tb_f, scale_f);
                 }
                 if (++x == W) x = 0;
               }
               static_assert(kRgbPx == 4, "one int4 per array and thread");
               if (i0 + kRgbPx <= n) {
                 *reinterpret_cast<int4*>(cs + i0) = make_int4(a[0], a[1], a[2], a[3]);
                 *reinterpret_cast<int4*>(ct + i0) = make_int4(b[0], b[1], b[2], b[3]);
                 *reinterpret_cast<int4*>(right + i0) = make_int4(rr[0], rr[1], rr[2], rr[3]);
                 *reinterpret_cast<int4*>(down + i0) = make_int4(dd[0], dd[1], dd[2], dd[3]);
               } else {
                 for (int k = 0; k < kRgbPx && i0 + k < n; ++k) {
                   cs[i0 + k] = a[k];
                   ct[i0 + k] = b[k];
                   right[i0 + k] = rr[k];
                   down[i0 + k] = dd[k];
                 }
               }
             });
}

// Zeroes the final-excess sums of the selected images: the first flood
// pass adds them up (a run whose recording stopped short floods again).
// Also zeroes the counts the flood's compactions and first pass append to,
// and the active count its labels list is built in.
__global__ void k_zero_posf(unsigned long long* sums, const Prep* __restrict__ prep,
                            int32_t* counts) {
  unsigned long long* r = sums + (size_t)prep[blockIdx.x].img * SUM_ROWS * kSumSlots;
  for (int j = threadIdx.x; j < kSumSlots; j += blockDim.x) r[SUM_POSF * kSumSlots + j] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts[CNT_FIRST] = counts[CNT_POS] = counts[CNT_ACT] = 0;
    counts[CNT_Q] = counts[CNT_Q + 1] = counts[CNT_Q + 2] = 0;
  }
}

// The start of a flood from scratch in one pass over the solve's tiles (sel,
// n of them): each image's final positive-excess sum (of its tiles' sums,
// as k_posf), the tiles holding excess (the first pass's list `out`, count
// counts[CNT_POS], zero before; none unless *done, when given), and every
// tile's reach bytes, import and changed flags cleared.  zmask: counts
// block 0 zeroes (none of them is appended to here).
constexpr int kFloodStartNT = 256;
template <int T>
__global__ void __launch_bounds__(kFloodStartNT)
k_flood_start_full(Geo g, CutDev s, const int32_t* __restrict__ sel, int n, int32_t* counts,
                   unsigned zmask, int32_t* out, const int32_t* __restrict__ done) {
  __shared__ int32_t tl[kFloodStartNT];
  zero_counts(counts, zmask);
  const bool gate = !done || *done != 0;
  const unsigned long long below = (1ull << lane_id()) - 1;
  const unsigned long long row = SUM_POSF * kSumSlots + blockIdx.x % kSumSlots;
  for (int base = blockIdx.x * kFloodStartNT; base < n; base += gridDim.x * kFloodStartNT) {
    const int i = base + threadIdx.x;
    int img = -1, t = 0;
    long long v = 0;
    bool f = false;
    if (i < n) {
      t = sel ? sel[i] : i;
      img = g.tile_img[t];
      v = s.tpos[t];
      f = gate && s.haspos[t] != 0;
      s.imp[t] = 0;
      s.fchg[t] = 0;
    }
    tl[threadIdx.x] = t;
    const int img0 = __shfl(img, 0);
    if (__all(img == img0 || img < 0)) {
      v = wave_sum64(v);
      if (lane_id() == 0 && img0 >= 0 && v)
        atomicAdd(&s.sums[(size_t)img0 * SUM_ROWS * kSumSlots + row], (unsigned long long)v);
    } else if (img >= 0 && v) {
      atomicAdd(&s.sums[(size_t)img * SUM_ROWS * kSumSlots + row], (unsigned long long)v);
    }
    const unsigned long long m = __ballot(f);
    int at = 0;
    if (lane_id() == 0 && m) at = atomicAdd(&counts[CNT_POS], __popcll(m));
    at = __shfl(at, 0);
    if (f) out[at + __popcll(m & below)] = t;
    __syncthreads();
    // the step's tiles' reach bytes, 16 per thread
    constexpr int V = T * T / 16;
    const int nv = min(kFloodStartNT, n - base) * V;
    for (int k = threadIdx.x; k < nv; k += kFloodStartNT)
      reinterpret_cast<uint4*>(s.reach + tbase<T>(tl[k / V]))[k % V] = make_uint4(0, 0, 0, 0);
    __syncthreads();  // tl is rewritten by the next step
  }
}

// The incremental flood's recompute set in one pass over the solve's tiles
// (sel, n of them): every tile flagged in fchg (the changed tiles and their
// import closure) gets its changed and import flags cleared, its reach
// bytes zeroed and lmark set (its labels are rewritten), and joins `out`,
// the first pass's list (count counts[CNT_FIRST], zero before), if it holds
// excess or belongs to an image reflowed warm (a fresh image's other tiles
// get reach from its own tiles, through the rounds).  Gated on *done when
// given (a device-driven solve floods only once it is done).  Block 0
// zeroes the queue counts of the flood's rounds.
constexpr int kFloodSetNT = 256;
template <int T>
__global__ void __launch_bounds__(kFloodSetNT)
k_flood_set(Geo g, CutDev s, const int32_t* __restrict__ sel, int n, int32_t* lmark,
            const int32_t* __restrict__ imgfresh, int32_t* out, int32_t* counts,
            const int32_t* __restrict__ done) {
  __shared__ int32_t tl[kFloodSetNT];
  __shared__ int ntl;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    counts[CNT_Q] = counts[CNT_Q + 1] = counts[CNT_Q + 2] = 0;
  if (done && *done == 0) n = 0;
  const unsigned long long below = (1ull << lane_id()) - 1;
  for (int base = blockIdx.x * kFloodSetNT; base < n; base += gridDim.x * kFloodSetNT) {
    if (threadIdx.x == 0) ntl = 0;
    __syncthreads();
    const int i = base + threadIdx.x;
    int t = 0;
    bool f = false, first = false;
    if (i < n) {
      t = sel ? sel[i] : i;
      f = s.fchg[t] != 0;
    }
    if (f) {
      s.fchg[t] = 0;
      s.imp[t] = 0;
      if (lmark) lmark[t] = 1;
      first = s.haspos[t] || !imgfresh[g.tile_img[t]];
      tl[atomicAdd(&ntl, 1)] = t;
    }
    const unsigned long long m = __ballot(first);
    int at = 0;
    if (lane_id() == 0 && m) at = atomicAdd(&counts[CNT_FIRST], __popcll(m));
    at = __shfl(at, 0);
    if (first) out[at + __popcll(m & below)] = t;
    __syncthreads();
    // the flagged tiles' reach bytes, 16 per thread
    constexpr int V = T * T / 16;
    const int nv = ntl * V;
    for (int k = threadIdx.x; k < nv; k += kFloodSetNT)
      reinterpret_cast<uint4*>(s.reach + tbase<T>(tl[k / V]))[k % V] = make_uint4(0, 0, 0, 0);
    __syncthreads();  // tl and ntl are rewritten by the next step
  }
}


template <int T, int NT>
struct PoolKernels {
  static void push(Pool* p, int grid, int sweeps, int kind) {
    if (kind != PUSH_LATER)
      hipLaunchKernelGGL((k_push<T, NT, true>), dim3(grid), dim3(NT), 0, p->stream, p->geo, p->dev,
                         p->lfresh, p->counts + CNT_FRESH, sweeps, kind == PUSH_FIRST ? 1 : 0);
    else
      hipLaunchKernelGGL((k_push<T, NT, false>), dim3(grid), dim3(NT), 0, p->stream, p->geo,
                         p->dev, p->alist, p->counts + CNT_ACT, sweeps, 1);
  }
  static void fold(Pool* p, int grid, const int32_t* list, int cnt) {
    hipLaunchKernelGGL((k_fold<T>), dim3(grid), dim3(kFoldNT<T>), 0, p->stream, p->geo, p->dev,
                       list, p->counts + cnt);
  }
  // (grid: the most tiles the pass can list; a block takes kBfsTPB tiles
  // at a time, and at most grid_cap blocks run)
  static int bfs_grid(Pool* p, int tiles) {
    return std::max(1, std::min((tiles + kBfsTPB<T> - 1) / kBfsTPB<T>, p->grid_cap));
  }
  static void bfs_init(Pool* p, int grid, const int32_t* list, int cnt, bool all,
                       const Queue& q) {
    grid = bfs_grid(p, grid);
    hipLaunchKernelGGL((k_bfs_init<T>), dim3(grid), dim3(kBfsNT), 0, p->stream, p->geo, p->dev,
                       list, p->counts + cnt, all ? 1 : 0, q, p->counts + CNT_ACT);
  }
  static void bfs_round(Pool* p, int grid, const Queue& q) {
    hipLaunchKernelGGL((k_bfs_round<T>), dim3(grid), dim3(kBfsNT), 0, p->stream, p->geo, p->dev,
                       q);
  }
  static void flood(Pool* p, int grid, const Queue& q, int mode, int32_t* lmark) {
    hipLaunchKernelGGL((k_flood<T, NT>), dim3(grid), dim3(kFloodNT), 0, p->stream, p->geo, p->dev, q,
                       mode, lmark);
  }
  // (tiles: the most tiles the list can hold)
  static void labels(Pool* p, int tiles, const int32_t* list, int cnt, int nflows) {
    const int grid =
        std::max(1, std::min((tiles + kLabelsTPB<T> - 1) / kLabelsTPB<T>, p->grid_cap));
    const FlowOut F{reinterpret_cast<const int32_t*>(p->d_prep), (int)(sizeof(Prep) / 4), nflows,
                    p->d_flows};
    hipLaunchKernelGGL((k_labels<T>), dim3(grid), dim3(kLabelsNT), 0, p->stream, p->geo, p->dev,
                       list, p->counts + cnt, F);
  }
  static void posf(Pool* p, int grid, unsigned zmask) {
    hipLaunchKernelGGL((k_posf<T>), dim3(grid), dim3(256), 0, p->stream, p->geo, p->dev, p->lsel,
                       p->counts + CNT_SEL, p->counts, zmask);
  }
  static void flood_start_full(Pool* p, int nsel, unsigned zmask, const int32_t* done) {
    const int grid = std::max(1, std::min((nsel + kFloodStartNT - 1) / kFloodStartNT, 2048));
    hipLaunchKernelGGL((k_flood_start_full<T>), dim3(grid), dim3(kFloodStartNT), 0, p->stream,
                       p->geo, p->dev, p->lsel, nsel, p->counts, zmask, p->plist, done);
  }
  static void flood_set(Pool* p, int nsel, int32_t* lmark, const int32_t* done) {
    const int grid = std::max(1, std::min((nsel + kFloodSetNT - 1) / kFloodSetNT, 2048));
    hipLaunchKernelGGL((k_flood_set<T>), dim3(grid), dim3(kFloodSetNT), 0, p->stream, p->geo,
                       p->dev, p->lsel, nsel, lmark, p->d_imgfresh, p->alist, p->counts, done);
  }
  static void edit(Pool* p, int img, int n, const int32_t* idx, const int32_t* cs,
                   const int32_t* ct) {
    const int grid = std::min((n + 255) / 256, 1024);
    hipLaunchKernelGGL((k_edit<T>), dim3(grid > 0 ? grid : 1), dim3(256), 0, p->stream, p->geo,
                       p->dev, img, n, idx, cs, ct, p->imgs[img].fresh ? 0 : 1);
  }
  static void paint(Pool* p, const PaintJob* d_jobs, int njobs, int nchunks) {
    // a wave per (stroke, segment, kPaintRows rows, piece of kPaintCols
    // columns), four per block
    const int grid = std::min((nchunks + 3) / 4, 8192);
    hipLaunchKernelGGL((k_paint<T>), dim3(grid > 0 ? grid : 1), dim3(256), 0, p->stream, p->geo,
                       p->dev, d_jobs, njobs, nchunks);
  }
  static void untile(Pool* p, int img, const int32_t* src, int32_t* out) {
    const PoolImage& im = p->imgs[img];
    const int PW = im.tw * T;
    hipLaunchKernelGGL((k_untile<T>), dim3((PW + 255) / 256, im.th * T), dim3(256), 0, p->stream,
                       p->geo, img, src, out);
  }
};

template <int T, int NT>
static const PoolOps* ops_for() {
  static const PoolOps ops = {T,
                              NT,
                              &PoolKernels<T, NT>::push,
                              &PoolKernels<T, NT>::fold,
                              &PoolKernels<T, NT>::bfs_init,
                              &PoolKernels<T, NT>::bfs_round,
                              &PoolKernels<T, NT>::flood,
                              &PoolKernels<T, NT>::labels,
                              &PoolKernels<T, NT>::posf,
                              &PoolKernels<T, NT>::flood_set,
                              &PoolKernels<T, NT>::flood_start_full,
                              &PoolKernels<T, NT>::edit,
                              &PoolKernels<T, NT>::paint,
                              &PoolKernels<T, NT>::untile};
  return &ops;
}

// Threads per tile for each tile size (a build-time tuning constant).
#ifndef GCMI_NT32
#define GCMI_NT32 256
#endif

const PoolOps* pool_ops(int tile) {
  switch (tile) {
    case 16: return ops_for<16, 256>();
    case 32: return ops_for<32, GCMI_NT32>();
    case 64: return ops_for<64, 1024>();
    default: return nullptr;
  }
}

// ------------------------------------------------------------- allocator

void RangeAlloc::init(long long c) {
  cap = c;
  free_.clear();
  if (c > 0) free_.push_back({0, c});
}

long long RangeAlloc::alloc(long long len) {
  for (size_t i = 0; i < free_.size(); ++i) {
    if (free_[i].second >= len) {
      const long long s = free_[i].first;
      free_[i].first += len;
      free_[i].second -= len;
      if (free_[i].second == 0) free_.erase(free_.begin() + i);
      return s;
    }
  }
  return -1;
}

void RangeAlloc::release(long long start, long long len) {
  if (len <= 0) return;
  auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(start, (long long)0));
  it = free_.insert(it, {start, len});
  size_t i = it - free_.begin();
  if (i + 1 < free_.size() && free_[i].first + free_[i].second == free_[i + 1].first) {
    free_[i].second += free_[i + 1].second;
    free_.erase(free_.begin() + i + 1);
  }
  if (i > 0 && free_[i - 1].first + free_[i - 1].second == free_[i].first) {
    free_[i - 1].second += free_[i].second;
    free_.erase(free_.begin() + i);
  }
}

long long RangeAlloc::largest() const {
  long long m = 0;
  for (auto& f : free_) m = std::max(m, f.second);
  return m;
}

// ------------------------------------------------------------- pool

int tiles_for(int W, int H, int tile, int* tw, int* th) {
  if (W < 1 || H < 1) {
    set_error("width and height must be >= 1 (got %d x %d)", W, H);
    return GCMI_EINVAL;
  }
  *tw = (W + tile - 1) / tile;
  *th = (H + tile - 1) / tile;
  if (*tw > 0xffff || *th > 0xffff) {
    set_error("an image is at most 65535 tiles across and down (got %d x %d)", *tw, *th);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

size_t pool_bytes(int tile, int cap_tiles, long long cap_pix, int cap_imgs) {
  const size_t np = (size_t)cap_tiles * tile * tile;
  return np * (6 * 4 + 2) + (size_t)cap_tiles * (4 * tile * 4 + 16 * 4 + 8 + 32) +
         (size_t)cap_pix * (4 * 4 + 1) + (size_t)cap_pix / 8 +
         (size_t)cap_imgs *
             (sizeof(ImgDesc) + SUM_ROWS * kSumSlots * 8 + sizeof(Prep) + 8 + kPixAlign * 17) +
         (size_t)compact_scratch_ints(cap_tiles) * 4 + 8 * kSumSlots * kRgbSumStride * kRgbJobsMax +
         4096;
}

// Synchronises the pool's stream: every staging copy queued on it is done.
static int stream_sync(Pool* p) {
  const auto t0 = std::chrono::steady_clock::now();
  GCMI_HIP(hipStreamSynchronize(p->stream));
  p->ns_wait += std::chrono::duration_cast<std::chrono::nanoseconds>(
                    std::chrono::steady_clock::now() - t0).count();
  for (bool& b : p->stage_pending) b = false;
  return GCMI_OK;
}

// Before staging buffer k is rewritten: waits for the stream only if a copy
// from it was queued since the last synchronisation.
static int stage_wait(Pool* p, int k) {
  return p->stage_pending[k] ? stream_sync(p) : GCMI_OK;
}

static int dev_alloc(void** ptr, size_t bytes, size_t* total) {
  if (hipMalloc(ptr, bytes ? bytes : 4) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of %zu bytes failed", bytes);
    return GCMI_ENOMEM;
  }
  *total += bytes;
  return GCMI_OK;
}

int pool_init(Pool* p, int tile, int cap_tiles, long long cap_pix, int cap_imgs,
              hipStream_t stream) {
  // room for every image's range to start on a multiple of kPixAlign
  cap_pix = (cap_pix + kPixAlign - 1) / kPixAlign * kPixAlign + (long long)kPixAlign * cap_imgs;
  p->ops = pool_ops(tile);
  if (!p->ops) {
    set_error("tile must be 16, 32 or 64 (got %d)", tile);
    return GCMI_EINVAL;
  }
  if (cap_tiles < 1 || cap_pix < 1 || cap_imgs < 1) {
    set_error("empty pool");
    return GCMI_EINVAL;
  }
  const long long np = (long long)cap_tiles * tile * tile;
  if (np >= (1LL << 30)) {
    set_error("pool of %lld padded pixels exceeds 2^30", np);
    return GCMI_EINVAL;
  }
  p->T = tile;
  p->cap_tiles = cap_tiles;
  p->cap_pix = cap_pix;
  p->cap_imgs = cap_imgs;
  if (stream) {
    p->stream = stream;
  } else {
    GCMI_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->own_stream = true;
  }
  int dev = 0, ncu = 0;
  GCMI_HIP(hipGetDevice(&dev));
  GCMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  p->grid_cap = ncu > 0 ? ncu * 8 : 2048;
  {
    const char* tr = getenv("GCMI_TRACE");
    p->trace = tr && tr[0] == '1';
  }
  size_t tot = 0;
  int rc = GCMI_OK;
  CutDev& d = p->dev;
#define ALLOC(ptr, bytes) \
  if (rc == GCMI_OK) rc = dev_alloc((void**)&(ptr), (bytes), &tot)
  ALLOC(d.e, np * 4);
  ALLOC(d.h, np * 4);
  ALLOC(d.cu, np * 4);
  ALLOC(d.cd, np * 4);
  ALLOC(d.cl, np * 4);
  ALLOC(d.cr, np * 4);
  ALLOC(d.mask, np);
  ALLOC(d.reach, np);
  ALLOC(d.pend, (size_t)cap_tiles * 4 * tile * 4);
  ALLOC(d.pflag, (size_t)cap_tiles * 4);
  ALLOC(d.act, (size_t)cap_tiles * 4);
  ALLOC(d.stamp, (size_t)cap_tiles * 4);
  ALLOC(d.bflag, (size_t)cap_tiles * 4);
  ALLOC(d.dirty, (size_t)cap_tiles * 4);
  ALLOC(d.rflag, (size_t)cap_tiles * 4);
  ALLOC(d.haspos, (size_t)cap_tiles * 4);
  ALLOC(d.packed, (size_t)cap_pix / 8);
  ALLOC(d.fchg, (size_t)cap_tiles * 4);
  ALLOC(d.imp, (size_t)cap_tiles * 4);
  ALLOC(d.tpos, (size_t)cap_tiles * 8);
  ALLOC(d.sums, (size_t)cap_imgs * SUM_ROWS * kSumSlots * 8);
  ALLOC(d.cs, (size_t)cap_pix * 4);
  ALLOC(d.ct, (size_t)cap_pix * 4);
  int32_t* right = nullptr;
  int32_t* down = nullptr;
  ALLOC(right, (size_t)cap_pix * 4);
  ALLOC(down, (size_t)cap_pix * 4);
  d.right = right;
  d.down = down;
  ALLOC(d.labels, (size_t)cap_pix);
  ALLOC(p->d_img, (size_t)cap_imgs * sizeof(ImgDesc));
  ALLOC(p->tile_img, (size_t)cap_tiles * 4);
  ALLOC(p->tgeo, (size_t)cap_tiles * 32);
  ALLOC(p->sel, (size_t)cap_tiles * 4);
  ALLOC(p->fresh, (size_t)cap_tiles * 4);
  ALLOC(p->q, (size_t)cap_tiles * 3 * 4);
  ALLOC(p->dlist, (size_t)cap_tiles * 4);
  ALLOC(p->plist, (size_t)cap_tiles * 4);
  ALLOC(p->alist, (size_t)cap_tiles * 4);
  ALLOC(p->d_status, kStatusFlows + (size_t)cap_imgs * 8);
  ALLOC(p->d_rgbsum, 8 * kSumSlots * kRgbSumStride * kRgbJobsMax);
  ALLOC(p->d_imgfresh, (size_t)cap_imgs * 4);
  ALLOC(p->scratch, (size_t)compact_scratch_ints(cap_tiles) * 4);
  ALLOC(p->d_prep, (size_t)cap_imgs * sizeof(Prep));
#undef ALLOC
  if (rc == GCMI_OK &&
      (hipHostMalloc((void**)&p->h_status, kStatusFlows + (size_t)cap_imgs * 8,
                     hipHostMallocDefault) != hipSuccess ||
       hipHostMalloc((void**)&p->h_prep, (size_t)cap_imgs * sizeof(Prep), hipHostMallocDefault) !=
           hipSuccess)) {
    (void)hipGetLastError();
    set_error("pinned host allocation failed");
    rc = GCMI_ENOMEM;
  }
  if (p->d_status) {
    p->counts = reinterpret_cast<int32_t*>(p->d_status);
    d.tally = reinterpret_cast<unsigned long long*>(p->d_status + kStatusTally);
    p->d_flows = reinterpret_cast<long long*>(p->d_status + kStatusFlows);
  }
  if (p->h_status) {
    p->h_counts = reinterpret_cast<int32_t*>(p->h_status);
    p->h_tally = reinterpret_cast<unsigned long long*>(p->h_status + kStatusTally);
    p->h_flows = reinterpret_cast<long long*>(p->h_status + kStatusFlows);
  }
  p->device_bytes = tot;
  if (rc != GCMI_OK) {
    pool_free(p);
    return rc;
  }
  p->geo.maxh = (int32_t)np;
  p->geo.ntiles = cap_tiles;
  p->geo.img = p->d_img;
  p->geo.tile_img = p->tile_img;
  p->geo.tgeo = p->tgeo;
  p->imgs.assign(cap_imgs, PoolImage());
  p->tiles_alloc.init(cap_tiles);
  p->pix_alloc.init(cap_pix);
  // pending slots and flags start (and stay, between solves) zero
  GCMI_HIP(hipMemsetAsync(d.pend, 0, (size_t)cap_tiles * 4 * tile * 4, p->stream));
  int32_t* flags[] = {d.pflag, d.act, d.stamp, d.bflag, d.dirty, d.rflag, d.haspos, d.fchg, d.imp};
  for (int32_t* f : flags) GCMI_HIP(hipMemsetAsync(f, 0, (size_t)cap_tiles * 4, p->stream));
  GCMI_HIP(hipMemsetAsync(d.tpos, 0, (size_t)cap_tiles * 8, p->stream));
  GCMI_HIP(hipMemsetAsync(p->d_status, 0, kStatusFlows, p->stream));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

void pool_free(Pool* p) {
  CutDev& d = p->dev;
  void* ptrs[] = {d.e,     d.h,       d.cu,     d.cd,       d.cl,         d.cr,       d.mask,
                  d.reach, d.pend,    d.pflag,  d.act,      d.stamp,      d.bflag,    d.dirty,
                  d.rflag, d.haspos,  d.fchg,   d.imp,      d.packed,
                  d.tpos,  d.sums,    d.cs,     d.ct,       (void*)d.right,
                  (void*)d.down,      d.labels, p->d_img,   p->tile_img,  p->sel,     p->fresh,
                  p->q,    p->dlist,  p->plist, p->alist,   p->d_status,  p->scratch, p->d_prep,
                  p->tgeo,
                  p->d_edit, p->d_rgbsum, p->d_rgbjobs, p->d_rgb, p->d_jobs, p->d_imgfresh};
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  p->graph_exec = nullptr;
  for (void* x : ptrs)
    if (x) (void)hipFree(x);
  void* hp[] = {p->h_status, p->h_prep, p->h_edit, p->h_jobs, p->h_rgbjobs};
  for (void* x : hp)
    if (x) (void)hipHostFree(x);
  p->prof.release();
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  *p = Pool();
}

int pool_add_image(Pool* p, int W, int H, int* out) {
  int tw = 0, th = 0;
  GCMI_TRY(tiles_for(W, H, p->T, &tw, &th));
  int id = -1;
  for (int i = 0; i < p->cap_imgs; ++i)
    if (!p->imgs[i].live) {
      id = i;
      break;
    }
  if (id < 0) {
    set_error("pool holds %d images already", p->cap_imgs);
    return GCMI_EBUDGET;
  }
  const long long nt = (long long)tw * th;
  const long long t0 = p->tiles_alloc.alloc(nt);
  if (t0 < 0) {
    set_error("no room for %lld tiles (largest free run %lld)", nt, p->tiles_alloc.largest());
    return GCMI_EBUDGET;
  }
  const long long p0 = p->pix_alloc.alloc(pix_len(W, H));
  if (p0 < 0) {
    p->tiles_alloc.release(t0, nt);
    set_error("no room for %lld pixels", (long long)W * H);
    return GCMI_EBUDGET;
  }
  PoolImage& im = p->imgs[id];
  im = PoolImage();
  im.live = true;
  im.W = W;
  im.H = H;
  im.tw = tw;
  im.th = th;
  im.tile0 = (int)t0;
  im.ntiles = (int)nt;
  im.pix0 = p0;
  ImgDesc d{W, H, tw, th, (int)t0, (int)nt, p0};
  GCMI_HIP(hipMemcpyAsync(p->d_img + id, &d, sizeof d, hipMemcpyHostToDevice, p->stream));
  // a clean packed mask (its last word's bits past W * H stay 0)
  const long long words = ((long long)W * H + 31) / 32;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<long long>((words + 255) / 256, 4096)),
                     dim3(256), 0, p->stream, (int32_t*)(p->dev.packed + p0 / 32), words, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<long long>((nt + 255) / 256, 4096)),
                     dim3(256), 0, p->stream, p->tile_img + t0, nt, (int32_t)id);
  hipLaunchKernelGGL(k_tgeo_fill, dim3((unsigned)std::min<long long>((nt + 255) / 256, 4096)),
                     dim3(256), 0, p->stream, p->tgeo, d, id);
  GCMI_HIP(hipGetLastError());
  GCMI_HIP(hipStreamSynchronize(p->stream));
  *out = id;
  return GCMI_OK;
}

int pool_remove_image(Pool* p, int img) {
  if (img < 0 || img >= p->cap_imgs || !p->imgs[img].live) {
    set_error("no image %d", img);
    return GCMI_EINVAL;
  }
  PoolImage& im = p->imgs[img];
  p->tiles_alloc.release(im.tile0, im.ntiles);
  p->pix_alloc.release(im.pix0, pix_len(im.W, im.H));
  im = PoolImage();
  return GCMI_OK;
}

int validate_caps(const int32_t* cs, const int32_t* ct, const int32_t* right,
                  const int32_t* down, size_t n, int32_t* max_pair) {
  int64_t mt = 0, mw = 0;
  for (size_t i = 0; i < n; ++i) {
    if (cs[i] < 0 || ct[i] < 0 || right[i] < 0 || down[i] < 0) {
      set_error("capacities must be >= 0 (pixel %zu)", i);
      return GCMI_EINVAL;
    }
    mt = std::max<int64_t>(mt, std::max(cs[i], ct[i]));
    mw = std::max<int64_t>(mw, std::max(right[i], down[i]));
  }
  if (mt + 4 * mw >= (int64_t(1) << 31)) {
    set_error("max(source,sink) + 4*max(pairwise) = %lld must be < 2^31",
              (long long)(mt + 4 * mw));
    return GCMI_EINVAL;
  }
  *max_pair = (int32_t)mw;
  return GCMI_OK;
}

static int check_img(Pool* p, int img) {
  if (img < 0 || img >= p->cap_imgs || !p->imgs[img].live) {
    set_error("no image %d", img);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

int pool_upload(Pool* p, int img, const int32_t* cs, const int32_t* ct, const int32_t* right,
                const int32_t* down, bool device_src) {
  p->step_img = -1;  // drops a stepped solve in progress
  GCMI_TRY(check_img(p, img));
  PoolImage& im = p->imgs[img];
  const size_t n = (size_t)im.W * im.H;
  if (!device_src) GCMI_TRY(validate_caps(cs, ct, right, down, n, &im.max_pair));
  const hipMemcpyKind kind = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  int32_t* dst[4] = {p->dev.cs, p->dev.ct, (int32_t*)p->dev.right, (int32_t*)p->dev.down};
  const int32_t* src[4] = {cs, ct, right, down};
  for (int k = 0; k < 4; ++k)
    GCMI_HIP(hipMemcpyAsync(dst[k] + im.pix0, src[k], n * 4, kind, p->stream));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  if (device_src) im.max_pair = -1;  // unknown: edits re-check against 2^29
  im.uploaded = true;
  im.fresh = true;
  im.solved = false;
  return GCMI_OK;
}

int pool_capacities(Pool* p, int img, int32_t* cs, int32_t* ct, int32_t* right, int32_t* down) {
  GCMI_TRY(check_img(p, img));
  const PoolImage& im = p->imgs[img];
  if (!im.uploaded) {
    set_error("image %d has no capacities", img);
    return GCMI_EINVAL;
  }
  const size_t n = (size_t)im.W * im.H;
  const int32_t* src[4] = {p->dev.cs, p->dev.ct, (const int32_t*)p->dev.right,
                           (const int32_t*)p->dev.down};
  int32_t* dst[4] = {cs, ct, right, down};
  for (int k = 0; k < 4; ++k)
    GCMI_HIP(hipMemcpyAsync(dst[k], src[k] + im.pix0, n * 4, hipMemcpyDeviceToHost, p->stream));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

int check_rgb_model(const gcmi_rgb_model* m, int32_t* max_pair) {
  if (!m) {
    set_error("null colour model");
    return GCMI_EINVAL;
  }
  bool ok = m->pair_kind == GCMI_PAIR_EXP || m->pair_kind == GCMI_PAIR_INV;
  const double v[] = {m->mu_bg[0], m->mu_bg[1], m->mu_bg[2], m->mu_fg[0], m->mu_fg[1],
                      m->mu_fg[2], m->unary_scale, m->unary_div, m->unary_cap, m->pair_scale,
                      m->inv_beta, m->inv_eps};
  for (double x : v) ok = ok && std::isfinite(x);
  ok = ok && m->unary_scale >= 0 && m->unary_div > 0 && m->unary_cap >= 0 &&
       m->unary_scale * m->unary_cap <= double(1 << 26) && m->pair_scale >= 0;
  double wmax = 0;
  if (ok && m->pair_kind == GCMI_PAIR_EXP) wmax = m->pair_scale;
  if (ok && m->pair_kind == GCMI_PAIR_INV) {
    ok = m->inv_beta >= 0 && m->inv_eps > 0;
    if (ok) wmax = m->pair_scale / m->inv_eps;
  }
  ok = ok && wmax <= double(1 << 24);
  if (!ok) {
    set_error("colour model out of range");
    return GCMI_EINVAL;
  }
  if (max_pair) *max_pair = (int32_t)std::ceil(wmax + 0.5);
  return GCMI_OK;
}

int pool_upload_rgb_many(Pool* p, int count, const RgbReq* reqs) {
  p->step_img = -1;  // drops a stepped solve in progress
  if (count <= 0) return GCMI_OK;
  static_assert(kPixAlign % 4 == 0, "colour capacities are stored four pixels at a time");
  std::vector<int32_t> wmax(count);
  for (int i = 0; i < count; ++i) {
    const RgbReq& r = reqs[i];
    GCMI_TRY(check_img(p, r.img));
    if (r.channels != 3 && r.channels != 4) {
      set_error("a colour image has 3 or 4 channels (got %d)", r.channels);
      return GCMI_EINVAL;
    }
    GCMI_TRY(check_rgb_model(r.model, &wmax[i]));
    if (!r.d_rgb || reinterpret_cast<uintptr_t>(r.d_rgb) % 4 != 0) {
      set_error("a colour image on the device must start 4-byte aligned");
      return GCMI_EINVAL;
    }
    for (int k = 0; k < i; ++k)
      if (reqs[k].img == r.img) {
        set_error("image %d has two colour uploads in one call", r.img);
        return GCMI_EINVAL;
      }
  }
  // staging: the job table of kRgbJobsMax images (allocated on first use)
  if (!p->h_rgbjobs) {
    GCMI_HIP(hipStreamSynchronize(p->stream));
    if (hipHostMalloc((void**)&p->h_rgbjobs, sizeof(RgbJobs), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&p->d_rgbjobs, sizeof(RgbJobs)) != hipSuccess) {
      (void)hipGetLastError();
      set_error("colour staging allocation failed");
      return GCMI_ENOMEM;
    }
  }
  RgbJobs* hj = reinterpret_cast<RgbJobs*>(p->h_rgbjobs);
  const RgbJobs* dj = reinterpret_cast<const RgbJobs*>(p->d_rgbjobs);
  for (int base = 0; base < count; base += kRgbJobsMax) {
    const int m = std::min(count - base, kRgbJobsMax);
    // the job table may still be read by the previous launch's copy
    GCMI_TRY(stage_wait(p, STAGE_RGBJOBS));
    long long blocks[2] = {0, 0};
    for (int i = 0; i < m; ++i) {
      const RgbReq& r = reqs[base + i];
      const PoolImage& im = p->imgs[r.img];
      hj->job[i] = RgbJob{r.d_rgb, im.pix0, im.W, im.H, r.channels, {0, 0}, *r.model};
      for (int k = 0; k < 2; ++k) {
        const long long px_per_block =
            (long long)(k == RGB_SUM ? kRgbSumChunks : kRgbCapsChunks) * kRgbBlockPx;
        hj->block0[k][i] = hj->job[i].block0[k] = (int)blocks[k];
        blocks[k] += ((long long)im.W * im.H + px_per_block - 1) / px_per_block;
      }
      if (blocks[RGB_CAPS] > 0x7fffffffLL) {
        set_error("colour images too large for one launch");
        return GCMI_EINVAL;
      }
    }
    GCMI_HIP(hipMemcpyAsync(p->d_rgbjobs, hj, offsetof(RgbJobs, job) + (size_t)m * sizeof(RgbJob),
                            hipMemcpyHostToDevice, p->stream));
    p->stage_pending[STAGE_RGBJOBS] = true;
    hipLaunchKernelGGL(k_fill_i32, dim3(std::max(1, m * kSumSlots * kRgbSumStride * 2 / 256)),
                       dim3(256), 0, p->stream, (int32_t*)p->d_rgbsum,
                       2LL * kSumSlots * kRgbSumStride * m, 0);
    if (blocks[RGB_CAPS] > 0) {
      hipLaunchKernelGGL(k_rgb_pair_sum, dim3((unsigned)blocks[RGB_SUM]), dim3(kRgbNT), 0,
                         p->stream, dj, m, p->d_rgbsum);
      hipLaunchKernelGGL(k_rgb_caps, dim3((unsigned)blocks[RGB_CAPS]), dim3(kRgbNT), 0, p->stream,
                         dj, m,
                         (const unsigned long long*)p->d_rgbsum, p->dev.cs, p->dev.ct,
                         (int32_t*)p->dev.right, (int32_t*)p->dev.down);
    }
    GCMI_HIP(hipGetLastError());
    for (int i = 0; i < m; ++i) {
      PoolImage& im = p->imgs[reqs[base + i].img];
      im.max_pair = wmax[base + i];
      im.uploaded = true;
      im.fresh = true;
      im.solved = false;
    }
  }
  return GCMI_OK;
}

int pool_upload_rgb(Pool* p, int img, const uint8_t* d_rgb, int channels,
                    const gcmi_rgb_model* m) {
  const RgbReq r{img, d_rgb, channels, m};
  return pool_upload_rgb_many(p, 1, &r);
}

int pool_stage_rgb(Pool* p, int img, const uint8_t* h_rgb, int channels,
                   const uint8_t** d_out) {
  GCMI_TRY(check_img(p, img));
  if (!h_rgb) {
    set_error("null colour image");
    return GCMI_EINVAL;
  }
  if (channels != 3 && channels != 4) {
    set_error("a colour image has 3 or 4 channels (got %d)", channels);
    return GCMI_EINVAL;
  }
  const PoolImage& im = p->imgs[img];
  const long long bytes = (long long)im.W * im.H * channels;
  if (p->rgb_cap < bytes) {
    GCMI_HIP(hipStreamSynchronize(p->stream));
    if (p->d_rgb) (void)hipFree(p->d_rgb);
    p->d_rgb = nullptr;
    p->rgb_cap = 0;
    if (hipMalloc((void**)&p->d_rgb, (size_t)bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %lld bytes failed", bytes);
      return GCMI_ENOMEM;
    }
    p->rgb_cap = bytes;
  }
  GCMI_HIP(hipMemcpyAsync(p->d_rgb, h_rgb, (size_t)bytes, hipMemcpyHostToDevice, p->stream));
  *d_out = p->d_rgb;
  return GCMI_OK;
}

int pool_upload_rgb_host(Pool* p, int img, const uint8_t* h_rgb, int channels,
                         const gcmi_rgb_model* m) {
  GCMI_TRY(check_rgb_model(m, nullptr));
  const uint8_t* d = nullptr;
  GCMI_TRY(pool_stage_rgb(p, img, h_rgb, channels, &d));
  GCMI_TRY(pool_upload_rgb(p, img, d, channels, m));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

int check_edit(long long npx, int n, const int32_t* idx, const int32_t* cs, const int32_t* ct,
               int64_t wmax, std::vector<uint64_t>& mark) {
  if (n < 0 || (n > 0 && (!idx || !cs || !ct))) {
    set_error("bad edit arguments");
    return GCMI_EINVAL;
  }
  if ((long long)mark.size() * 64 < npx) mark.assign((size_t)((npx + 63) / 64), 0);
  int bad = -1, kind = 0;
  int i = 0;
  for (; i < n; ++i) {
    const int32_t q = idx[i];
    if (q < 0 || q >= npx) {
      bad = i;
      kind = 0;
      break;
    }
    uint64_t& w = mark[q >> 6];
    const uint64_t b = uint64_t(1) << (q & 63);
    if (w & b) {
      bad = i;
      kind = 0;
      break;
    }
    w |= b;
    if (cs[i] < 0 || ct[i] < 0 || std::max<int64_t>(cs[i], ct[i]) + 4 * wmax >= (1LL << 31)) {
      bad = i;
      kind = 1;
      ++i;
      break;
    }
  }
  // clear the bits this call set (the first i indices, bad one excluded when
  // it was a duplicate or out of range)
  const int set = (bad >= 0 && kind == 0) ? bad : i;
  for (int k = 0; k < set; ++k) mark[idx[k] >> 6] &= ~(uint64_t(1) << (idx[k] & 63));
  if (bad >= 0) {
    if (kind == 0)
      set_error("edit indices must be distinct and in [0, %lld) (entry %d)", npx, bad);
    else
      set_error("edit %d: capacities out of range", bad);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

int pool_edit_many(Pool* p, int count, const EditReq* e) {
  p->step_img = -1;  // drops a stepped solve in progress
  long long total = 0;
  for (int k = 0; k < count; ++k) {
    GCMI_TRY(check_img(p, e[k].img));
    const PoolImage& im = p->imgs[e[k].img];
    if (!im.uploaded) {
      set_error("edit before upload");
      return GCMI_EINVAL;
    }
    const int64_t wmax = im.max_pair >= 0 ? im.max_pair : (1 << 27);
    GCMI_TRY(check_edit((long long)im.W * im.H, e[k].n, e[k].idx, e[k].cs, e[k].ct, wmax,
                        p->mark));
    total += e[k].n;
  }
  if (total == 0) return GCMI_OK;
  if (total > (1LL << 30)) {
    set_error("edit of %lld pixels too large", total);
    return GCMI_EINVAL;
  }
  if (p->edit_cap < total) {
    // the old staging may still be read by an earlier copy
    GCMI_HIP(hipStreamSynchronize(p->stream));
    if (p->d_edit) (void)hipFree(p->d_edit);
    if (p->h_edit) (void)hipHostFree(p->h_edit);
    p->d_edit = nullptr;
    p->h_edit = nullptr;
    p->edit_cap = 0;
    const long long cap = std::max<long long>(total, 1 << 16);
    if (hipMalloc((void**)&p->d_edit, (size_t)cap * 12) != hipSuccess ||
        hipHostMalloc((void**)&p->h_edit, (size_t)cap * 12, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      set_error("edit staging allocation failed");
      return GCMI_ENOMEM;
    }
    p->edit_cap = cap;
  } else {
    // the staging buffer is reused: wait for the previous edit's copy
    GCMI_TRY(stage_wait(p, STAGE_EDIT));
  }
  long long off = 0;
  for (int k = 0; k < count; ++k) {
    const int n = e[k].n;
    memcpy(p->h_edit + 3 * off, e[k].idx, (size_t)n * 4);
    memcpy(p->h_edit + 3 * off + n, e[k].cs, (size_t)n * 4);
    memcpy(p->h_edit + 3 * off + 2 * n, e[k].ct, (size_t)n * 4);
    off += n;
  }
  GCMI_HIP(hipMemcpyAsync(p->d_edit, p->h_edit, (size_t)total * 12, hipMemcpyHostToDevice,
                          p->stream));
  p->stage_pending[STAGE_EDIT] = true;
  off = 0;
  for (int k = 0; k < count; ++k) {
    const int n = e[k].n;
    PoolImage& im = p->imgs[e[k].img];
    if (n > 0) {
      int32_t* d = p->d_edit + 3 * off;
      p->ops->edit(p, e[k].img, n, d, d + n, d + 2 * n);
      GCMI_HIP(hipGetLastError());
      if (!im.fresh) im.edited = true;
      im.solved = false;
    }
    off += n;
  }
  return GCMI_OK;
}

int pool_edit(Pool* p, int img, int n, const int32_t* idx, const int32_t* cs, const int32_t* ct) {
  const EditReq e{img, n, idx, cs, ct};
  return pool_edit_many(p, 1, &e);
}

int make_polyline(int W, int H, int npts, const double* xy, double width, int32_t cs, int32_t ct,
                  int64_t wmax, Polyline* out) {
  if (npts < 0 || npts > kMaxStrokePoints || (npts > 0 && !xy)) {
    set_error("a stroke has 0..%d points (got %d)", kMaxStrokePoints, npts);
    return GCMI_EINVAL;
  }
  if (!(width >= 0.0) || !std::isfinite(width) || width > 1e6) {
    set_error("stroke width must be finite and in [0, 1e6]");
    return GCMI_EINVAL;
  }
  if (cs < 0 || ct < 0 || std::max<int64_t>(cs, ct) + 4 * wmax >= (1LL << 31)) {
    set_error("stroke capacities out of range");
    return GCMI_EINVAL;
  }
  Polyline& pl = *out;
  pl = Polyline();
  pl.npts = npts;
  pl.r = width / 2.0;
  pl.cs = cs;
  pl.ct = ct;
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (int k = 0; k < npts; ++k) {
    const double x = xy[2 * k], y = xy[2 * k + 1];
    if (!std::isfinite(x) || !std::isfinite(y) || std::fabs(x) > 1e9 || std::fabs(y) > 1e9) {
      set_error("stroke point %d is not finite or too far out", k);
      return GCMI_EINVAL;
    }
    pl.x[k] = x;
    pl.y[k] = y;
    x0 = std::min(x0, x);
    x1 = std::max(x1, x);
    y0 = std::min(y0, y);
    y1 = std::max(y1, y);
  }
  // pixels farther than r + 1 from every point's box cannot be hit
  const long long bx0 = std::max<long long>((long long)std::floor(x0 - pl.r), 0);
  const long long bx1 = std::min<long long>((long long)std::ceil(x1 + pl.r), W - 1);
  const long long by0 = std::max<long long>((long long)std::floor(y0 - pl.r), 0);
  const long long by1 = std::min<long long>((long long)std::ceil(y1 + pl.r), H - 1);
  if (npts < 2 || bx0 > bx1 || by0 > by1) {
    pl.bw = pl.bh = 0;
    return GCMI_OK;
  }
  pl.bx0 = (int)bx0;
  pl.by0 = (int)by0;
  pl.bw = (int)(bx1 - bx0 + 1);
  pl.bh = (int)(by1 - by0 + 1);
  pl.W = W;
  // rows each segment's capsule can reach (as the box, per segment), each
  // in pieces of kPaintCols columns covering the box's width
  long long total = 0;
  for (int k = 0; k + 1 < npts; ++k) {
    const double ylo = std::min(pl.y[k], pl.y[k + 1]) - pl.r;
    const double yhi = std::max(pl.y[k], pl.y[k + 1]) + pl.r;
    const double xlo = std::min(pl.x[k], pl.x[k + 1]) - pl.r;
    const double xhi = std::max(pl.x[k], pl.x[k + 1]) + pl.r;
    const long long ya = std::max<long long>((long long)std::floor(ylo), 0);
    const long long yb = std::min<long long>((long long)std::ceil(yhi), H - 1);
    const bool cols = (long long)std::ceil(xhi) >= 0 && (long long)std::floor(xlo) <= W - 1;
    const long long rows = (ya <= yb && cols) ? yb - ya + 1 : 0;
    pl.ya[k] = rows > 0 ? (int)ya : 0;
    pl.nr[k] = (int)rows;
    // a row's candidate columns: the x extent of the segment's part within
    // r + 1 rows of it, widened by r + 1 each way, plus rounding; at most
    // the box's width
    const long long xa = std::max<long long>((long long)std::floor(xlo) - 1, 0);
    const long long xb = std::min<long long>((long long)std::ceil(xhi) + 1, W - 1);
    const double r1 = pl.r + 1.0;
    const double adx = std::fabs(pl.x[k + 1] - pl.x[k]), ady = std::fabs(pl.y[k + 1] - pl.y[k]);
    const double span = (ady > 0.0 ? adx * std::min(1.0, 2.0 * r1 / ady) : adx) + 2.0 * r1 + 4.0;
    const double width = std::min(span, (double)(xb - xa + 1));
    pl.xc[k] = std::max(1, (int)std::ceil(width / kPaintCols));
    // the capsule's pixel box, one pixel wider each way than needed
    pl.sbox[k] = make_int4((int)std::max<long long>((long long)std::floor(xlo) - 1, -1),
                           (int)std::min<long long>((long long)std::ceil(xhi) + 1, W),
                           (int)std::max<long long>((long long)std::floor(ylo) - 1, -1),
                           (int)std::min<long long>((long long)std::ceil(yhi) + 1, H));
    pl.c0[k] = (int)std::min<long long>(total, INT32_MAX);
    total += (rows + kPaintRows - 1) / kPaintRows * pl.xc[k];
  }
  if (total > INT32_MAX) {
    set_error("stroke spans too many rows");
    return GCMI_EINVAL;
  }
  pl.nchunks = (int)total;
  return GCMI_OK;
}

int pool_paint(Pool* p, int img, int npts, const double* xy, double width, int32_t cs,
               int32_t ct) {
  const PaintReq r{img, npts, xy, width, cs, ct};
  return pool_paint_many(p, 1, &r);
}

int pool_paint_many(Pool* p, int count, const PaintReq* reqs) {
  p->step_img = -1;  // drops a stepped solve in progress
  if (count <= 0) return GCMI_OK;
  if (!reqs) {
    set_error("null stroke list");
    return GCMI_EINVAL;
  }
  // check every stroke first, so that a refused one paints nothing
  std::vector<Polyline> pls((size_t)count);
  for (int i = 0; i < count; ++i) {
    const PaintReq& r = reqs[i];
    GCMI_TRY(check_img(p, r.img));
    const PoolImage& im = p->imgs[r.img];
    if (!im.uploaded) {
      set_error("stroke before upload");
      return GCMI_EINVAL;
    }
    GCMI_TRY(make_polyline(im.W, im.H, r.npts, r.xy, r.width, r.cs, r.ct,
                           im.max_pair >= 0 ? im.max_pair : (1 << 27), &pls[i]));
  }
  if (p->jobs_cap < count) {
    GCMI_HIP(hipStreamSynchronize(p->stream));
    if (p->h_jobs) (void)hipHostFree(p->h_jobs);
    if (p->d_jobs) (void)hipFree(p->d_jobs);
    p->h_jobs = nullptr;
    p->d_jobs = nullptr;
    p->jobs_cap = 0;
    const int cap = std::max(count, 64);
    if (hipHostMalloc((void**)&p->h_jobs, (size_t)cap * sizeof(PaintJob), hipHostMallocDefault) !=
            hipSuccess ||
        hipMalloc((void**)&p->d_jobs, (size_t)cap * sizeof(PaintJob)) != hipSuccess) {
      (void)hipGetLastError();
      set_error("paint staging allocation failed");
      return GCMI_ENOMEM;
    }
    p->jobs_cap = cap;
  }
  // the staging may still be read by a previous launch's copy
  GCMI_TRY(stage_wait(p, STAGE_JOBS));
  std::vector<char> seen((size_t)p->cap_imgs, 0);
  int m = 0, base = 0;
  long long rows = 0;
  auto launch = [&]() -> int {
    if (m == base) return GCMI_OK;
    GCMI_HIP(hipMemcpyAsync(p->d_jobs + base, p->h_jobs + base, (size_t)(m - base) * sizeof(PaintJob),
                            hipMemcpyHostToDevice, p->stream));
    p->stage_pending[STAGE_JOBS] = true;
    p->ops->paint(p, p->d_jobs + base, m - base, (int)rows);
    GCMI_HIP(hipGetLastError());
    std::fill(seen.begin(), seen.end(), 0);
    base = m;
    rows = 0;
    return GCMI_OK;
  };
  for (int i = 0; i < count; ++i) {
    const Polyline& pl = pls[i];
    if (pl.bw == 0 || pl.nchunks == 0) continue;
    // one launch paints each image once; chunks are counted in int
    if (seen[reqs[i].img] || rows + pl.nchunks > INT32_MAX) GCMI_TRY(launch());
    seen[reqs[i].img] = 1;
    PoolImage& im = p->imgs[reqs[i].img];
    PaintJob& J = p->h_jobs[m++];
    J.pl = pl;
    J.img = reqs[i].img;
    J.warm = im.fresh ? 0 : 1;
    J.chunk0 = (int)rows;
    rows += pl.nchunks;
    if (!im.fresh) im.edited = true;
    im.solved = false;
  }
  return launch();
}

// The done gate of a device-driven relabel, run by the last block of the
// relabel's compaction of the active flags (c: null for a plain compaction).
// Sets the solve's done flag once a complete relabel (its queue, list
// `qidx` at `q`, drained) found no active tile; a done solve zeroes the
// active and relabel counts, so every later push and relabel launch of a
// recorded solve exits at once.  A relabel whose queue did not drain marks
// the queued tiles dirty: their distances are the only ones that may
// violate h(p) <= h(q) + 1 across an arc (a neighbour's fell and they were
// not recomputed yet), so a local relabel that recomputes them -- and the
// rounds that follow it -- starts again from valid stored distances.  Also
// zeroes the counts the next relabel appends to.
struct Gate {
  int32_t* c;
  int qidx;
  const int32_t* q;
  int32_t* dirty;
};

__device__ __forceinline__ int load_count(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by every thread of every block after its last append.  The last
// block to finish runs the gate: each block's count append returned before
// its finish was counted, so the active count it reads is complete.
__device__ __forceinline__ void gate_last_block(const Gate& G) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&G.c[CNT_BLOCKS], 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  int32_t* c = G.c;
  const int nq = load_count(c + CNT_Q + G.qidx);
  for (int i = threadIdx.x; i < nq; i += blockDim.x) G.dirty[G.q[i]] = 1;
  __syncthreads();  // every thread has read the count before it is zeroed
  if (threadIdx.x == 0) {
    if (nq == 0 && load_count(c + CNT_ACT) == 0) c[CNT_DONE] = 1;
    if (load_count(c + CNT_DONE)) {
      c[CNT_ACT] = 0;
      c[CNT_GSEL] = 0;
    }
    // the next relabel's queue and local-set counts; the next gated
    // compaction's block count
    c[CNT_Q] = c[CNT_Q + 1] = c[CNT_Q + 2] = c[CNT_DIRTY] = 0;
    c[CNT_BLOCKS] = 0;
  }
}

// The selected tiles (positions of `sel`, or tiles 0..n-1) whose flag is
// set, appended to `out` in any order with one atomic per block of
// kFlagNT * kFlagPer positions; *count must be zero before.  clear: each
// listed flag is zeroed.  imgskip (may be null): tiles of images flagged
// there are left out (and not cleared).  done (may be null): unless *done
// is set, nothing is listed or cleared (a device-driven solve's flood runs
// only once the solve is done).
constexpr int kFlagNT = 256;
constexpr int kFlagPer = 4;

__global__ void __launch_bounds__(kFlagNT)
k_flag_list(const int32_t* __restrict__ sel, int n, int32_t* flags, int clear, int32_t* out,
            int32_t* count, const int32_t* __restrict__ tile_img,
            const int32_t* __restrict__ imgskip, const int32_t* __restrict__ done, Gate G) {
  __shared__ int s_off[kFlagNT / kWave];
  __shared__ int s_base;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  if (done && *done == 0) n = 0;
  const unsigned long long below = (1ull << lane) - 1;
  for (int base = blockIdx.x * kFlagNT * kFlagPer; base < n;
       base += gridDim.x * kFlagNT * kFlagPer) {
    int t[kFlagPer];
    bool keep[kFlagPer];
#pragma unroll
    for (int j = 0; j < kFlagPer; ++j) {
      const int i = base + j * kFlagNT + threadIdx.x;
      t[j] = 0;
      keep[j] = false;
      if (i < n) t[j] = sel ? sel[i] : i;
    }
#pragma unroll
    for (int j = 0; j < kFlagPer; ++j)
      if (base + j * kFlagNT + threadIdx.x < n)
        keep[j] = flags[t[j]] != 0 && !(imgskip && imgskip[tile_img[t[j]]]);
    unsigned long long m[kFlagPer];
    int tot = 0;
#pragma unroll
    for (int j = 0; j < kFlagPer; ++j) {
      m[j] = __ballot(keep[j]);
      tot += __popcll(m[j]);
    }
    if (lane == 0) s_off[wave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
      int sum = 0;
      for (int w = 0; w < kFlagNT / kWave; ++w) {
        const int c = s_off[w];
        s_off[w] = sum;
        sum += c;
      }
      s_base = sum ? atomicAdd(count, sum) : 0;
    }
    __syncthreads();
    int at = s_base + s_off[wave];
#pragma unroll
    for (int j = 0; j < kFlagPer; ++j) {
      if (keep[j]) {
        out[at + __popcll(m[j] & below)] = t[j];
        if (clear) flags[t[j]] = 0;
      }
      at += __popcll(m[j]);
    }
    __syncthreads();  // s_off and s_base are rewritten by the next pass
  }
  if (G.c) gate_last_block(G);
}

// The solve's tiles (sel) whose flag is set (leaving out the images flagged
// in imgskip, if given), as a list + count for the solver's own launches,
// which do not depend on the list's order.  zeroed: an earlier launch of
// the solve already zeroed the count (no fill launch).
enum { COUNT_FILL = 0, COUNT_ZEROED = 1 };
static int compact_sel(Pool* p, int32_t* flags, int32_t* out, int cnt, bool clear, int nsel,
                       const int32_t* imgskip, int zeroed, const Gate& gate = Gate{},
                       const int32_t* done = nullptr) {
  Prof* P = p->prof.on ? &p->prof : nullptr;
  if (P) P->begin(p->stream, K_COMPACT);
  if (zeroed != COUNT_ZEROED)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, p->stream, p->counts + cnt, 1LL, 0);
  constexpr int per_block = kFlagNT * kFlagPer;
  hipLaunchKernelGGL(k_flag_list,
                     dim3(std::max(1, std::min((nsel + per_block - 1) / per_block, 2048))),
                     dim3(kFlagNT), 0, p->stream, p->lsel, nsel, flags, clear ? 1 : 0, out,
                     p->counts + cnt, p->tile_img, imgskip, done, gate);
  if (P) P->end(p->stream);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

// The queue of a BFS or flood round reading rotating list `cur`; the epoch
// slots alternate with every round.
static Queue round_queue(Pool* p, int cur) {
  const int nxt = (cur + 1) % 3, clr = (cur + 2) % 3;
  const size_t cap = (size_t)p->cap_tiles;
  const Queue q{p->q + cur * cap,          p->counts + CNT_Q + cur,
                p->q + nxt * cap,          p->counts + CNT_Q + nxt,
                p->counts + CNT_Q + clr,   p->counts + CNT_EP + p->ep_par,
                p->counts + CNT_EP + 1 - p->ep_par};
  p->ep_par ^= 1;
  return q;
}

// The counts to the host (and with nres > 0 the tallies and the first nres
// images' flows, the rest of the status block, in the same copy); waits.
static int read_counts(Pool* p, int nres = 0) {
  const size_t bytes = nres > 0 ? kStatusFlows + (size_t)nres * 8 : CNT_N * 4;
  GCMI_HIP(hipMemcpyAsync(p->h_status, p->d_status, bytes, hipMemcpyDeviceToHost, p->stream));
  GCMI_TRY(stream_sync(p));
  p->stats[STAT_SYNCS]++;
  // stamps hold epochs: restart both before the counter nears overflow
  if (std::max(p->h_counts[CNT_EP], p->h_counts[CNT_EP + 1]) >= (1 << 30)) p->ep_reset = true;
  return GCMI_OK;
}

namespace {

// One solve's fixed facts, shared by its phases.
struct Solve {
  Pool* p;
  const PoolOps* ops;
  int n, nsel, nfresh, sweeps, bfs;
  int grid_relabel;
  bool dd;  // device-driven: fixed launch shapes, no host reads (recordable)
  Prof* P;
  void B(int k) const {
    if (P) P->begin(p->stream, k);
  }
  void E() const {
    if (P) P->end(p->stream);
  }
  int grid_for(int m) const {
    return dd ? p->grid_cap : (m < p->grid_cap ? (m > 0 ? m : 1) : p->grid_cap);
  }
  int grid_clear() const { return std::min((nsel + 255) / 256, 4096); }
};

// Start of every solve: the image lists, sums, tallies and counts.
int phase_prepare(const Solve& S) {
  Pool* p = S.p;
  // (the same image list as the last solve's -- a run of warm strokes on
  // one image -- finds its table on the device already; a recording copies
  // it on every replay)
  const size_t pb = (size_t)S.n * sizeof(Prep);
  if (p->capturing || p->last_prep.size() != pb ||
      memcmp(p->last_prep.data(), p->h_prep, pb) != 0) {
    GCMI_HIP(hipMemcpyAsync(p->d_prep, p->h_prep, pb, hipMemcpyHostToDevice, p->stream));
    p->stage_pending[STAGE_PREP] = true;
    p->last_prep.assign(reinterpret_cast<const uint8_t*>(p->h_prep),
                        reinterpret_cast<const uint8_t*>(p->h_prep) + pb);
  }
  int maxnt = 1;
  for (int i = 0; i < S.n; ++i) maxnt = std::max(maxnt, p->imgs[p->h_prep[i].img].ntiles);
  const int gy = std::min(1024, (maxnt + 63) / 64);
  hipLaunchKernelGGL(k_prepare, dim3(S.n, gy), dim3(256), 0, p->stream, p->geo, p->dev, p->d_prep,
                     S.nsel, S.nfresh, p->sel, p->fresh, p->counts, p->d_imgfresh, 4 * p->T);
  GCMI_HIP(hipGetLastError());
  p->ep_par = 0;
  p->flood_ran = false;
  return GCMI_OK;
}

// Whether round 1 of the fresh images may leave its heights unstored: with
// one push round per relabel and no warm image (whose pre-push lists every
// active tile, fresh ones too), the full relabel of those tiles comes next,
// in the host loop and in a device-driven run alike.
bool first_relabel_next(const Solve& S) { return S.bfs <= 1 && S.nsel == S.nfresh; }

// The first push round: every tile of the fresh images, then the tiles the
// edits of the warm images made active.
int phase_first(const Solve& S, int32_t* nact) {
  Pool* p = S.p;
  if (S.nfresh > 0) {
    // (k_prepare reset the fresh images' tiles)
    S.B(K_PUSH_FIRST);
    S.ops->push(p, S.nfresh, S.sweeps, first_relabel_next(S) ? PUSH_FIRST_NO_H : PUSH_FIRST);
    S.E();
    GCMI_HIP(hipGetLastError());
  }
  if (S.nsel > S.nfresh) {
    // (k_prepare zeroed the count)
    GCMI_TRY(compact_sel(p, p->dev.act, p->alist, CNT_ACT, true, S.nsel, nullptr, COUNT_ZEROED));
    if (!S.dd) {
      GCMI_TRY(read_counts(p));
      *nact = p->h_counts[CNT_ACT];
      if (p->trace) fprintf(stderr, "gcmi: warm pre-push over %d tiles\n", *nact);
    }
    if (S.dd || *nact > 0) {
      S.B(K_PUSH);
      S.ops->push(p, S.grid_for(*nact), S.sweeps, PUSH_LATER);
      S.E();
      GCMI_HIP(hipGetLastError());
    }
  }
  return GCMI_OK;
}

// A relabel and its BFS rounds; returns the rotating list the next round
// would read in *cur.  Host-driven: rounds run until the queue drains, the
// active list is compacted and its count read (*nact).  Device-driven: a
// fixed number of rounds, then the compaction and the done gate.
int phase_relabel(const Solve& S, bool local, int rounds_dd, int* cur_out, int32_t* nact) {
  Pool* p = S.p;
  const int32_t* glist = p->lsel;
  int gcnt = S.dd ? CNT_GSEL : CNT_SEL;
  // the queue counts and the local set's count start at zero (a
  // device-driven run: zeroed by k_prepare or the last relabel's gate)
  static_assert(CNT_DIRTY == CNT_Q + 3, "one fill zeroes the queue and local-set counts");
  if (!S.dd)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, p->stream, p->counts + CNT_Q, 4LL, 0);
  if (local) {
    // (marked with the epoch slot the first pass will read)
    hipLaunchKernelGGL(k_dilate, dim3(S.grid_clear()), dim3(256), 0, p->stream, p->geo, p->dev,
                       p->lsel, p->counts + CNT_SEL, p->dlist, p->counts + CNT_DIRTY,
                       p->counts + CNT_EP + p->ep_par);
    glist = p->dlist;
    gcnt = CNT_DIRTY;
  }
  S.B(K_FOLD);
  S.ops->fold(p, S.grid_relabel, glist, gcnt);
  S.E();
  // the neighbours the first pass's distances may improve start queue 0
  const Queue first{nullptr,
                    nullptr,
                    p->q,
                    p->counts + CNT_Q,
                    nullptr,
                    p->counts + CNT_EP + p->ep_par,
                    p->counts + CNT_EP + 1 - p->ep_par};
  p->ep_par ^= 1;
  S.B(K_BFS_INIT);
  S.ops->bfs_init(p, S.nsel, glist, gcnt, !local, first);
  S.E();
  GCMI_HIP(hipGetLastError());
  int cur = 0;
  const int32_t* hc = p->h_counts;
  if (S.dd) {
    for (int b = 0; b < rounds_dd; ++b) {
      S.ops->bfs_round(p, p->grid_cap, round_queue(p, cur));
      cur = (cur + 1) % 3;
    }
    // the one compaction of a device-driven relabel clears the flags it
    // lists (nothing reads them before the push round that follows), and
    // its last block runs the done gate
    const Gate gate{p->counts, cur, p->q + (size_t)cur * p->cap_tiles, p->dev.dirty};
    GCMI_TRY(
        compact_sel(p, p->dev.act, p->alist, CNT_ACT, true, S.nsel, nullptr, COUNT_ZEROED, gate));
    *cur_out = cur;
    return GCMI_OK;
  }
  if (p->trace) {
    GCMI_TRY(read_counts(p));
    fprintf(stderr, "gcmi: relabel %s over %d tiles, first queue %d\n", local ? "local" : "full",
            local ? hc[CNT_DIRTY] : S.nsel, hc[CNT_Q]);
  }
  // later BFS rounds, several per host check (a round with an empty input
  // queue exits at once); the active list rides along with the check
  int zeroed = COUNT_ZEROED;
  for (int batch = p->trace ? 1 : kRoundBatch;; batch = p->trace ? 1 : kRoundBatchMax) {
    for (int b = 0; b < batch; ++b) {
      S.B(K_BFS_ROUND);
      S.ops->bfs_round(p, p->grid_cap, round_queue(p, cur));
      S.E();
      GCMI_HIP(hipGetLastError());
      cur = (cur + 1) % 3;
    }
    GCMI_TRY(compact_sel(p, p->dev.act, p->alist, CNT_ACT, false, S.nsel, nullptr, zeroed));
    zeroed = COUNT_FILL;
    if (p->trace)
      GCMI_HIP(hipMemcpyAsync(p->h_tally, p->dev.tally, TALLY_N * 8, hipMemcpyDeviceToHost,
                              p->stream));
    GCMI_TRY(read_counts(p));
    if (p->trace)
      fprintf(stderr, "gcmi:   bfs round -> queue %d (tiles so far %llu)\n", hc[CNT_Q + cur],
              p->h_tally[TALLY_BFS]);
    if (hc[CNT_Q + cur] == 0) break;
  }
  *nact = hc[CNT_ACT];
  *cur_out = cur;
  return GCMI_OK;
}

// A push round over the active list built by the last relabel.  A
// host-driven relabel compacts after every batch of rounds without clearing
// the flags, so they are cleared here; a device-driven one cleared them.
int phase_push(const Solve& S, int32_t nact) {
  Pool* p = S.p;
  if (!S.dd)
    hipLaunchKernelGGL(k_clear_listed, dim3((nact + 255) / 256), dim3(256), 0, p->stream,
                       p->dev.act, p->alist, p->counts + CNT_ACT);
  S.B(K_PUSH);
  S.ops->push(p, S.grid_for(nact), S.sweeps, PUSH_LATER);
  S.E();
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

// Further push rounds before a relabel (bfs > 1).  Host-driven: stops early
// (returning false) when no tile is active.
int phase_extra_pushes(const Solve& S, bool* any) {
  Pool* p = S.p;
  *any = true;
  for (int r = 1; r < S.bfs; ++r) {
    GCMI_TRY(compact_sel(p, p->dev.act, p->alist, CNT_ACT, true, S.nsel, nullptr, COUNT_FILL));
    int32_t nact = 0;
    if (!S.dd) {
      GCMI_TRY(read_counts(p));
      nact = p->h_counts[CNT_ACT];
      if (nact == 0) {
        *any = false;
        return GCMI_OK;
      }
    }
    S.B(K_PUSH);
    S.ops->push(p, S.grid_for(nact), S.sweeps, PUSH_LATER);
    S.E();
    GCMI_HIP(hipGetLastError());
  }
  return GCMI_OK;
}

// Minimal source set: the first flood pass from the tiles holding excess,
// then rounds (device-driven: a fixed number), labels and flows.  Returns
// the list the next round would read in *cur_out.
int phase_flood_start(const Solve& S, int rounds_dd, int* cur_out) {
  Pool* p = S.p;
  // the images' final sum of max(e, 0) (k_prepare zeroed it; a flood run
  // before in the same solve added to it) and the counts the flood's
  // compactions, first pass, queues and labels list append to
  if (p->flood_ran)
    hipLaunchKernelGGL(k_zero_posf, dim3(S.n), dim3(64), 0, p->stream, p->dev.sums, p->d_prep,
                       p->counts);
  p->flood_ran = true;
  const unsigned zmask = (1u << CNT_FIRST) | (1u << CNT_POS) | (1u << CNT_ACT) | (1u << CNT_Q) |
                         (1u << (CNT_Q + 1)) | (1u << (CNT_Q + 2));
  const size_t cap = (size_t)p->cap_tiles;
  const bool inc = S.nfresh < S.nsel;
  p->flood_inc = inc;
  // A device-driven solve's flood is gated on its done flag: its lists stay
  // empty, and the changed flags and reach bytes untouched, unless the
  // solve finished.  (A flood over an unfinished solve's state would leave
  // reach computed for that state -- and, with its rounds cut short,
  // incomplete -- in tiles whose changed flags it consumed, which the
  // incremental flood after the host finishes the solve would then trust.)
  const int32_t* done = S.dd ? p->counts + CNT_DONE : nullptr;
  // (an incremental flood flags the tiles it processes in bflag, all zero
  // since the last relabel's compaction, for their labels)
  int32_t* lmark = inc ? p->dev.bflag : nullptr;
  // a small incremental flood starts in one block (sums, changed tiles and
  // their closure); otherwise the sums, a compaction and the closure
  // (a flood from scratch starts in one pass: sums, first list, reach)
  const bool small = inc && S.nsel <= kSmallSel;
  if (small)
    hipLaunchKernelGGL(k_flood_start_small, dim3(1), dim3(1024), 0, p->stream, p->geo, p->dev,
                       p->lsel, S.nsel, p->counts, zmask, p->d_imgfresh, done, p->q, p->q + cap,
                       p->q + 2 * cap);
  else if (inc)
    S.ops->posf(p, std::min((S.nsel + 255) / 256, 1024), zmask);
  else
    S.ops->flood_start_full(p, S.nsel, zmask & ~(1u << CNT_POS), done);
  if (inc) {
    // incremental: recompute the changed tiles and the tiles whose reach
    // may have come through them; every other tile keeps its reach bytes
    // (they can only grow, and the rounds grow them)
    // (a fresh image's tiles are all in the set already, and so are their
    // neighbours: the closure starts from the warm images' changed tiles)
    // (k_zero_posf zeroed the counts of these compactions)
    if (!small) {
      GCMI_TRY(compact_sel(p, p->dev.fchg, p->q, CNT_Q, false, S.nsel, p->d_imgfresh,
                           COUNT_ZEROED, Gate{}, done));
      hipLaunchKernelGGL(k_imp_closure, dim3(1), dim3(1024), 0, p->stream, p->geo, p->dev, p->q,
                         p->counts + CNT_Q, p->q + cap, p->q + 2 * cap);
    }
    // the set's tiles: flags cleared, reach zeroed, marked for labels, and
    // the first pass's list
    S.ops->flood_set(p, S.nsel, lmark, done);
  }
  // (from scratch: the tiles holding excess seed the first pass; haspos
  // persists, a clean tile keeping the flag its last relabel wrote)
  GCMI_HIP(hipGetLastError());
  // (the queue counts are zero: k_zero_posf, and k_flood_set after the
  // incremental path's use of the queue as a compaction target)
  S.B(K_FLOOD);
  const Queue first{inc ? p->alist : p->plist,
                    p->counts + (inc ? CNT_FIRST : CNT_POS),
                    p->q,
                    p->counts + CNT_Q,
                    nullptr,
                    p->counts + CNT_EP + p->ep_par,
                    p->counts + CNT_EP + 1 - p->ep_par};
  p->ep_par ^= 1;
  S.ops->flood(p, p->grid_cap, first, inc ? FLOOD_FIRST_INC : FLOOD_FIRST, lmark);
  S.E();
  GCMI_HIP(hipGetLastError());
  int cur = 0;
  for (int b = 0; b < rounds_dd; ++b) {
    S.B(K_FLOOD);
    S.ops->flood(p, p->grid_cap, round_queue(p, cur), FLOOD_ROUND, lmark);
    S.E();
    GCMI_HIP(hipGetLastError());
    cur = (cur + 1) % 3;
  }
  *cur_out = cur;
  return GCMI_OK;
}

// Labels and flows in one launch (no synchronisation; read_counts with the
// solve's image count then brings flows and tallies to the host in the
// counts' copy).
// zeroed: the active count is zero (right after the flood's start).
int phase_results(const Solve& S, int zeroed) {
  Pool* p = S.p;
  if (p->flood_inc) {
    // the labels of the tiles the flood processed (the others' reach, and
    // so their labels, did not change); the active list is free by now
    GCMI_TRY(compact_sel(p, p->dev.bflag, p->alist, CNT_ACT, true, S.nsel, nullptr, zeroed, Gate{},
                         S.dd ? p->counts + CNT_DONE : nullptr));
    S.B(K_LABELS);
    S.ops->labels(p, S.nsel, p->alist, CNT_ACT, S.n);
  } else {
    S.B(K_LABELS);
    S.ops->labels(p, S.nsel, p->lsel, CNT_SEL, S.n);
  }
  S.E();
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

// Host-driven flood rounds from list `cur`, each batch closed by labels,
// flows and a count check.
int flood_rounds_host(const Solve& S, int cur) {
  Pool* p = S.p;
  const int32_t* hc = p->h_counts;
  for (int batch = p->trace ? 1 : kRoundBatch;; batch = p->trace ? 1 : kRoundBatchMax) {
    for (int b = 0; b < batch; ++b) {
      S.B(K_FLOOD);
      S.ops->flood(p, p->grid_cap, round_queue(p, cur), FLOOD_ROUND,
                   p->flood_inc ? p->dev.bflag : nullptr);
      S.E();
      GCMI_HIP(hipGetLastError());
      cur = (cur + 1) % 3;
    }
    // labels and flows ride along with the check
    GCMI_TRY(phase_results(S, COUNT_FILL));
    GCMI_TRY(read_counts(p, S.n));
    if (p->trace)
      fprintf(stderr, "gcmi:   flood round -> queue %d (tiles so far %llu)\n", hc[CNT_Q + cur],
              p->h_tally[TALLY_FLOOD]);
    if (hc[CNT_Q + cur] == 0) break;
  }
  return GCMI_OK;
}

// Push/relabel iterations of a device-driven solve (the host loop finishes
// a solve that needs more): runs from scratch mostly finish after the first
// relabel or the second, and so do warm reflows of one or two images; in a
// larger batch of reflows some image mostly needs a third, which costs less
// than a host round trip.
int dd_iters(const Solve& S) { return S.nfresh == S.nsel || S.n <= 2 ? 2 : 3; }

// The device-driven solve (the sequence a recording holds): fixed launch
// shapes, `rounds` BFS and flood rounds per relabel, no host reads.
// *flood_cur: the flood list the next round would read.
int emit_device_driven(const Solve& S, int rounds, int* flood_cur) {
  int32_t nact = 0;
  GCMI_TRY(phase_prepare(S));
  GCMI_TRY(phase_first(S, &nact));
  for (int it = 0; it < dd_iters(S); ++it) {
    if (it > 0) {
      bool any;
      GCMI_TRY(phase_extra_pushes(S, &any));
    }
    // local relabels as in the host loop (their dirty lists are empty, so
    // they are gated too, once the solve is done)
    const bool local = S.p->local_relabel && (S.nfresh < S.nsel || it > 0) && it < kLocalRelabels;
    int cur;
    GCMI_TRY(phase_relabel(S, local, rounds, &cur, &nact));
    GCMI_TRY(phase_push(S, 0));
  }
  GCMI_TRY(phase_flood_start(S, rounds, flood_cur));
  return phase_results(S, COUNT_ZEROED);
}

}  // namespace

// Records the device-driven solve of the current image list (or reuses the
// recording of the same list and parameters) and replays it.  The caller
// reads the counts and finishes the solve if the recording did not.
static int run_recorded(Pool* p, const Solve& S, int* flood_cur) {
  std::vector<int> key;
  key.reserve(4 * S.n + 8);
  key.push_back(S.n);
  key.push_back(S.nsel);
  key.push_back(S.nfresh);
  for (int i = 0; i < S.n; ++i) {
    const PoolImage& im = p->imgs[p->h_prep[i].img];
    key.push_back(p->h_prep[i].img);
    key.push_back(p->h_prep[i].fresh_off);
    key.push_back(im.tile0);
    key.push_back(im.ntiles);
  }
  key.push_back(S.sweeps);
  key.push_back(S.bfs);
  key.push_back(p->dd_rounds);
  key.push_back(p->lsel == nullptr);
  key.push_back(p->lfresh == nullptr);
  key.push_back(p->local_relabel);
  if (!p->graph_exec || key != p->graph_key) {
    if (p->graph_exec) {
      (void)hipGraphExecDestroy(p->graph_exec);
      p->graph_exec = nullptr;
    }
    GCMI_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
    int cur = 0;
    p->capturing = true;
    const int rc = emit_device_driven(S, p->dd_rounds, &cur);
    p->capturing = false;
    p->last_prep.clear();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(p->stream, &graph);
    if (rc != GCMI_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) {
      set_error("solve recording failed: %s", hipGetErrorString(e));
      return GCMI_EHIP;
    }
    const hipError_t ei = hipGraphInstantiate(&p->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      p->graph_exec = nullptr;
      set_error("solve recording could not be instantiated: %s", hipGetErrorString(ei));
      return GCMI_EHIP;
    }
    p->graph_key = key;
    p->graph_flood_cur = cur;
    p->graph_ep_par = p->ep_par;
  }
  GCMI_HIP(hipGraphLaunch(p->graph_exec, p->stream));
  // the epoch slot a round after the recording reads
  p->ep_par = p->graph_ep_par;
  *flood_cur = p->graph_flood_cur;
  return GCMI_OK;
}

// A completed solve of ids[0..n): the images' flows and state flags, and
// the run's counters (from the device tallies read with the flows).
static void finish_images(Pool* p, const int* ids, int n, int64_t* flows_out) {
  const int T = p->T;
  for (int i = 0; i < n; ++i) {
    PoolImage& im = p->imgs[ids[i]];
    im.flow = p->h_flows[i];
    im.fresh = false;
    im.edited = false;
    im.solved = true;
    if (flows_out) flows_out[i] = im.flow;
  }
  const unsigned long long* ht = p->h_tally;
  p->stats[STAT_ROUNDS] = (int64_t)ht[TALLY_PUSH_ROUNDS];
  p->stats[STAT_RESIDENCIES] = (int64_t)ht[TALLY_PUSH_TILES];
  p->stats[STAT_GLOBAL_RELABELS] = (int64_t)ht[TALLY_RELABELS];
  p->stats[STAT_BFS_PASSES] = (int64_t)(ht[TALLY_RELABEL_TILES] + ht[TALLY_BFS]);
  p->stats[STAT_FLOOD_PASSES] = (int64_t)ht[TALLY_FLOOD];
  p->stats[STAT_RELABEL_TILES] = (int64_t)ht[TALLY_RELABEL_TILES];
  p->stats[STAT_BFS_ROUND_PASSES] = (int64_t)ht[TALLY_BFS];
  // algorithmic bytes: first round 16 per image pixel read + 25 per padded
  // pixel written (state + mask; 21 without its heights); later residencies 49 (24 read, 25 written);
  // relabel first pass 5 (mask read, h written); later BFS passes 9 (mask +
  // h read, h written); flood passes 2 (mask read, reach written); labels 2.
  int64_t img_px = 0, fresh_px = 0, nfresh = 0;
  for (int i = 0; i < n; ++i) {
    const PoolImage& im = p->imgs[ids[i]];
    img_px += (int64_t)im.W * im.H;
    if (p->h_prep[i].fresh_off >= 0) {
      fresh_px += (int64_t)im.W * im.H;
      nfresh += im.ntiles;
    }
  }
  const int64_t tp = (int64_t)T * T;
  p->stats[STAT_ALG_BYTES] = fresh_px * 16 + nfresh * tp * (p->first_no_h ? 21 : 25) +
                             (p->stats[STAT_RESIDENCIES] - nfresh) * tp * 49 +
                             (int64_t)ht[TALLY_RELABEL_TILES] * tp * 5 +
                             (int64_t)ht[TALLY_BFS] * tp * 9 +
                             (int64_t)ht[TALLY_FLOOD] * tp * 2 + img_px * 2;
}

int pool_solve(Pool* p, const int* ids, int n, int sweeps, int bfs, int64_t max_rounds,
               int64_t* flows_out) {
  p->step_img = -1;  // a stepped solve in progress is dropped
  if (n <= 0 || n > p->cap_imgs || !ids) {
    set_error("bad image list");
    return GCMI_EINVAL;
  }
  if (sweeps <= 0) sweeps = kDefaultSweeps;
  if (bfs <= 0) bfs = kDefaultBfs;
  int nsel = 0, nfresh = 0;
  for (int i = 0; i < n; ++i) {
    GCMI_TRY(check_img(p, ids[i]));
    PoolImage& im = p->imgs[ids[i]];
    if (!im.uploaded) {
      set_error("image %d solved before upload", ids[i]);
      return GCMI_EINVAL;
    }
    for (int j = 0; j < i; ++j)
      if (ids[j] == ids[i]) {
        set_error("image %d listed twice", ids[i]);
        return GCMI_EINVAL;
      }
    // the pinned prep buffer may still be read by a previous solve's copy
    if (i == 0) GCMI_TRY(stage_wait(p, STAGE_PREP));
    // warm only if edited since the last solve; otherwise start over from
    // the stored capacities (which include every edit so far)
    const bool fresh = im.fresh || !im.edited;
    p->h_prep[i] = Prep{ids[i], nsel, fresh ? nfresh : -1, 0};
    nsel += im.ntiles;
    if (fresh) nfresh += im.ntiles;
  }
  // the whole pool in tile order: kernels and compactions skip the lists
  bool ident = nsel == p->cap_tiles;
  for (int i = 0; i < n && ident; ++i) ident = p->imgs[ids[i]].tile0 == p->h_prep[i].sel_off;
  p->lsel = ident ? nullptr : p->sel;
  p->lfresh = (ident && nfresh == nsel) ? nullptr : p->fresh;
  Prof* P = p->prof.on ? &p->prof : nullptr;
  if (P) P->reset();
  for (int k = 0; k < STAT_N; ++k) p->stats[k] = 0;
  p->stats[STAT_TILES] = nsel;
  if (p->ep_reset) {
    // stamps and local-relabel marks hold epochs
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min((p->cap_tiles + 255) / 256, 4096)), dim3(256), 0,
                       p->stream, p->dev.stamp, (long long)p->cap_tiles, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min((p->cap_tiles + 255) / 256, 4096)), dim3(256), 0,
                       p->stream, p->dev.rflag, (long long)p->cap_tiles, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, p->stream, p->counts + CNT_EP, 2LL, 0);
    GCMI_HIP(hipGetLastError());
    p->ep_reset = false;
  }
  // the relabel's first pass runs one block per listed tile (their number
  // is known on the device only; surplus blocks exit at once)
  Solve S{p, p->ops, n, nsel, nfresh, sweeps, bfs, std::min(nsel, kRelabelGridCap), false, P};
  // (set here too for a replayed recording, whose host code does not run)
  p->first_no_h = nfresh > 0 && first_relabel_next(S);
  const int32_t* hc = p->h_counts;
  int32_t nact = 0;
  int64_t rounds = 0;
  bool solved = false, flooded = false;
  int cur = 0;
  const bool dd_run = p->dd_rounds > 0 && !P && !p->trace;
  if (dd_run) {
    Solve R = S;
    R.dd = true;
    // gated launches of a done solve exit at once: one block per CU slot
    R.grid_relabel = std::min(nsel, p->grid_cap);
    if (p->graph)
      GCMI_TRY(run_recorded(p, R, &cur));
    else
      GCMI_TRY(emit_device_driven(R, p->dd_rounds, &cur));
    GCMI_TRY(read_counts(p, n));
    p->flood_ran = true;  // (a replayed recording ran its flood too)
    solved = hc[CNT_DONE] != 0;
    flooded = solved && hc[CNT_Q + cur] == 0;
    rounds = dd_iters(R) + 1;
    if (!solved) {
      // go on from the state after the recording's last push round
      nact = 1;
    }
  } else {
    GCMI_TRY(phase_prepare(S));
    GCMI_TRY(phase_first(S, &nact));
    rounds = (nfresh > 0) + (nsel > nfresh && nact > 0);
  }
  int nlocal = dd_run ? dd_iters(S) : 0;  // relabels that may have been local so far
  int relabels = dd_run ? dd_iters(S) : 0;
  while (!solved) {
    bool any = true;
    if (rounds > 0) GCMI_TRY(phase_extra_pushes(S, &any));
    // a global relabel: over every selected tile, or (local) only over the
    // tiles that changed and a ring around them -- in solves with edited
    // images, or after a first relabel (before it, a fresh image's tiles
    // hold no distances), and for kLocalRelabels relabels at most, which
    // bounds the cost of the stale distances a local relabel leaves in
    // clean tiles
    const bool local =
        p->local_relabel && (nfresh < nsel || relabels > 0) && nlocal < kLocalRelabels;
    if (local) ++nlocal;
    ++relabels;
    GCMI_TRY(phase_relabel(S, local, 0, &cur, &nact));
    if (p->trace) fprintf(stderr, "gcmi: push round over %d active tiles\n", nact);
    if (nact == 0) break;
    if (max_rounds > 0 && rounds >= max_rounds) {
      set_error("cut did not terminate within %lld push rounds", (long long)max_rounds);
      return GCMI_ENOTCONVERGED;
    }
    GCMI_TRY(phase_push(S, nact));
    ++rounds;
  }
  if (!flooded) {
    if (solved) {
      // the recording's flood did not drain: its rounds go on
      GCMI_TRY(flood_rounds_host(S, cur));
    } else {
      GCMI_TRY(phase_flood_start(S, 0, &cur));
      GCMI_TRY(flood_rounds_host(S, cur));
    }
  }
  if (P) P->collect();
  finish_images(p, ids, n, flows_out);
  return GCMI_OK;
}

// Solving one image step by step, as the reference's GraphCut hooks do
// (GraphCut.pyx:187-346): the solve's state lives in the pool between the
// calls.  begin() sets it up: the image list and counts, and for an image
// cut from scratch its tile state built from the capacities (round 1 of the
// push kernel with no sweeps: e, caps, h = 0 on sinks and 1 elsewhere).
static Solve step_solve(Pool* p) {
  const PoolImage& im = p->imgs[p->step_img];
  const int nfresh = p->step_fresh ? im.ntiles : 0;
  return Solve{p, p->ops, 1, im.ntiles, nfresh, p->step_sweeps, kDefaultBfs,
               std::min(im.ntiles, kRelabelGridCap), false, nullptr};
}

static int step_begin(Pool* p, int img) {
  if (p->step_img == img) return GCMI_OK;
  GCMI_TRY(check_img(p, img));
  PoolImage& im = p->imgs[img];
  if (!im.uploaded) {
    set_error("image %d stepped before upload", img);
    return GCMI_EINVAL;
  }
  GCMI_TRY(stage_wait(p, STAGE_PREP));
  const bool fresh = im.fresh || !im.edited;
  p->h_prep[0] = Prep{img, 0, fresh ? 0 : -1, 0};
  const bool ident = im.ntiles == p->cap_tiles && im.tile0 == 0;
  p->lsel = ident ? nullptr : p->sel;
  p->lfresh = (ident && fresh) ? nullptr : p->fresh;
  for (int k = 0; k < STAT_N; ++k) p->stats[k] = 0;
  p->stats[STAT_TILES] = im.ntiles;
  if (p->ep_reset) {
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min((p->cap_tiles + 255) / 256, 4096)), dim3(256), 0,
                       p->stream, p->dev.stamp, (long long)p->cap_tiles, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min((p->cap_tiles + 255) / 256, 4096)), dim3(256), 0,
                       p->stream, p->dev.rflag, (long long)p->cap_tiles, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, p->stream, p->counts + CNT_EP, 2LL, 0);
    GCMI_HIP(hipGetLastError());
    p->ep_reset = false;
  }
  p->step_img = img;
  p->step_fresh = fresh;
  p->step_last = STEP_BEGIN;
  p->step_nact = -1;
  p->step_relabels = 0;
  const Solve S = step_solve(p);
  GCMI_TRY(phase_prepare(S));
  if (fresh) {
    // (a stepped solve may push again before it relabels)
    p->first_no_h = false;
    S.ops->push(p, S.nfresh, 0, PUSH_FIRST);
    GCMI_HIP(hipGetLastError());
  }
  return GCMI_OK;
}

int pool_step_push(Pool* p, int img, int sweeps) {
  if (p->step_img != img) p->step_sweeps = sweeps > 0 ? sweeps : kDefaultSweeps;
  GCMI_TRY(step_begin(p, img));
  if (sweeps > 0) p->step_sweeps = sweeps;
  const Solve S = step_solve(p);
  // the tiles holding active pixels now (the flags are cleared as listed;
  // the push sets them again where pixels stay active)
  GCMI_TRY(compact_sel(p, p->dev.act, p->alist, CNT_ACT, true, S.nsel, nullptr, COUNT_FILL));
  GCMI_TRY(read_counts(p));
  const int32_t nact = p->h_counts[CNT_ACT];
  if (nact > 0) {
    S.ops->push(p, S.grid_for(nact), S.sweeps, PUSH_LATER);
    GCMI_HIP(hipGetLastError());
  }
  p->step_last = STEP_PUSH;
  return GCMI_OK;
}

int pool_step_relabel(Pool* p, int img, bool full, int32_t* nact_out) {
  if (p->step_img != img) p->step_sweeps = kDefaultSweeps;
  GCMI_TRY(step_begin(p, img));
  const Solve S = step_solve(p);
  // a local relabel keeps the stored distances of clean tiles: valid once a
  // relabel has run in this solve, or from the solve before a warm edit
  const bool local = !full && (S.nfresh < S.nsel || p->step_relabels > 0);
  int cur = 0;
  int32_t nact = 0;
  GCMI_TRY(phase_relabel(S, local, 0, &cur, &nact));
  p->step_relabels++;
  p->step_nact = nact;
  p->step_last = STEP_RELABEL;
  if (nact_out) *nact_out = nact;
  return GCMI_OK;
}

int pool_step_completed(Pool* p, int img, int* done, int64_t* flow) {
  GCMI_TRY(check_img(p, img));
  *done = 0;
  if (p->step_img != img) {
    // not stepping: done if a cut of the current capacities is complete
    const PoolImage& im = p->imgs[img];
    *done = im.solved ? 1 : 0;
    if (flow) *flow = im.flow;
    return GCMI_OK;
  }
  // termination is decided right after a relabel that drained and left no
  // active tile (the same rule as pool_solve)
  if (p->step_last != STEP_RELABEL || p->step_nact != 0) return GCMI_OK;
  const Solve S = step_solve(p);
  int cur = 0;
  GCMI_TRY(phase_flood_start(S, 0, &cur));
  GCMI_TRY(flood_rounds_host(S, cur));
  const int ids[1] = {img};
  finish_images(p, ids, 1, flow);
  p->step_img = -1;
  *done = 1;
  return GCMI_OK;
}

int pool_labels(Pool* p, int img, uint8_t* out, bool device_dst) {
  GCMI_TRY(check_img(p, img));
  const PoolImage& im = p->imgs[img];
  if (!im.solved) {
    set_error("no completed cut for image %d", img);
    return GCMI_EINVAL;
  }
  GCMI_HIP(hipMemcpyAsync(out, p->dev.labels + im.pix0, (size_t)im.W * im.H,
                          device_dst ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          p->stream));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

const uint32_t* packed_mask(const Pool* p, int img) {
  return p->dev.packed + p->imgs[img].pix0 / 32;
}

int pool_labels_packed(Pool* p, int img, uint32_t* out) {
  GCMI_TRY(check_img(p, img));
  const PoolImage& im = p->imgs[img];
  if (!im.solved) {
    set_error("no completed cut for image %d", img);
    return GCMI_EINVAL;
  }
  const long long words = ((long long)im.W * im.H + 31) / 32;
  GCMI_HIP(hipMemcpyAsync(out, packed_mask(p, img), (size_t)words * 4, hipMemcpyDeviceToHost,
                          p->stream));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

int pool_set_option(Pool* p, int option, int value) {
  switch (option) {
    case GCMI_OPT_LOCAL_RELABEL:
      p->local_relabel = value != 0;
      return GCMI_OK;
    case GCMI_OPT_DEVICE_ROUNDS:
      if (value < 0 || value > kDeviceRoundsMax) {
        set_error("device-driven rounds must be in [0, %d] (got %d)", kDeviceRoundsMax, value);
        return GCMI_EINVAL;
      }
      p->dd_rounds = value;
      return GCMI_OK;
    case GCMI_OPT_GRAPH:
      p->graph = value != 0;
      if (p->graph && p->dd_rounds == 0) p->dd_rounds = kDeviceRoundsDefault;
      return GCMI_OK;
    default:
      set_error("unknown option %d", option);
      return GCMI_EINVAL;
  }
}

int pool_forget(Pool* p, int img) {
  p->step_img = -1;  // drops a stepped solve in progress
  GCMI_TRY(check_img(p, img));
  PoolImage& im = p->imgs[img];
  im.fresh = true;
  im.edited = false;
  im.solved = false;
  return GCMI_OK;
}

int pool_heights(Pool* p, int img, int32_t* out) { return pool_field(p, img, GCMI_FIELD_HEIGHT, out); }

int pool_field(Pool* p, int img, int which, int32_t* out) {
  GCMI_TRY(check_img(p, img));
  const int32_t* src = nullptr;
  switch (which) {
    case GCMI_FIELD_HEIGHT: src = p->dev.h; break;
    case GCMI_FIELD_EXCESS: src = p->dev.e; break;
    case GCMI_FIELD_CAP_UP: src = p->dev.cu; break;
    case GCMI_FIELD_CAP_DOWN: src = p->dev.cd; break;
    case GCMI_FIELD_CAP_LEFT: src = p->dev.cl; break;
    case GCMI_FIELD_CAP_RIGHT: src = p->dev.cr; break;
    default:
      set_error("unknown state field %d", which);
      return GCMI_EINVAL;
  }
  const PoolImage& im = p->imgs[img];
  const size_t np = (size_t)im.ntiles * p->T * p->T;
  int32_t* tmp = nullptr;
  if (hipMalloc(&tmp, np * 4) != hipSuccess) {
    (void)hipGetLastError();
    set_error("allocation failed");
    return GCMI_ENOMEM;
  }
  p->ops->untile(p, img, src, tmp);
  hipError_t e1 = hipMemcpyAsync(out, tmp, np * 4, hipMemcpyDeviceToHost, p->stream);
  hipError_t e2 = hipStreamSynchronize(p->stream);
  (void)hipFree(tmp);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    set_error("height readback failed");
    return GCMI_EHIP;
  }
  return GCMI_OK;
}

}  // namespace gcmi

// Device kernels of the tiled push-relabel grid cut (gfx950).
//
// A pool holds one or more images (a plain cut is a pool of one).  Every
// image is padded to whole T x T tiles and owns a contiguous range of pool
// tiles; every per-pixel array is tile-blocked -- pool tile t occupies
// [t*T*T, (t+1)*T*T), row-major inside the tile -- so one tile of one field
// is T*T contiguous int32 (whole 128-byte lines).  Row-major inputs and
// labels of image i live at pixel offset pix0(i) of their pool arrays.
//
// Per-pixel state (int32): e (excess; < 0 means remaining sink capacity),
// h (height / BFS distance), cu, cd, cl, cr (residual capacity of the arc to
// the up / down / left / right neighbour), plus a residual-mask byte.  Flow
// pushed across a tile edge is added atomically to a pending slot of the
// receiving tile (pend[(t*4 + side)*T + k]) instead of into its state, which
// another work-group may hold at the same time.
//
// Tile kernels take a list of pool tiles (a null list means positions are
// tile numbers: the whole pool in order) and map blocks to list positions
// XCD by XCD (xcd_pos), so that neighbouring tiles -- whose halos a block
// reads -- are processed on the same XCD and share its L2.
//
// Reference counterparts: push_up/down/left/right + add_border
// (graphcut.cl:114-536), relabel (:538-577), init_bfs / bfs_intratile /
// bfs_intertile (:582-1016), check_completion (:20-60).
#pragma once

#include "gcmi_common.h"

namespace gcmi {

enum { D_UP = 0, D_DOWN = 1, D_LEFT = 2, D_RIGHT = 3 };
enum { M_UP = 1, M_DOWN = 2, M_LEFT = 4, M_RIGHT = 8, M_SINK = 16, M_EXCESS = 32 };

// Per-image 64-bit sums, each spread over kSumSlots addresses so that many
// tiles do not serialise on one atomic.  Rows: sum of min(cs, ct); sum of
// max(e0, 0); sum of max(e, 0) after the last relabel.
constexpr int kSumSlots = 64;
enum { SUM_MIN = 0, SUM_POS0 = 1, SUM_POSF = 2, SUM_ROWS = 3 };
// Pool-wide tallies (added by block 0 of each launch): tiles processed by
// later BFS / flood passes; push launches with work and their tiles;
// relabel first passes with work and their tiles
enum {
  TALLY_BFS = 0,
  TALLY_FLOOD = 1,
  TALLY_PUSH_TILES = 2,
  TALLY_PUSH_ROUNDS = 3,
  TALLY_RELABELS = 4,
  TALLY_RELABEL_TILES = 5,
  TALLY_N = 6
};
// Number of XCDs blocks are dealt to round-robin.
constexpr int kXcds = 8;

struct ImgDesc {
  int W, H;        // image
  int tw, th;      // tiles across / down
  int tile0;       // first pool tile
  int ntiles;      // tw * th
  long long pix0;  // offset of the image's row-major pixels
};

struct Geo {
  int32_t maxh;             // MAX_HEIGHT: pool capacity in pixels (> any distance)
  int ntiles;               // pool capacity in tiles
  const ImgDesc* img;       // per image
  const int32_t* tile_img;  // per pool tile: its image
  // per pool tile, its geometry in two int4 (one 32-byte line):
  // {img, tx | ty << 16, tw | th << 16, W}, {H, pix0 low, pix0 high, 0}
  const int4* tgeo;
};

// Device pointers of a pool's state, passed by value.
struct CutDev {
  int32_t *e, *h, *cu, *cd, *cl, *cr;  // tile-blocked state
  uint8_t* mask;                       // residual bits + sink/excess bits
  uint8_t* reach;                      // flood result
  int32_t* pend;                       // pending edge inflow, 4*T per tile
  int32_t* pflag;                      // tile has (possibly) pending inflow
  int32_t* act;                        // tile flagged for the next push round
  int32_t* stamp;                      // per tile: epoch of its last queue entry
  int32_t* bflag;                      // tile flagged by the relabel's first pass
  int32_t* dirty;                      // residual arcs or excess changed since the
                                       // tile's distances were last computed
  int32_t* rflag;                      // epoch of the local relabel recomputing the tile
  int32_t* haspos;                     // tile holds excess
  uint32_t* packed;                    // per pool pixel: its label as one bit (row-major,
                                       // each image from a word of its own)
  int32_t* fchg;                       // arcs or excess changed since the tile's reach was
                                       // last computed
  int32_t* imp;                        // the tile's reach is not all reached from its own
                                       // excess through its own arcs (some came over an edge)
  long long* tpos;                     // per tile: sum of max(e, 0)
  unsigned long long* sums;            // per image [SUM_ROWS][kSumSlots]
  unsigned long long* tally;           // [TALLY_N]
  int32_t *cs, *ct;                    // row-major unaries (edits rewrite them)
  const int32_t *right, *down;         // row-major pairwise weights
  uint8_t* labels;                     // row-major labels
};

// Block-uniform geometry of one pool tile.
struct TG {
  int t, img, tx, ty, tw, th, W, H;
  long long pix0;
};

// (one load of the tile's geometry line, no dependent load of its image's
// descriptor)
__device__ __forceinline__ TG tile_geo(const Geo& g, int t) {
  const int4 a = g.tgeo[2 * (size_t)t], b = g.tgeo[2 * (size_t)t + 1];
  TG r;
  r.t = t;
  r.img = a.x;
  r.tx = a.y & 0xffff;
  r.ty = (int)((unsigned)a.y >> 16);
  r.tw = a.z & 0xffff;
  r.th = (int)((unsigned)a.z >> 16);
  r.W = a.w;
  r.H = b.x;
  r.pix0 = (long long)(((unsigned long long)(unsigned)b.z << 32) | (unsigned)b.y);
  return r;
}


// The same two reads through the scalar cache (the geometry lines and tile
// lists do not change while a kernel runs), for a block-uniform tile: they
// count against the scalar-load counter, so waiting for them never waits
// for the wave's vector loads and stores in flight.
using ConstI32 = const __attribute__((address_space(4))) int32_t;
__device__ __forceinline__ TG tile_geo_s(const Geo& g, int t) {
  t = __builtin_amdgcn_readfirstlane(t);
  ConstI32* c = (ConstI32*)g.tgeo + 8 * (size_t)t;
  TG r;
  r.t = t;
  r.img = c[0];
  r.tx = c[1] & 0xffff;
  r.ty = (int)((unsigned)c[1] >> 16);
  r.tw = c[2] & 0xffff;
  r.th = (int)((unsigned)c[2] >> 16);
  r.W = c[3];
  r.H = c[4];
  r.pix0 = (long long)(((unsigned long long)(unsigned)c[6] << 32) | (unsigned)c[5]);
  return r;
}
__device__ __forceinline__ int list_tile_s(const int32_t* list, int pos) {
  return list ? __builtin_amdgcn_readfirstlane(((ConstI32*)list)[pos]) : pos;
}

template <int T>
__device__ __forceinline__ size_t tbase(int t) {
  return (size_t)t * (T * T);
}

// The tile at position pos of a tile list (no list: position = tile).
// The per-tile loops load the next position's entry while they work on the
// current tile, so a tile's first dependent load is not the list read.
__device__ __forceinline__ int list_tile(const int32_t* __restrict__ list, int pos) {
  return list ? list[pos] : pos;
}

// Blocks b, b+8, b+16, ... are dealt to the same XCD; give each XCD a
// contiguous run of list positions so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_pos(int b, int grid) {
  const int x = b % kXcds, k = b / kXcds;
  const int per = grid / kXcds, rem = grid % kXcds;
  return x * per + min(x, rem) + k;
}

// The list positions block b of `grid` takes, XCD by XCD: XCD x (the blocks
// b % 8 == x) owns a contiguous eighth of the n positions and its blocks
// take them in turn, so the tiles an XCD holds at once are neighbours and
// the row above them was its own recent work (in its L2).  With one block
// per position this is xcd_pos.
struct XcdRun {
  int pos, step, end;
};
__device__ __forceinline__ XcdRun xcd_run(int b, int grid, int n) {
  const int x = b % kXcds, k = b / kXcds;
  const int per = n / kXcds, rem = n % kXcds;
  const int lo = x * per + min(x, rem);
  return XcdRun{lo + k, grid / kXcds + (x < grid % kXcds ? 1 : 0), lo + per + (x < rem ? 1 : 0)};
}

// Block-wide sum of a 64-bit value; result valid in thread 0.
template <int NT>
__device__ __forceinline__ long long block_sum64(long long v, long long* red) {
  v = wave_sum64(v);
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  long long s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / kWave; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Block sums of N 64-bit values at once (two barriers for all of them);
// thread 0 receives the sums, red holds N * NT / 64 entries.
template <int NT, int N>
__device__ __forceinline__ void block_sums64(long long (&v)[N], long long* red) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    v[i] = wave_sum64(v[i]);
    if (lane_id() == 0) red[i * (NT / kWave) + threadIdx.x / kWave] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      long long s = 0;
      for (int w = 0; w < NT / kWave; ++w) s += red[i * (NT / kWave) + w];
      v[i] = s;
    }
  __syncthreads();
}

// Zeroes counts[i] for every bit i of `mask` (the solver's list counts a
// later launch appends to), from thread 0 of block 0.
__device__ __forceinline__ void zero_counts(int32_t* counts, unsigned mask) {
  if (counts && blockIdx.x == 0 && threadIdx.x == 0)
    for (int i = 0; mask; ++i, mask >>= 1)
      if (mask & 1) counts[i] = 0;
}

// Thread-0 accumulator of per-image sums that flushes (one slotted atomic)
// whenever the image changes.
struct ImgAcc {
  int img = -1;
  long long v = 0;
  __device__ __forceinline__ void add(unsigned long long* sums, int row, int im, long long x) {
    if (im != img) {
      flush(sums, row);
      img = im;
    }
    v += x;
  }
  __device__ __forceinline__ void flush(unsigned long long* sums, int row) {
    if (img >= 0 && v)
      atomicAdd(&sums[((size_t)img * SUM_ROWS + row) * kSumSlots + blockIdx.x % kSumSlots],
                (unsigned long long)v);
    v = 0;
  }
};

__device__ __forceinline__ int pixel_mask(int e, int cu, int cd, int cl, int cr) {
  return (cu > 0 ? M_UP : 0) | (cd > 0 ? M_DOWN : 0) | (cl > 0 ? M_LEFT : 0) |
         (cr > 0 ? M_RIGHT : 0) | (e < 0 ? M_SINK : 0) | (e > 0 ? M_EXCESS : 0);
}

// Halo position of thread j < 4T (side = j / T): the neighbour tile (-1 if
// outside the image), the pixel inside it, the LDS index in a (T+2)^2
// halo-padded array and the pixel's image coordinates.
struct Halo {
  int nb, nx, ny, idx, gx, gy;
};

template <int T>
__device__ __forceinline__ Halo halo_slot(const TG& q, int j) {
  constexpr int HP = T + 2;
  const int side = j / T, k = j % T;
  Halo h;
  h.nb = -1;
  h.nx = 0;
  h.ny = 0;
  if (side == 0) {
    if (q.ty > 0) { h.nb = q.t - q.tw; h.nx = k; h.ny = T - 1; }
    h.idx = k + 1;
    h.gx = q.tx * T + k;
    h.gy = q.ty * T - 1;
  } else if (side == 1) {
    if (q.ty < q.th - 1) { h.nb = q.t + q.tw; h.nx = k; h.ny = 0; }
    h.idx = (T + 1) * HP + k + 1;
    h.gx = q.tx * T + k;
    h.gy = q.ty * T + T;
  } else if (side == 2) {
    if (q.tx > 0) { h.nb = q.t - 1; h.nx = T - 1; h.ny = k; }
    h.idx = (k + 1) * HP;
    h.gx = q.tx * T - 1;
    h.gy = q.ty * T + k;
  } else {
    if (q.tx < q.tw - 1) { h.nb = q.t + 1; h.nx = 0; h.ny = k; }
    h.idx = (k + 1) * HP + T + 1;
    h.gx = q.tx * T + T;
    h.gy = q.ty * T + k;
  }
  return h;
}

// ----------------------------------------------------------------- push
// One residency of each listed tile: load state (FIRST: build it from the
// row-major capacities), fold pending inflow, freeze the halo heights, run
// up to `sweeps` synchronous push/relabel sweeps in LDS, write back state,
// residual mask and the tile's positive-excess sum, and post edge outflow to
// the neighbours' pending slots.
//
// FIRST (an image's first round, over all its tiles): e = cs - ct, caps from
// the pairwise weights, h = 0 on sink pixels (e < 0) and 1 elsewhere -- a
// valid labelling, since every residual arc then satisfies h(p) <= h(q) + 1
// -- and the image's sums get sum min(cs, ct) and sum max(e0, 0).  Its
// heights are stored only if store_h (a full relabel that comes next
// rewrites them unread; later rounds always store).
//
// A sweep: every active pixel (e > 0, h < MAX) pushes min(rem, cap) along
// each admissible arc (h(p) == h(q) + 1) in the order down, up, right, left
// (the reference's pass order, GraphCut.pyx:278-315); inflows are gathered;
// then every pixel that is still active and has no admissible arc relabels
// to min(h(q) + 1) over residual arcs (MAX if none), all from the heights
// of the sweep's start (graphcut.cl:538-577 is the same Jacobi relabel).
// Heights are double-buffered in LDS, so a sweep needs two barriers.
//
// Pending protocol (no fences: an agent-scope fence writes back the whole
// L2 on this part): a sender atomically adds its edge outflow to the
// receiver's slot and sets pflag[receiver] and act[receiver] with plain
// stores.  A resident receiver reads pflag as a hint and, if set, exchanges
// its slots with 0; it never clears pflag.  Whatever lands in a slot after
// that exchange stays there with pflag still set, and k_fold (a separate
// launch, so every store of this kernel is visible) folds it and clears
// the flag.  No inflow is lost or counted twice.
// R consecutive int32 of a 4R-byte aligned address, and back.
template <int R>
__device__ __forceinline__ void load_row(const int32_t* __restrict__ p, int (&v)[R]) {
  if constexpr (R == 4) {
    const int4 a = *reinterpret_cast<const int4*>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  } else if constexpr (R == 2) {
    const int2 a = *reinterpret_cast<const int2*>(p);
    v[0] = a.x; v[1] = a.y;
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = p[k];
  }
}

template <int R>
__device__ __forceinline__ void store_row(int32_t* p, const int (&v)[R]) {
  if constexpr (R == 4) {
    *reinterpret_cast<int4*>(p) = make_int4(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 2) {
    *reinterpret_cast<int2*>(p) = make_int2(v[0], v[1]);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) p[k] = v[k];
  }
}

// Waves per SIMD the push kernel is compiled for: at T = 32 five blocks of
// 256 threads per CU (96 registers).
template <int T>
constexpr int kPushWaves = T == 32 ? 5 : 1;

// The value `v` of the lane one to the left (right) in this lane's 16-lane
// row, by DPP (0 where that lane lies outside the row).  A tile row's
// threads are consecutive lanes aligned to their count (8 or 16), so for
// every thread but the first (last) of a tile row this is the thread
// holding the pixels to its left (right).  Every lane of the wave must be
// active.
__device__ __forceinline__ int from_left_lane(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
}
__device__ __forceinline__ int from_right_lane(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x101, 0xf, 0xf, false);  // row_shl:1
}

// Round 1's inputs of one thread's segment (row-major capacities): unaries,
// right and down weights, the down weight of the row above and the right
// weight left of the segment (0 outside the image), and for the halo
// loaders the unaries of their halo pixel.
template <int R>
struct PushIn {
  int a[R], b[R], rr[R], dd[R], du[R];
  int rl;
  int ha, hb;
  bool hin;
};

template <int T, int R>
__device__ __forceinline__ void push_in_load(const CutDev& s, const TG& q, int ly, int x0,
                                             PushIn<R>& in) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    in.a[k] = 0; in.b[k] = 0; in.rr[k] = 0; in.dd[k] = 0; in.du[k] = 0;
  }
  in.rl = 0;
  const int y = q.ty * T + ly, xg = q.tx * T + x0;
  if (y < q.H && xg < q.W) {
    const size_t p = q.pix0 + (size_t)y * q.W + xg;
    if (R > 1 && q.W % R == 0) {
      // whole aligned segment (pixel ranges start on a multiple of 4)
      load_row<R>(s.cs + p, in.a);
      load_row<R>(s.ct + p, in.b);
      load_row<R>(s.right + p, in.rr);
      load_row<R>(s.down + p, in.dd);
      if (y > 0) load_row<R>(s.down + p - q.W, in.du);
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (xg + k < q.W) {
          in.a[k] = s.cs[p + k];
          in.b[k] = s.ct[p + k];
          in.rr[k] = s.right[p + k];
          in.dd[k] = s.down[p + k];
          if (y > 0) in.du[k] = s.down[p + k - q.W];
        }
    }
    if (xg > 0) in.rl = s.right[p - 1];
  }
  in.hin = false;
  in.ha = in.hb = 0;
  if (threadIdx.x < 4 * T) {
    const Halo hs = halo_slot<T>(q, threadIdx.x);
    if (hs.nb >= 0 && hs.gx < q.W && hs.gy < q.H) {
      const size_t p = q.pix0 + (size_t)hs.gy * q.W + hs.gx;
      in.hin = true;
      in.ha = s.cs[p];
      in.hb = s.ct[p];
    }
  }
}

template <int T, int NT, bool FIRST>
__global__ void __launch_bounds__(NT, kPushWaves<T>)
k_push(Geo g, CutDev s, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
       int sweeps, int store_h) {
  // thread -> R consecutive pixels of one tile row: row ly, columns x0 .. x0+R-1
  constexpr int R = T * T / NT;
  constexpr int CPR = T / R;  // threads per row
  // heights in LDS: T + 2 rows (halo above and below) of HS entries, pixel
  // (x, y) at (y + 1) * HS + 4 + x, so each thread's segment is 16-byte
  // aligned; the halo columns sit at 3 and 4 + T
  constexpr int HS = T + 8;
  constexpr int HN = (T + 2) * HS;
  static_assert(4 * T <= NT, "halo loaders");
  static_assert(T % R == 0 && 16 % CPR == 0, "a tile row's threads share a 16-lane row");
  __shared__ __align__(16) int32_t H0[HN], H1[HN];
  // a sweep's vertical flows (the horizontal ones pass between lanes)
  __shared__ __align__(16) int32_t FU[T * T], FD[T * T];
  // the residual of each edge pixel's arc out of the tile at the start of
  // the sweeps (up, down, left, right side); written and read back by the
  // thread owning the pixel: the flow it sent out is the start value minus
  // the final one (nothing flows in over such an arc during a residency)
  __shared__ __align__(16) int32_t EC[4][T];
  __shared__ long long red[3 * (NT / kWave)];
  __shared__ int sflag;
  const int seg = threadIdx.x % CPR;
  const int ly = threadIdx.x / CPR, x0 = seg * R;
  const bool row0 = ly == 0, rowL = ly == T - 1, seg0 = seg == 0, segL = seg == CPR - 1;
  const int32_t MAXH = g.maxh;
  const int n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) {
    atomicAdd(&s.tally[TALLY_PUSH_TILES], (unsigned long long)n);
    atomicAdd(&s.tally[TALLY_PUSH_ROUNDS], 1ull);
  }
  const XcdRun run = xcd_run(blockIdx.x, gridDim.x, n);
  const int G = run.step, rpos = run.pos, rend = run.end;
  ImgAcc amin, apos;
  int tcur = rpos < rend ? list_tile_s(list, rpos) : 0;
  int tnext = rpos + G < rend ? list_tile_s(list, rpos + G) : 0;
  const int hrow = (ly + 1) * HS + 4 + x0;  // this thread's segment in H0 / H1
  const int i0 = ly * T + x0;               // ... in the flow planes
  const int iu = (row0 ? 0 : ly - 1) * T + x0, id = (rowL ? ly : ly + 1) * T + x0;
  for (int pos = rpos; pos < rend; pos += G) {
    const TG q = tile_geo_s(g, tcur);
    const int t = q.t;
    const size_t row = tbase<T>(t) + ly * T + x0;  // this thread's segment, tile-blocked
    int e[R], h[R], cu[R], cd[R], cl[R], cr[R];
    long long smin = 0, spos = 0;
    int hv0 = MAXH;  // the halo loaders' halo height
    if constexpr (FIRST) {
      PushIn<R> in;
      push_in_load<T, R>(s, q, ly, x0, in);
      const int y = q.ty * T + ly, xg = q.tx * T + x0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const bool in_x = xg + k < q.W;
        cr[k] = xg + k < q.W - 1 ? in.rr[k] : 0;
        cl[k] = in_x ? (k == 0 ? in.rl : in.rr[k - 1]) : 0;
        cd[k] = y < q.H - 1 ? in.dd[k] : 0;
        cu[k] = in.du[k];
        e[k] = in.a[k] - in.b[k];
        smin += min(in.a[k], in.b[k]);
        spos += e[k] > 0 ? e[k] : 0;
        h[k] = e[k] < 0 ? 0 : 1;
      }
      if (in.hin) hv0 = in.ha - in.hb < 0 ? 0 : 1;
      // the image sums' wave totals go to LDS now (their registers are
      // free for the sweeps)
      smin = wave_sum64(smin);
      spos = wave_sum64(spos);
      if (lane_id() == 0) {
        red[1 * (NT / kWave) + threadIdx.x / kWave] = smin;
        red[2 * (NT / kWave) + threadIdx.x / kWave] = spos;
      }
    } else {
      load_row<R>(s.e + row, e);
      load_row<R>(s.h + row, h);
      load_row<R>(s.cu + row, cu);
      load_row<R>(s.cd + row, cd);
      load_row<R>(s.cl + row, cl);
      load_row<R>(s.cr + row, cr);
      if (threadIdx.x < 4 * T) {
        const Halo hs = halo_slot<T>(q, threadIdx.x);
        if (hs.nb >= 0) hv0 = s.h[tbase<T>(hs.nb) + hs.ny * T + hs.nx];
      }
    }
    tcur = tnext;
    if (pos + 2 * G < rend) tnext = list_tile_s(list, pos + 2 * G);
    store_row<R>(H0 + hrow, h);
    store_row<R>(H1 + hrow, h);
    if (threadIdx.x < 4 * T) {
      const int side = threadIdx.x / T, j = threadIdx.x % T;
      const int hidx = side == 0   ? 4 + j
                       : side == 1 ? (T + 1) * HS + 4 + j
                       : side == 2 ? (j + 1) * HS + 3
                                   : (j + 1) * HS + 4 + T;
      H0[hidx] = hv0;
      H1[hidx] = hv0;
    }
    // round 1 starts from reset tiles: nothing pending
    if (threadIdx.x == 0) sflag = FIRST ? 0 : s.pflag[t];
    __syncthreads();
    // (later rounds) nonzero once a flow or a height of this thread's
    // pixels changed: a thread whose pixels kept their state stores nothing
    int chg = 0;
    // the halo heights at this row's ends (frozen for the residency)
    const int hL = H0[(ly + 1) * HS + 3], hR = H0[(ly + 1) * HS + 4 + T];
    if (sflag) {
      // issue every exchange first, then use the results (one wait)
      int vu[R], vd[R], vl[R], vr[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int x = x0 + k;
        vu[k] = row0 ? atomicExch(&s.pend[((size_t)t * 4 + 0) * T + x], 0) : 0;
        vd[k] = rowL ? atomicExch(&s.pend[((size_t)t * 4 + 1) * T + x], 0) : 0;
        vl[k] = x == 0 ? atomicExch(&s.pend[((size_t)t * 4 + 2) * T + ly], 0) : 0;
        vr[k] = x == T - 1 ? atomicExch(&s.pend[((size_t)t * 4 + 3) * T + ly], 0) : 0;
      }
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if constexpr (!FIRST) chg |= vu[k] | vd[k] | vl[k] | vr[k];
        e[k] += vu[k] + vd[k] + vl[k] + vr[k];
        cu[k] += vu[k];
        cd[k] += vd[k];
        cl[k] += vl[k];
        cr[k] += vr[k];
      }
    }
    if (row0 || rowL) store_row<R>(&EC[row0 ? 0 : 1][x0], row0 ? cu : cd);
    if (seg0) EC[2][ly] = cl[0];
    if (segL) EC[3][ly] = cr[R - 1];
    int any0 = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) any0 |= e[k] > 0 && h[k] < MAXH;
    int active = __syncthreads_or(any0);

    for (int sw = 0; sw < sweeps && active; ++sw) {
      const int32_t* Hc = (sw & 1) ? H1 : H0;
      int32_t* Hn = (sw & 1) ? H0 : H1;
      // heights at the start of the sweep: rows above / below from LDS,
      // left / right from this thread or the neighbouring lanes (the halo
      // at the tile's edges)
      int hu[R], hd[R];
      const int hlx = from_left_lane(h[R - 1]), hrx = from_right_lane(h[0]);
      const int hl0 = seg0 ? hL : hlx, hrR = segL ? hR : hrx;
      // push along admissible arcs, down / up / right / left; an inactive
      // pixel pushes min(0, cap) = 0 (capacities are never negative); a
      // wave with no active pixel pushes nothing
      int fu[R], fd[R], fl[R], fr[R];
      int wact = 0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        fu[k] = fd[k] = fl[k] = fr[k] = 0;
        wact |= e[k] > 0 && h[k] < MAXH;
      }
      if (__any(wact)) {
        load_row<R>(Hc + hrow - HS, hu);
        load_row<R>(Hc + hrow + HS, hd);
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int hv = h[k];
          const int hl = k > 0 ? h[k - 1] : hl0;
          const int hr = k < R - 1 ? h[k + 1] : hrR;
          int rem = (e[k] > 0 && hv < MAXH) ? e[k] : 0;
          fd[k] = hv == hd[k] + 1 ? min(rem, cd[k]) : 0;
          rem -= fd[k];
          fu[k] = hv == hu[k] + 1 ? min(rem, cu[k]) : 0;
          rem -= fu[k];
          fr[k] = hv == hr + 1 ? min(rem, cr[k]) : 0;
          rem -= fr[k];
          fl[k] = hv == hl + 1 ? min(rem, cl[k]) : 0;
          if constexpr (!FIRST) chg |= fd[k] | fu[k] | fr[k] | fl[k];
          cd[k] -= fd[k];
          cu[k] -= fu[k];
          cr[k] -= fr[k];
          cl[k] -= fl[k];
          e[k] -= fd[k] + fu[k] + fr[k] + fl[k];
        }
      }
      store_row<R>(FU + i0, fu);
      store_row<R>(FD + i0, fd);
      // inflow from the left / right neighbours inside the tile: from the
      // neighbouring lanes at the segment's ends, and inside the segment
      // now (every pixel has pushed), so only two values outlive the barrier
      const int vlx = from_left_lane(fr[R - 1]), vrx = from_right_lane(fl[0]);
      const int vl0 = seg0 ? 0 : vlx, vrR = segL ? 0 : vrx;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int vl = k > 0 ? fr[k - 1] : 0;
        const int vr = k < R - 1 ? fl[k + 1] : 0;
        e[k] += vl + vr;
        cl[k] += vl;
        cr[k] += vr;
      }
      __syncthreads();
      // gather inflow: the sender's arc toward this pixel is spent, so the
      // reverse arc (back to the sender) gains the same capacity
      int vu[R], vd[R];
      load_row<R>(FD + iu, vu);
      load_row<R>(FU + id, vd);
      int any = 0;
      wact = 0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int inu = row0 ? 0 : vu[k], ind = rowL ? 0 : vd[k];
        const int vl = k == 0 ? vl0 : 0;
        const int vr = k == R - 1 ? vrR : 0;
        if constexpr (!FIRST) chg |= vl | vr | inu | ind;
        e[k] += vl + vr + inu + ind;
        cl[k] += vl;
        cr[k] += vr;
        cu[k] += inu;
        cd[k] += ind;
        wact |= e[k] > 0 && h[k] < MAXH;
      }
      // relabel each pixel still active with no admissible arc, to the
      // lowest neighbour over a residual arc + 1 (MAX if none); a wave with
      // no active pixel keeps its heights
      if (__any(wact)) {
        load_row<R>(Hc + hrow - HS, hu);
        load_row<R>(Hc + hrow + HS, hd);
        int hn[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int hv = h[k];
          const int hl = k > 0 ? h[k - 1] : hl0;
          const int hr = k < R - 1 ? h[k + 1] : hrR;
          const int nd = cd[k] > 0 ? hd[k] + 1 : MAXH;
          const int nu = cu[k] > 0 ? hu[k] + 1 : MAXH;
          const int nr = cr[k] > 0 ? hr + 1 : MAXH;
          const int nl = cl[k] > 0 ? hl + 1 : MAXH;
          const bool act = e[k] > 0 && hv < MAXH;
          // (hv < MAX for an active pixel, so nX == hv means arc X is
          // residual and admissible)
          const bool adm = nd == hv || nu == hv || nr == hv || nl == hv;
          const int nh = min(min(min(nd, nu), min(nr, nl)), MAXH);
          const int nh2 = act && !adm ? nh : hv;
          any |= act && nh2 < MAXH;
          hn[k] = nh2;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if constexpr (!FIRST) chg |= hn[k] ^ h[k];
          h[k] = hn[k];
        }
      }
      store_row<R>(Hn + hrow, h);
      active = __syncthreads_or(any);
    }

    long long tp = 0;
    uint32_t mk = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      mk |= (uint32_t)pixel_mask(e[k], cu[k], cd[k], cl[k], cr[k]) << (8 * k);
      tp += e[k] > 0 ? e[k] : 0;
    }
    if (FIRST || chg) {
      store_row<R>(s.e + row, e);
      if (!FIRST || store_h) store_row<R>(s.h + row, h);
      store_row<R>(s.cu + row, cu);
      store_row<R>(s.cd + row, cd);
      store_row<R>(s.cl + row, cl);
      store_row<R>(s.cr + row, cr);
      if constexpr (R == 4) {
        *reinterpret_cast<uint32_t*>(s.mask + row) = mk;
      } else if constexpr (R == 2) {
        *reinterpret_cast<uint16_t*>(s.mask + row) = (uint16_t)mk;
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) s.mask[row + k] = (uint8_t)(mk >> (8 * k));
      }
    }
    // edge outflow -> neighbour's pending slot (the receiving side is the
    // opposite one); a nonzero outflow implies the arc and so the neighbour
    if ((row0 && q.ty > 0) || (rowL && q.ty < q.th - 1)) {
      const int nb = row0 ? t - q.tw : t + q.tw;
      int c0[R];
      load_row<R>(&EC[row0 ? 0 : 1][x0], c0);
      bool sent = false;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int out = c0[k] - (row0 ? cu[k] : cd[k]);
        if (out) {
          atomicAdd(&s.pend[((size_t)nb * 4 + (row0 ? 1 : 0)) * T + x0 + k], out);
          sent = true;
        }
      }
      if (sent) {
        s.pflag[nb] = 1;
        s.act[nb] = 1;
        s.dirty[nb] = 1;
      }
    }
    if ((seg0 && q.tx > 0) || (segL && q.tx < q.tw - 1)) {
      const int nb = seg0 ? t - 1 : t + 1;
      const int out = seg0 ? EC[2][ly] - cl[0] : EC[3][ly] - cr[R - 1];
      if (out) {
        atomicAdd(&s.pend[((size_t)nb * 4 + (seg0 ? 3 : 2)) * T + ly], out);
        s.pflag[nb] = 1;
        s.act[nb] = 1;
        s.dirty[nb] = 1;
      }
    }
    if (FIRST) {
      // (rows 1 and 2 of red hold the image sums' wave totals)
      tp = wave_sum64(tp);
      if (lane_id() == 0) red[threadIdx.x / kWave] = tp;
      __syncthreads();
      if (threadIdx.x == 0) {
        tp = smin = spos = 0;
        for (int w = 0; w < NT / kWave; ++w) {
          tp += red[w];
          smin += red[NT / kWave + w];
          spos += red[2 * (NT / kWave) + w];
        }
      }
      __syncthreads();
    } else {
      tp = block_sum64<NT>(tp, red);
    }
    if (threadIdx.x == 0) {
      s.tpos[t] = tp;
      s.dirty[t] = 1;
      if (active) s.act[t] = 1;
      if (FIRST) {
        amin.add(s.sums, SUM_MIN, q.img, smin);
        apos.add(s.sums, SUM_POS0, q.img, spos);
      }
    }
  }
  if (FIRST && threadIdx.x == 0) {
    amin.flush(s.sums, SUM_MIN);
    apos.flush(s.sums, SUM_POS0);
  }
}

// Folds the pending inflow of every flagged tile into its state (between
// launches, when no tile is resident): e, the reverse capacity of each side
// that received inflow, the mask byte and the tile's positive-excess sum.
// A block walks listed tiles; thread j < 4T - 4 owns perimeter pixel j (the
// top and bottom rows, then the left and right columns), so corners are not
// raced.  Inflow only raises a residual, so the mask's residual bits are set
// for the sides that received some and kept for the others, and only the
// one residual of each receiving side is read and written.  Thread 0 alone
// reads and clears the tiles' flags, the next tile's flag loading while the
// block folds the current one.  The block is the perimeter rounded up to
// whole waves: 64, 128 and 256 threads at T = 16, 32 and 64.  (One wave per
// tile, with four tiles per block, was measured 2x slower on the 4096^2
// cut, and a grid of twice the blocks 40% slower: the scattered column
// accesses, not latency, bound the kernel.)
template <int T>
constexpr int kFoldNT = (4 * T - 4 + 63) / 64 * 64;

template <int T>
__global__ void __launch_bounds__(kFoldNT<T>)
k_fold(Geo g, CutDev s, const int32_t* __restrict__ list, const int32_t* __restrict__ count) {
  constexpr int P = 4 * T - 4;
  __shared__ int sflag[2];
  const int n = *count;
  const int j = threadIdx.x;
  // this thread's pixel, and the pending slot of its vertical side (top /
  // bottom row) and of its horizontal side (left / right column), -1 where
  // it has none; each slot belongs to one pixel
  int x = 0, y = 0;
  if (j < T) { x = j; y = 0; }
  else if (j < 2 * T) { x = j - T; y = T - 1; }
  else if (j < 3 * T - 2) { x = 0; y = j - 2 * T + 1; }
  else if (j < P) { x = T - 1; y = j - (3 * T - 2) + 1; }
  const int off = y * T + x;
  const int vs = j >= P ? -1 : (y == 0 ? x : (y == T - 1 ? T + x : -1));
  const int hs = j >= P ? -1 : (x == 0 ? 2 * T + y : (x == T - 1 ? 3 * T + y : -1));
  const bool up = y == 0, lf = x == 0;
  const int p0 = xcd_pos(blockIdx.x, gridDim.x);
  int tnext = p0 < n ? list_tile(list, p0) : 0;
  if (j == 0) sflag[0] = p0 < n ? s.pflag[tnext] : 0;
  int par = 0;
  for (int pos = p0; pos < n; pos += gridDim.x, par ^= 1) {
    const int t = tnext;
    const bool more = pos + (int)gridDim.x < n;
    if (more) tnext = list_tile(list, pos + gridDim.x);
    __syncthreads();  // sflag[par] is written; every thread has read sflag[par ^ 1]
    const int flagged = sflag[par];
    if (j == 0) {
      if (flagged) s.pflag[t] = 0;
      sflag[par ^ 1] = more ? s.pflag[tnext] : 0;
    }
    if (!flagged) continue;
    int32_t* pe = s.pend + (size_t)t * 4 * T;
    const size_t b = tbase<T>(t);
    const int sv = vs >= 0 ? pe[vs] : 0;
    const int sh = hs >= 0 ? pe[hs] : 0;
    // unconditional state loads (a thread with nothing to fold reads the
    // tile's first pixel, whose line other threads read anyway), so that
    // no select waits for a load
    const size_t i = (sv | sh) ? b + off : b;
    const size_t iv = sv ? i : b, ih = sh ? i : b;
    const int e0 = s.e[i];
    const int m0 = s.mask[i];
    const int cv = (up ? s.cu : s.cd)[iv];
    const int ch = (lf ? s.cl : s.cr)[ih];
    long long delta = 0;
    if (sv | sh) {
      const int e1 = e0 + sv + sh;
      s.e[i] = e1;
      if (sv) (up ? s.cu : s.cd)[i] = cv + sv;
      if (sh) (lf ? s.cl : s.cr)[i] = ch + sh;
      const int m = (m0 & (M_UP | M_DOWN | M_LEFT | M_RIGHT)) |
                    (sv ? (up ? M_UP : M_DOWN) : 0) | (sh ? (lf ? M_LEFT : M_RIGHT) : 0) |
                    (e1 < 0 ? M_SINK : 0) | (e1 > 0 ? M_EXCESS : 0);
      s.mask[i] = (uint8_t)m;
      if (sv) pe[vs] = 0;
      if (sh) pe[hs] = 0;
      delta = (long long)(e1 > 0 ? e1 : 0) - (e0 > 0 ? e0 : 0);
    }
    // the change of the tile's positive-excess sum, one atomic per wave
    delta = wave_sum64(delta);
    if (lane_id() == 0 && delta)
      atomicAdd(reinterpret_cast<unsigned long long*>(&s.tpos[t]), (unsigned long long)delta);
  }
}

// -------------------------------------------------------- tile queues
// The BFS and flood rounds pass work on through queues of tiles: a round
// reads its input list and appends the neighbour tiles it changed to its
// output list.  A tile enters an output list once: the first appender swaps
// the round's epoch into the tile's stamp.  Appends are gathered in LDS and
// reserved with one atomic per block flush.  Rounds rotate three lists: a
// round reads list i, appends to list i+1 and zeroes the count of list
// i+2 (read by the previous round, written by the next).
// The epochs live on the device, so that a recorded sequence of rounds can
// be replayed: a round reads its epoch from one slot and writes epoch + 1 to
// the other, which the next round reads (the slot it writes was last read by
// the previous round).
struct Queue {
  const int32_t* in;     // input tiles (null: input positions are tiles)
  const int32_t* in_count;
  int32_t* out;
  int32_t* out_count;
  int32_t* clear_count;  // zeroed by the launch (may be null)
  const int32_t* ep_in;  // this round's epoch: the stamp of tiles appended to out
  int32_t* ep_out;       // receives epoch + 1
};

constexpr int kQCap = 256;
struct BlockQ {
  int32_t buf[kQCap];
  int n, base, dirs, epoch;
};

__device__ __forceinline__ void bq_init(BlockQ& b, const Queue& Q) {
  if (threadIdx.x == 0) {
    b.n = 0;
    b.dirs = 0;
    b.epoch = *Q.ep_in;
    if (blockIdx.x == 0) {
      if (Q.clear_count) *Q.clear_count = 0;
      *Q.ep_out = b.epoch + 1;
    }
  }
  __syncthreads();
}

// Copies the block's buffer to the output list.  Called by all threads, a
// barrier after the last change of b.n.
__device__ __forceinline__ void bq_flush(BlockQ& b, const Queue& Q) {
  const int n = b.n;
  if (n == 0) return;
  __syncthreads();
  if (threadIdx.x == 0) b.base = atomicAdd(Q.out_count, n);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) Q.out[b.base + i] = b.buf[i];
  __syncthreads();
  if (threadIdx.x == 0) b.n = 0;
  __syncthreads();
}

// ------------------------------------------------------------ bit rows
// The relabel and flood kernels work on bit rows: lane y of a T-lane
// segment holds row y of one tile as a T-bit word (a frontier, a visited or
// reached set, and each arc kind of the mask), so a wave holds 64 / T tiles.
template <int T>
struct BitRow {
  using type = unsigned;
};
template <>
struct BitRow<64> {
  using type = unsigned long long;
};

// Bit k of each of w's four bytes, as four bits.
__device__ __forceinline__ unsigned byte_bits(unsigned w, int k) {
  return (((w >> k) & 0x01010101u) * 0x01020408u) >> 24;
}
// Four bits as four bytes of 0 / 1.
__device__ __forceinline__ unsigned bits_bytes(unsigned n) {
  return (n * 0x00204081u) & 0x01010101u;
}

// Row y of a tile's byte array (T bytes, 16-byte aligned) as bit rows of
// bits k0 .. k0 + NK - 1 of each byte.
template <int T, int NK>
__device__ __forceinline__ void load_bit_rows(const uint8_t* row, int k0,
                                              typename BitRow<T>::type (&out)[NK]) {
  using Row = typename BitRow<T>::type;
#pragma unroll
  for (int k = 0; k < NK; ++k) out[k] = 0;
#pragma unroll
  for (int c = 0; c < T / 16; ++c) {
    const uint4 v = reinterpret_cast<const uint4*>(row)[c];
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < NK; ++k) out[k] |= (Row)byte_bits(w[i], k0 + k) << (16 * c + 4 * i);
  }
}

template <int T>
__device__ __forceinline__ void store_bit_row(uint8_t* row, typename BitRow<T>::type r) {
#pragma unroll
  for (int c = 0; c < T / 16; ++c) {
    const unsigned n = (unsigned)(r >> (16 * c));
    reinterpret_cast<uint4*>(row)[c] = make_uint4(bits_bytes(n & 15), bits_bytes((n >> 4) & 15),
                                                  bits_bytes((n >> 8) & 15),
                                                  bits_bytes((n >> 12) & 15));
  }
}

template <typename Row>
__device__ __forceinline__ Row seg_shfl_up(Row v, int d, int T) {
  return __shfl_up(v, d, T);
}
template <typename Row>
__device__ __forceinline__ Row seg_shfl_down(Row v, int d, int T) {
  return __shfl_down(v, d, T);
}

template <typename Row>
__device__ __forceinline__ int row_ctz(Row v) {
  return sizeof(Row) == 8 ? __builtin_ctzll((unsigned long long)v) : __builtin_ctz((unsigned)v);
}

// Word of a predicate over the T lanes of this lane's segment (bit j: lane j
// of the segment).  Every lane of the wave calls it.
template <int T>
__device__ __forceinline__ typename BitRow<T>::type seg_word(bool pred) {
  using Row = typename BitRow<T>::type;
  constexpr unsigned long long SEG = T == 64 ? ~0ull : ((1ull << T) - 1);
  return (Row)((__ballot(pred) >> (lane_id() / T * T)) & SEG);
}

// Minimum over the T lanes of this lane's segment (every lane of the
// segment calls it).
template <int T>
__device__ __forceinline__ int seg_min(int v) {
#pragma unroll
  for (int o = T / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, T));
  return v;
}

// ------------------------------------------------------ global relabel
// Distance of every pixel to the sink set {e < 0} over residual arcs
// (MAX if unreachable): what init_bfs / bfs_intratile / bfs_intertile
// (graphcut.cl:582-1016) compute, without their 30-round cap.  Inside a
// tile, a level-synchronous BFS on bit rows: a level moves the frontier one
// step against the arcs (shifts along the rows, a shuffle across them) and
// takes in the edge pixels whose halo neighbour sits one level below; an
// empty frontier jumps to the next level some halo neighbour seeds.  A
// pixel's distance is the level at which the frontier first reaches it,
// written then into the block's LDS copy of the tile.  Across tiles, a tile
// whose edge distances fell queues its neighbours for another round, until
// no tile is queued.

constexpr int kBfsNT = 256;
constexpr int kInf = 0x7fffffff;

// A block's LDS copy of a tile holds row y's 16-byte pieces in a rotated
// order: piece c of the row at piece c ^ (y % kBfsSwz<T>).  The level loop's
// lanes (one per row) often store to the same column of their rows at once;
// row-major rows T words apart would put those stores in one LDS bank, the
// rotation spreads them over kBfsSwz<T> pieces.
template <int T>
constexpr int kBfsSwz = T / 4 < 8 ? T / 4 : 8;
template <int T>
__device__ __forceinline__ int bfs_px(int y, int x) {
  return y * T + (x ^ ((y & (kBfsSwz<T> - 1)) << 2));
}
template <int T>
constexpr int kBfsTPB = (kWave / T) * (kBfsNT / kWave);  // tiles per block

// One pass over listed tiles.  INIT (the relabel's first pass): seeds are
// the tile's sinks (distance 0) and the halo: a neighbour relabelled in the
// same pass (all, or its rflag holds the pass's epoch) counts as 0 at its
// sinks and MAX elsewhere, a clean one keeps its stored distances (valid:
// its arcs and excess have not changed since they were computed, and the
// rounds lower any of them a recomputed tile now undercuts).  Writes every
// distance and, for the completion test (check_completion,
// graphcut.cl:20-60),
//   act[t]    = some pixel has e > 0 and a finite distance
//   haspos[t] = some pixel has e > 0
// and moves a dirty flag to the flood's changed flag.  Later rounds (!INIT):
// the stored distances are the tile's in-tile fixed point for the halo it
// had, and halo distances only fall, so the new distances are the minimum
// of the stored ones and a BFS from the halo alone (seeded only where a
// halo distance + 1 beats the edge pixel's stored one); a pass writes the
// rows that fell and flags the tile active if an excess pixel's fell.
// A neighbour is queued when an edge pixel's distance (new, for a round) +
// 1 beats the value the halo holds for the pixel across the edge and that
// pixel's arc into this tile is residual; in the first pass a sink is no
// news to a neighbour relabelled in the same pass (it starts from 0 at
// sinks).  Distances only fall, so a stale halo value can only queue a
// neighbour needlessly.
template <int T, bool INIT>
__device__ __forceinline__ void bfs_pass(const Geo& g, CutDev& s, const int32_t* __restrict__ list,
                                         int n, int all, const Queue& Q, BlockQ& bq,
                                         int32_t* lds) {
  using Row = typename BitRow<T>::type;
  constexpr int TPB = kBfsTPB<T>;
  constexpr Row TOP = (Row)1 << (T - 1);
  const int MAXH = g.maxh;
  const int lane = lane_id(), y = lane % T;
  const int slot = (threadIdx.x / kWave) * (kWave / T) + lane / T;
  int32_t* buf = lds + slot * T * T;  // this segment's tile, row-major
  const int stride = gridDim.x * TPB;
  const int b0 = xcd_pos(blockIdx.x, gridDim.x) * TPB;
  int pos = b0 + slot;
  int tnext = pos < n ? list_tile(list, pos) : 0;
  for (int base = b0; base < n; base += stride, pos += stride) {
    const bool valid = pos < n;
    const int t = tnext;
    if (pos + stride < n) tnext = list_tile(list, pos + stride);
    Row mr[6] = {0, 0, 0, 0, 0, 0};
    TG q{};
    // per side (D_UP, D_DOWN, D_LEFT, D_RIGHT): this lane's edge pixel on
    // that side -- (y, 0), (y, T-1), (0, y), (T-1, y) -- and the halo pixel
    // across: its distance, whether its arc into this tile is residual, and
    // (INIT) whether it is relabelled in this pass; (!INIT) the edge
    // pixel's stored distance
    int hv[4] = {MAXH, MAXH, MAXH, MAXH}, hin[4] = {0, 0, 0, 0}, hls[4] = {0, 0, 0, 0};
    int e0[4] = {kInf, kInf, kInf, kInf};
    bool ex[4] = {false, false, false, false};
    int nbt[4] = {0, 0, 0, 0};
    const int epx[4] = {y, (T - 1) * T + y, y * T, y * T + T - 1};
    if (valid) {
      q = tile_geo(g, t);
      const size_t tb = tbase<T>(t);
      load_bit_rows<T, 6>(s.mask + tb + (size_t)y * T, 0, mr);
      ex[D_UP] = q.ty > 0;
      ex[D_DOWN] = q.ty < q.th - 1;
      ex[D_LEFT] = q.tx > 0;
      ex[D_RIGHT] = q.tx < q.tw - 1;
      nbt[D_UP] = t - q.tw;
      nbt[D_DOWN] = t + q.tw;
      nbt[D_LEFT] = t - 1;
      nbt[D_RIGHT] = t + 1;
      const int npx[4] = {(T - 1) * T + y, y, y * T + T - 1, y * T};
#pragma unroll
      for (int side = 0; side < 4; ++side) {
        if (!ex[side]) continue;
        const size_t p = tbase<T>(nbt[side]) + npx[side];
        const int mq = s.mask[p];
        if (INIT) {
          // (a neighbour relabelled in the same pass needs only its mask;
          // the halo columns cost a line per pixel, so its heights are not
          // loaded)
          const bool listed = all || s.rflag[nbt[side]] == bq.epoch;
          hv[side] = listed ? ((mq & M_SINK) ? 0 : MAXH) : s.h[p];
          hls[side] = listed ? 1 : 0;
        } else {
          hv[side] = s.h[p];
          e0[side] = s.h[tb + epx[side]];
        }
        // (the arc of the halo pixel that points into this tile)
        hin[side] = (mq & (1 << (side ^ 1))) ? 1 : 0;
      }
    }
    const Row U = mr[0], D = mr[1], Lf = mr[2], Rt = mr[3], SNK = mr[4], X = mr[5];
    // the first and last rows' arc and sink words (lanes hold columns there)
    const Row u0 = __shfl(U, 0, T), dl = __shfl(D, T - 1, T);
    const Row s0 = __shfl(SNK, 0, T), sl = __shfl(SNK, T - 1, T);
    // the level at which each edge pixel of this lane is seeded from the
    // halo (kInf: not), if its arc toward the halo pixel is residual
    const int arc[4] = {(int)((u0 >> y) & 1), (int)((dl >> y) & 1), (int)(Lf & 1),
                        (int)((Rt >> (T - 1)) & 1)};
    int inj[4];
#pragma unroll
    for (int side = 0; side < 4; ++side)
      inj[side] = (arc[side] && hv[side] < MAXH && hv[side] + 1 < e0[side]) ? hv[side] + 1 : kInf;
    // the tile's distances start as 0 at sinks and MAX elsewhere (INIT) or
    // as the stored ones (a round)
#pragma unroll
    for (int c = 0; c < T / 4; ++c) {
      int4 v;
      if (INIT) {
        int w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = ((SNK >> (4 * c + k)) & 1) ? 0 : MAXH;
        v = make_int4(w[0], w[1], w[2], w[3]);
      } else {
        v = valid ? reinterpret_cast<const int4*>(s.h + tbase<T>(t) + (size_t)y * T)[c]
                  : make_int4(MAXH, MAXH, MAXH, MAXH);
      }
      reinterpret_cast<int4*>(buf + y * T)[c ^ (y & (kBfsSwz<T> - 1))] = v;
    }
    // V: pixels the frontier has reached; a round's frontier keeps only the
    // pixels whose distance fell (a pixel at or below the level cannot
    // lower its neighbours: the stored distances are a fixed point)
    Row V = INIT ? SNK : 0, F = V, fell = 0;
    int lc = INIT ? 0 : -1;  // the segment's current level
    bool done = !valid;
    for (;;) {
      if (seg_word<T>(F != 0) == 0 && !done) {
        // frontier empty: jump to the next seeded level, or finish
        int m = kInf;
#pragma unroll
        for (int side = 0; side < 4; ++side)
          if (inj[side] > lc) m = min(m, inj[side]);
        m = seg_min<T>(m);
        if (m == kInf)
          done = true;
        else
          lc = m - 1;
      }
      if (!__any(!done)) break;
      const int L = lc + 1;
      const Row fu = seg_shfl_up(F, 1, T), fd = seg_shfl_down(F, 1, T);
      Row nw = ((F >> 1) & Rt) | ((F << 1) & Lf) | ((y > 0 ? fu : 0) & U) |
               ((y < T - 1 ? fd : 0) & D);
      if (inj[D_LEFT] == L) nw |= 1;
      if (inj[D_RIGHT] == L) nw |= TOP;
      const Row wt = seg_word<T>(inj[D_UP] == L), wb = seg_word<T>(inj[D_DOWN] == L);
      if (y == 0) nw |= wt;
      if (y == T - 1) nw |= wb;
      nw &= ~V;
      if (done) nw = 0;
      V |= nw;
      if (INIT) {
        for (Row m = nw; m; m &= m - 1) buf[bfs_px<T>(y, row_ctz(m))] = L;
      } else {
        for (Row m = nw; m; m &= m - 1) {
          const int x = row_ctz(m);
          if (atomicMin(&buf[bfs_px<T>(y, x)], L) <= L) nw &= ~((Row)1 << x);
        }
        fell |= nw;
      }
      F = nw;
      lc = L;
    }
    __syncthreads();  // the block's distances are in LDS
    // this lane's row: INIT writes it, a round the 16-byte pieces that fell
    if (valid) {
      const size_t rowg = tbase<T>(t) + (size_t)y * T;
#pragma unroll
      for (int c = 0; c < T / 4; ++c)
        if (INIT || ((fell >> (4 * c)) & 15))
          reinterpret_cast<int4*>(s.h + rowg)[c] =
              reinterpret_cast<const int4*>(buf + y * T)[c ^ (y & (kBfsSwz<T> - 1))];
    }
    // news for the neighbours
    int dirs = 0;
    const Row f0 = __shfl(fell, 0, T), fl = __shfl(fell, T - 1, T);
    if (valid) {
      const int sink[4] = {(int)((s0 >> y) & 1), (int)((sl >> y) & 1), (int)(SNK & 1),
                           (int)((SNK >> (T - 1)) & 1)};
      const int fel[4] = {(int)((f0 >> y) & 1), (int)((fl >> y) & 1), (int)(fell & 1),
                          (int)((fell >> (T - 1)) & 1)};
#pragma unroll
      for (int side = 0; side < 4; ++side) {
        if (!ex[side] || !hin[side]) continue;
        const int d = buf[side == D_UP     ? bfs_px<T>(0, y)
                          : side == D_DOWN ? bfs_px<T>(T - 1, y)
                          : side == D_LEFT ? bfs_px<T>(y, 0)
                                           : bfs_px<T>(y, T - 1)];
        const bool cand = INIT ? (d < MAXH && !(hls[side] && sink[side])) : fel[side] != 0;
        if (cand && d + 1 < hv[side]) dirs |= 1 << side;
      }
    }
    // lane y < 4 of the segment queues the neighbour across side y
    bool mine = false;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const bool any = seg_word<T>((dirs >> d) & 1) != 0;
      if (y == d) mine = any;
    }
    if (valid && mine) {
      const int nb = y == D_UP ? t - q.tw : y == D_DOWN ? t + q.tw : y == D_LEFT ? t - 1 : t + 1;
      if (atomicExch(&s.stamp[nb], bq.epoch) != bq.epoch) bq.buf[atomicAdd(&bq.n, 1)] = nb;
    }
    if (INIT) {
      const bool any_pos = seg_word<T>(X != 0) != 0;
      const bool any_act = seg_word<T>((X & V) != 0) != 0;
      if (valid && y == 0) {
        s.haspos[t] = any_pos;
        s.act[t] = any_act;
        // (no other block reads this tile's dirty flag during the pass)
        if (s.dirty[t]) {
          s.fchg[t] = 1;
          s.dirty[t] = 0;
        }
      }
    } else {
      const bool any_act = seg_word<T>((fell & X) != 0) != 0;
      if (valid && y == 0 && any_act) s.act[t] = 1;
    }
    __syncthreads();
    if (bq.n > kQCap - 4 * TPB) bq_flush(bq, Q);
  }
}

// First pass over the listed (dirty) tiles, from the residual masks
// (pending inflow has been folded).  all != 0: every tile of the images is
// listed (a full relabel); else the listed tiles are those whose rflag holds
// this pass's epoch (*Q.ep_in).  *zero (the count the relabel's active list
// is compacted into) is zeroed.  The neighbours this pass's distances may
// improve start the rounds' queue (Q.out; Q.in is not read).
template <int T>
__global__ void __launch_bounds__(kBfsNT)
k_bfs_init(Geo g, CutDev s, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
           int all, Queue Q, int32_t* zero) {
  __shared__ __align__(16) int32_t lds[kBfsTPB<T> * T * T];
  __shared__ BlockQ bq;
  bq_init(bq, Q);
  const int n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *zero = 0;
    if (n > 0) {
      atomicAdd(&s.tally[TALLY_RELABELS], 1ull);
      atomicAdd(&s.tally[TALLY_RELABEL_TILES], (unsigned long long)n);
    }
  }
  bfs_pass<T, true>(g, s, list, n, all, Q, bq, lds);
  bq_flush(bq, Q);
}

// Later passes over queued tiles, from the stored distances.
template <int T>
__global__ void __launch_bounds__(kBfsNT)
k_bfs_round(Geo g, CutDev s, Queue Q) {
  __shared__ __align__(16) int32_t lds[kBfsTPB<T> * T * T];
  __shared__ BlockQ bq;
  bq_init(bq, Q);
  const int n = *Q.in_count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0)
    atomicAdd(&s.tally[TALLY_BFS], (unsigned long long)n);
  bfs_pass<T, false>(g, s, Q.in, n, 0, Q, bq, lds);
  bq_flush(bq, Q);
}

// ------------------------------------------------------ minimal source set
// reach = pixels reachable over residual arcs from {e > 0}: the minimal
// source set once no pixel with e > 0 can reach a sink.  Same tile scheme
// as the BFS: in-tile fixed point in LDS, flags across edges.

// mode FLOOD_FIRST: over the tiles holding excess, reach bytes all cleared:
// seeds from the masks' excess bit, no halo reach, no import.
// FLOOD_FIRST_INC: over listed tiles whose reach bytes were cleared (other
// tiles keep theirs): seeds, halo reach from the neighbours' stored reach
// bytes; the tile's import flag is set when its reach differs from what its
// own excess reaches through its own arcs (no halo).  FLOOD_ROUND: over
// queued tiles, from their stored reach bytes, halo reach; a tile that
// gains reach has imported it.  A pixel newly reached on an edge with a
// residual arc across it queues that neighbour, unless the halo pixel there
// is reached already.  lmark (may be null): every processed tile is flagged
// there (its labels are rewritten after).
enum { FLOOD_ROUND = 0, FLOOD_FIRST = 1, FLOOD_FIRST_INC = 2 };

// Reach of the segment's rows closed under the arcs: g seeds; rt, lf, dn,
// up the rows' arc bits (bit x of rt: arc x -> x+1 residual, ...); y the
// lane's row.  Every lane of the wave calls it (the loop is wave-uniform).
template <int T>
__device__ __forceinline__ typename BitRow<T>::type flood_fill(typename BitRow<T>::type g,
                                                                 typename BitRow<T>::type rt,
                                                                 typename BitRow<T>::type lf,
                                                                 typename BitRow<T>::type dn,
                                                                 typename BitRow<T>::type up,
                                                                 int y) {
  using Row = typename BitRow<T>::type;
  constexpr Row SEG = T == 64 ? ~0ull : (Row)((1ull << T) - 1);
  // into x from x - 1 / x + 1; into row y from row y - 1 / y + 1
  const Row pr0 = (rt << 1) & SEG, pl0 = lf >> 1;
  Row qd0 = seg_shfl_up(dn, 1, T), qu0 = seg_shfl_down(up, 1, T);
  if (y == 0) qd0 = 0;
  if (y == T - 1) qu0 = 0;
  for (;;) {
    const Row before = g;
    Row gr = g, gl = g, pr = pr0, pl = pl0;
#pragma unroll
    for (int s = 1; s < T; s <<= 1) {
      gr |= (gr << s) & pr;
      pr &= pr << s;
      gl |= (gl >> s) & pl;
      pl &= pl >> s;
    }
    g = (gr | gl) & SEG;
    Row gd = g, gu = g, qd = qd0, qu = qu0;
#pragma unroll
    for (int s = 1; s < T; s <<= 1) {
      Row a = seg_shfl_up(gd, s, T), b = seg_shfl_up(qd, s, T);
      Row c = seg_shfl_down(gu, s, T), d = seg_shfl_down(qu, s, T);
      if (y < s) a = b = 0;
      if (y + s >= T) c = d = 0;
      gd |= a & qd;
      qd &= b;
      gu |= c & qu;
      qu &= d;
    }
    g = gd | gu;
    if (!__any(g != before)) return g;
  }
}

constexpr int kFloodNT = 256;

template <int T, int NT>
__global__ void __launch_bounds__(kFloodNT)
k_flood(Geo g, CutDev s, Queue Q, int mode, int32_t* lmark) {
  using Row = typename BitRow<T>::type;
  constexpr int TPW = kWave / T;                 // tiles per wave
  constexpr int TPB = TPW * (kFloodNT / kWave);  // tiles per block
  constexpr Row SEG = T == 64 ? ~0ull : (Row)((1ull << T) - 1);
  constexpr Row TOP = (Row)1 << (T - 1);
  const bool first = mode != FLOOD_ROUND;
  __shared__ BlockQ bq;
  const int n = *Q.in_count;
  bq_init(bq, Q);
  if (!first && blockIdx.x == 0 && threadIdx.x == 0 && n > 0)
    atomicAdd(&s.tally[TALLY_FLOOD], (unsigned long long)n);
  const int lane = lane_id();
  const int seg = lane / T, y = lane % T;
  const int slot = (threadIdx.x / kWave) * TPW + seg;  // this segment's tile of the block's TPB
  const int stride = gridDim.x * TPB;
  int pos = xcd_pos(blockIdx.x, gridDim.x) * TPB + slot;
  int tnext = pos < n ? list_tile(Q.in, pos) : 0;
  for (int base = xcd_pos(blockIdx.x, gridDim.x) * TPB; base < n; base += stride, pos += stride) {
    const bool valid = pos < n;
    const int t = tnext;
    if (pos + stride < n) tnext = list_tile(Q.in, pos + stride);
    Row mr[6] = {0, 0, 0, 0, 0, 0}, r0 = 0;
    Row hin_top = 0, hin_bot = 0, hr_top = 0, hr_bot = 0;
    int hin_l = 0, hin_r = 0, hr_l = 0, hr_r = 0, any_halo_reach = 0;
    TG q{};
    if (valid) {
      q = tile_geo(g, t);
      const size_t rowp = tbase<T>(t) + (size_t)y * T;
      load_bit_rows<T, 6>(s.mask + rowp, 0, mr);
      if (!first) {
        Row rr[1];
        load_bit_rows<T, 1>(s.reach + rowp, 0, rr);
        r0 = rr[0];
      }
      // the halo pixels of this lane: (y, -1) above, (y, T) below, (-1, y)
      // left, (T, y) right; mask byte and reach byte of each
      int mt = 0, mb = 0, ml = 0, mrr = 0, rt_ = 0, rb_ = 0, rl_ = 0, rr_ = 0;
      const bool hreach = mode != FLOOD_FIRST;
      if (q.ty > 0) {
        const size_t p = tbase<T>(t - q.tw) + (size_t)(T - 1) * T + y;
        mt = s.mask[p];
        if (hreach) rt_ = s.reach[p];
      }
      if (q.ty < q.th - 1) {
        const size_t p = tbase<T>(t + q.tw) + y;
        mb = s.mask[p];
        if (hreach) rb_ = s.reach[p];
      }
      if (q.tx > 0) {
        const size_t p = tbase<T>(t - 1) + (size_t)y * T + (T - 1);
        ml = s.mask[p];
        if (hreach) rl_ = s.reach[p];
      }
      if (q.tx < q.tw - 1) {
        const size_t p = tbase<T>(t + 1) + (size_t)y * T;
        mrr = s.mask[p];
        if (hreach) rr_ = s.reach[p];
      }
      // reach coming in over an arc into this tile
      hin_l = (rl_ && (ml & M_RIGHT)) ? 1 : 0;
      hin_r = (rr_ && (mrr & M_LEFT)) ? 1 : 0;
      hr_l = rl_ ? 1 : 0;
      hr_r = rr_ ? 1 : 0;
      any_halo_reach = (rt_ | rb_ | rl_ | rr_) ? 1 : 0;
      // (the top / bottom rows' words gather below, by ballot)
      hin_top = (rt_ && (mt & M_DOWN)) ? 1 : 0;
      hin_bot = (rb_ && (mb & M_UP)) ? 1 : 0;
      hr_top = rt_ ? 1 : 0;
      hr_bot = rb_ ? 1 : 0;
    }
    // segment words of the top and bottom halo rows (bit x from lane x)
    const int sh = seg * T;
    const Row w_in_top = (Row)((__ballot(hin_top != 0) >> sh) & SEG);
    const Row w_in_bot = (Row)((__ballot(hin_bot != 0) >> sh) & SEG);
    const Row w_r_top = (Row)((__ballot(hr_top != 0) >> sh) & SEG);
    const Row w_r_bot = (Row)((__ballot(hr_bot != 0) >> sh) & SEG);
    const bool seg_halo = ((__ballot(any_halo_reach != 0) >> sh) & SEG) != 0;
    const Row U = mr[0], D = mr[1], L = mr[2], Rt = mr[3], X = mr[5];
    // halo reach entering this lane's row
    Row hin = (y == 0 ? w_in_top : 0) | (y == T - 1 ? w_in_bot : 0) | (Row)hin_l |
              (hin_r ? TOP : 0);
    Row r, rl;
    if (mode == FLOOD_FIRST) {
      r = rl = flood_fill<T>(X, Rt, L, D, U, y);
    } else if (mode == FLOOD_FIRST_INC) {
      rl = flood_fill<T>(X, Rt, L, D, U, y);
      // (a second fill only where some halo pixel is reached; the fill is
      // wave-uniform, so segments without halo reach add nothing)
      r = __any(seg_halo) ? flood_fill<T>(rl | (seg_halo ? hin : 0), Rt, L, D, U, y) : rl;
    } else {
      r = rl = flood_fill<T>(r0 | X | hin, Rt, L, D, U, y);
    }
    if (!valid) r = rl = r0 = 0;
    const Row nw = r & ~r0;
    if (nw) store_bit_row<T>(s.reach + tbase<T>(t) + (size_t)y * T, r);
    // news: a newly reached edge pixel with a residual arc across the edge,
    // the halo pixel there not reached
    int dirs = 0;
    if (valid) {
      if (y == 0 && q.ty > 0 && (nw & U & ~w_r_top)) dirs |= 1 << D_UP;
      if (y == T - 1 && q.ty < q.th - 1 && (nw & D & ~w_r_bot)) dirs |= 1 << D_DOWN;
      if (q.tx > 0 && (nw & L & 1) && !hr_l) dirs |= 1 << D_LEFT;
      if (q.tx < q.tw - 1 && (nw & Rt & TOP) && !hr_r) dirs |= 1 << D_RIGHT;
    }
    const int imported = first ? ((r & ~rl) != 0) : (nw != 0);
    // per segment: the OR of its lanes' news, and whether it imported
    unsigned long long bd[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) bd[d] = (__ballot((dirs >> d) & 1) >> sh) & SEG;
    const bool seg_imp = ((__ballot(imported) >> sh) & SEG) != 0;
    if (valid && y < 4 && bd[y]) {
      const int nb = y == D_UP ? t - q.tw : y == D_DOWN ? t + q.tw : y == D_LEFT ? t - 1 : t + 1;
      if (atomicExch(&s.stamp[nb], bq.epoch) != bq.epoch) bq.buf[atomicAdd(&bq.n, 1)] = nb;
    }
    if (valid && y == 0) {
      if (first)
        s.imp[t] = seg_imp ? 1 : 0;
      else if (seg_imp)
        s.imp[t] = 1;
      if (lmark) lmark[t] = 1;
    }
    __syncthreads();
    if (bq.n > kQCap - 4 * TPB) bq_flush(bq, Q);
  }
  bq_flush(bq, Q);
}

// The images' final sum of max(e, 0) from the listed tiles' sums: a wave
// whose 64 tiles are of one image (the lists hold each image's tiles in a
// run) adds them with one atomic.  Also zeroes the counts in zmask.
template <int T>
__global__ void k_posf(Geo g, CutDev s, const int32_t* __restrict__ list,
                       const int32_t* __restrict__ count, int32_t* counts, unsigned zmask) {
  zero_counts(counts, zmask);
  const int n = *count;
  const int stride = gridDim.x * blockDim.x;
  for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const int pos = base + threadIdx.x;
    int img = -1;
    long long v = 0;
    if (pos < n) {
      const int t = list ? list[pos] : pos;
      img = g.tile_img[t];
      v = s.tpos[t];
    }
    const int img0 = __shfl(img, 0);
    const unsigned long long row = SUM_POSF * kSumSlots + blockIdx.x % kSumSlots;
    if (__all(img == img0 || img < 0)) {
      v = wave_sum64(v);
      if (lane_id() == 0 && img0 >= 0 && v)
        atomicAdd(&s.sums[(size_t)img0 * SUM_ROWS * kSumSlots + row], (unsigned long long)v);
    } else if (img >= 0 && v) {
      atomicAdd(&s.sums[(size_t)img * SUM_ROWS * kSumSlots + row], (unsigned long long)v);
    }
  }
}

// Writes bits `b` under `m` (bit j: pixel bitpos + j, T <= 64 of them) into
// a bit-packed mask.  A word wholly under the mask belongs to this row
// segment alone and is stored; a partial one is merged atomically (other
// segments own its other bits).
__device__ __forceinline__ void put_bits(uint32_t* packed, long long bitpos,
                                         unsigned long long b, unsigned long long m) {
  uint32_t* w = packed + (bitpos >> 5);
  const int off = (int)(bitpos & 31);
  const unsigned long long mlo = m << off, blo = b << off;
  const uint32_t mw[3] = {(uint32_t)mlo, (uint32_t)(mlo >> 32),
                          off ? (uint32_t)(m >> (64 - off)) : 0u};
  const uint32_t bw[3] = {(uint32_t)blo, (uint32_t)(blo >> 32),
                          off ? (uint32_t)(b >> (64 - off)) : 0u};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (mw[j] == 0xffffffffu) {
      w[j] = bw[j];
    } else if (mw[j]) {
      atomicAnd(&w[j], ~mw[j]);
      if (bw[j] & mw[j]) atomicOr(&w[j], bw[j] & mw[j]);
    }
  }
}

// Row-major labels of the listed tiles' images from the reach bytes (a
// tile writes its own T x T block of its image), as bytes and as bits of
// the bit-packed masks.  A thread moves 16 pixels of a tile row: one
// 16-byte reach load and, where the image's rows are 16-byte aligned, one
// 16-byte label store; the T / 16 threads of a row gather its bits by
// shuffles and its first thread merges them into the packed words.  A
// block of kLabelsNT threads takes kLabelsTPB<T> tiles at a time, and the
// loop is software-pipelined: the next tiles' geometry and reach bytes load
// while this step's labels are stored (the memory a CU keeps in flight, not
// the bytes, bounds this kernel).
// The flows of a solve's images, which the last block of the labels launch
// adds up besides its tiles (the sums are final by then): image i of the
// solve is img[i * stride]; its flow is sum min(cs, ct) + sum max(e0, 0) -
// sum max(e, 0) over the slotted sums, one wave per image.
struct FlowOut {
  const int32_t* img;
  int stride, n;
  long long* out;
};
__device__ __forceinline__ void image_flows(const unsigned long long* __restrict__ sums,
                                            const FlowOut& F) {
  const int nw = blockDim.x / kWave;
  for (int i = threadIdx.x / kWave; i < F.n; i += nw) {
    const unsigned long long* r = sums + (size_t)F.img[(size_t)i * F.stride] * SUM_ROWS * kSumSlots;
    long long v = 0;
    for (int j = lane_id(); j < kSumSlots; j += kWave)
      v += (long long)r[SUM_MIN * kSumSlots + j] + (long long)r[SUM_POS0 * kSumSlots + j] -
           (long long)r[SUM_POSF * kSumSlots + j];
    v = wave_sum64(v);
    if (lane_id() == 0) F.out[i] = v;
  }
}

constexpr int kLabelsNT = 256;
template <int T>
constexpr int kLabelsTPB = kLabelsNT / (T * T / 16);

template <int T>
__global__ void __launch_bounds__(kLabelsNT)
k_labels(Geo g, CutDev s, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
         FlowOut F) {
  if (blockIdx.x == gridDim.x - 1) image_flows(s.sums, F);
  constexpr int TPR = T / 16;          // threads per tile row
  constexpr int TPT = T * T / 16;      // threads per tile
  constexpr int TPB = kLabelsTPB<T>;
  static_assert(kWave % TPR == 0, "a row's threads lie in one wave");
  const int slot = threadIdx.x / TPT, j = threadIdx.x % TPR, ly = (threadIdx.x % TPT) / TPR;
  const int n = *count;
  const int stride = gridDim.x * TPB;
  int pos = xcd_pos(blockIdx.x, gridDim.x) * TPB + slot;
  // the tile after next (its list entry), and the next tile (its geometry
  // and reach bytes)
  int t2 = pos + stride < n ? list_tile(list, pos + stride) : 0;
  TG qn{};
  uint4 vn = make_uint4(0, 0, 0, 0);
  if (pos < n) {
    const int t1 = list_tile(list, pos);
    qn = tile_geo(g, t1);
    vn = *reinterpret_cast<const uint4*>(s.reach + tbase<T>(t1) + ly * T + 16 * j);
  }
  // (the loop runs while any tile of the block's step is listed: every
  // thread takes part in the shuffles)
  for (int base = xcd_pos(blockIdx.x, gridDim.x) * TPB; base < n; base += stride, pos += stride) {
    const bool valid = pos < n;
    const TG q = qn;
    const uint4 v = vn;
    if (pos + stride < n) {
      const int t1 = t2;
      if (pos + 2 * stride < n) t2 = list_tile(list, pos + 2 * stride);
      qn = tile_geo(g, t1);
      vn = *reinterpret_cast<const uint4*>(s.reach + tbase<T>(t1) + ly * T + 16 * j);
    }
    const int x = q.tx * T + 16 * j, y = q.ty * T + ly;
    const int nin = (valid && y < q.H) ? min(max(q.W - x, 0), 16) : 0;
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
    unsigned bits = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) bits |= byte_bits(w[i], 0) << (4 * i);
    const unsigned m16 = nin >= 16 ? 0xffffu : ((1u << nin) - 1u);
    bits &= m16;
    if (nin > 0) {
      uint8_t* dst = s.labels + q.pix0 + (size_t)y * q.W + x;
      if (nin == 16 && (q.W & 15) == 0) {
        *reinterpret_cast<uint4*>(dst) = v;
      } else if ((q.W & 3) == 0) {
        // (pix0 is a multiple of 32, x of 16)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (4 * i + 4 <= nin) {
            reinterpret_cast<uint32_t*>(dst)[i] = w[i];
          } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
              if (4 * i + k < nin) dst[4 * i + k] = (uint8_t)(w[i] >> (8 * k));
          }
      } else {
#pragma unroll
        for (int k = 0; k < 16; ++k)
          if (k < nin) dst[k] = (uint8_t)(w[k / 4] >> (8 * (k % 4)));
      }
    }
    // the row's bits and mask: thread j's 16 at bit 16j (T = 64: two
    // 32-bit halves)
    const int sh = (16 * j) & 31;
    const bool hi = 16 * j >= 32;
    uint32_t blo = hi ? 0u : bits << sh, bhi = hi ? bits << sh : 0u;
    uint32_t mlo = hi ? 0u : m16 << sh, mhi = hi ? m16 << sh : 0u;
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) {
      blo |= __shfl_xor(blo, off, kWave);
      mlo |= __shfl_xor(mlo, off, kWave);
      if constexpr (T == 64) {
        bhi |= __shfl_xor(bhi, off, kWave);
        mhi |= __shfl_xor(mhi, off, kWave);
      }
    }
    const unsigned long long b = ((unsigned long long)bhi << 32) | blo;
    const unsigned long long m = ((unsigned long long)mhi << 32) | mlo;
    if (j == 0 && m) put_bits(s.packed, q.pix0 + (long long)y * q.W + q.tx * T, b, m);
  }
}


// Tile-blocked int32 field of one image -> row-major padded image.
template <int T>
__global__ void k_untile(Geo g, int img, const int32_t* __restrict__ src, int32_t* out) {
  const ImgDesc d = g.img[img];
  const int PW = d.tw * T;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= PW) return;
  const int t = d.tile0 + (y / T) * d.tw + x / T;
  out[(size_t)y * PW + x] = src[tbase<T>(t) + (y % T) * T + (x % T)];
}

}  // namespace gcmi

namespace gcmi {

// ------------------------------------------------------------- bookkeeping

// One selected image of a solve: its tiles go to sel[sel_off...] and, if
// fresh_off >= 0 (first solve of the image), also to fresh[fresh_off...].
struct Prep {
  int img, sel_off, fresh_off, pad;
};

// A pixel's edit in two steps, so that a thread can have the loads of
// several pixels in flight before it stores any: edit_load reads the
// pixel's unaries and state (in one round trip, whether or not the excess
// changes), edit_apply writes the new values.
struct PixEdit {
  size_t r, q;  // row-major unary index, tiled state index
  int t;        // tile
  int a0, b0, e0, m0;
};

template <int T>
__device__ __forceinline__ void edit_load(const CutDev& s, const ImgDesc& d, int x, int y,
                                          int warm, PixEdit& E) {
  E.r = d.pix0 + (size_t)y * d.W + x;
  E.t = d.tile0 + (y / T) * d.tw + x / T;
  E.q = tbase<T>(E.t) + (y % T) * T + (x % T);
  E.a0 = s.cs[E.r];
  E.b0 = s.ct[E.r];
  E.e0 = warm ? s.e[E.q] : 0;
  E.m0 = warm ? s.mask[E.q] : 0;
}

__device__ __forceinline__ void edit_apply(const CutDev& s, const PixEdit& E, int a1, int b1,
                                           int warm, long long& dmin, long long& dpos) {
  const int a0 = E.a0, b0 = E.b0;
  if (a1 == a0 && b1 == b0) return;  // (a pixel painted again with the same values)
  s.cs[E.r] = a1;
  s.ct[E.r] = b1;
  if (!warm) return;  // no state yet: the first round reads the unaries
  dmin += (long long)min(a1, b1) - min(a0, b0);
  dpos += (long long)(a1 - b1 > 0 ? a1 - b1 : 0) - (a0 - b0 > 0 ? a0 - b0 : 0);
  const int de = (a1 - b1) - (a0 - b0);
  if (de == 0) return;
  const int e1 = E.e0 + de;
  s.e[E.q] = e1;
  s.mask[E.q] = (uint8_t)((E.m0 & (M_UP | M_DOWN | M_LEFT | M_RIGHT)) | (e1 < 0 ? M_SINK : 0) |
                          (e1 > 0 ? M_EXCESS : 0));
  // a new deficit is a sink at height 0; the tile pushes before the first
  // relabel of the next solve
  if (e1 < 0) s.h[E.q] = 0;
  // the tile is active: the solve's first push round over the active tiles
  // recomputes its positive-excess sum tpos before anything reads it
  s.act[E.t] = 1;
  s.dirty[E.t] = 1;
}

// One pixel's share of a warm-start edit: new unaries (a1, b1) for pixel p
// (row-major index local to the image).  Updates the stored unaries and, if
// the image holds solved state (warm), e by the change of cs - ct, the
// mask's sink/excess bits and the tile's positive-excess sum; the changes
// of min(cs, ct) and max(cs - ct, 0) are added to dmin / dpos for the
// image's sums.  A pixel whose excess changed marks its tile active, and a
// new deficit gets height 0; other heights stay as they are.  The next
// solve pushes on the active tiles before its first global relabel; the
// labelling may be invalid until then, which costs pushes but not
// correctness, since termination is decided by an exact relabel.
template <int T>
__device__ __forceinline__ void edit_pixel(const CutDev& s, const ImgDesc& d, int x, int y, int a1,
                                           int b1, int warm, long long& dmin, long long& dpos) {
  PixEdit E;
  edit_load<T>(s, d, x, y, warm, E);
  edit_apply(s, E, a1, b1, warm, dmin, dpos);
}

// Adds a block's edit deltas to the image's sums (slotted atomics).
__device__ __forceinline__ void edit_sums(const CutDev& s, int img, long long dmin,
                                          long long dpos, long long* red) {
  dmin = block_sum64<256>(dmin, red);
  dpos = block_sum64<256>(dpos, red);
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x % kSumSlots;
    if (dmin)
      atomicAdd(&s.sums[((size_t)img * SUM_ROWS + SUM_MIN) * kSumSlots + slot],
                (unsigned long long)dmin);
    if (dpos)
      atomicAdd(&s.sums[((size_t)img * SUM_ROWS + SUM_POS0) * kSumSlots + slot],
                (unsigned long long)dpos);
  }
}

// Warm-start edit of n distinct pixels of image img: idx are row-major
// indices local to the image, (ncs, nct) their new unaries.
template <int T>
__global__ void __launch_bounds__(256)
k_edit(Geo g, CutDev s, int img, int n, const int32_t* __restrict__ idx,
       const int32_t* __restrict__ ncs, const int32_t* __restrict__ nct, int warm) {
  __shared__ long long red[256 / kWave];
  const ImgDesc d = g.img[img];
  long long dmin = 0, dpos = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int p = idx[i];
    edit_pixel<T>(s, d, p % d.W, p / d.W, ncs[i], nct[i], warm, dmin, dpos);
  }
  edit_sums(s, img, dmin, dpos, red);
}

// A brush stroke: the pixels whose centre (x, y) lies within width/2 of a
// polyline -- a union of capsules, one per segment -- get unaries (cs, ct).
constexpr int kMaxStrokePoints = 16;
constexpr int kPaintCols = 64;  // columns of one row of one segment painted by one wave
constexpr int kPaintRows = 4;    // rows of one segment painted by one wave
struct Polyline {
  double x[kMaxStrokePoints], y[kMaxStrokePoints];
  double r;  // width / 2
  int npts;
  int32_t cs, ct;
  int bx0, by0, bw, bh;  // pixel box covering the stroke, clipped to the image
  int W;                 // image width
  // rows of the image each segment's capsule can reach: segment k covers
  // nr[k] rows from ya[k], in groups of kPaintRows rows, each row in xc[k]
  // pieces of kPaintCols columns; segment k's (group, piece) chunks start at
  // chunk c0[k] of the stroke, group by group; nchunks chunks in all
  int ya[kMaxStrokePoints], nr[kMaxStrokePoints], xc[kMaxStrokePoints], c0[kMaxStrokePoints];
  int nchunks;
  // pixel box (x0, x1, y0, y1) of each segment's capsule: a pixel outside it
  // is not in it (one 16-byte scalar load per segment)
  int4 sbox[kMaxStrokePoints];
};

// Whether pixel centre (px, py) lies within r of segment k: the point of the
// segment nearest to it is a + t (b - a), t = clamp(((p - a).(b - a)) /
// |b - a|^2, 0, 1) (t = 0 for a zero-length segment), and the test is
// dx^2 + dy^2 <= r^2.  Products and sums are rounded one at a time, in this
// order, so that the pixel set is reproducible on any IEEE float64 host.
__device__ __forceinline__ bool in_capsule(const Polyline& pl, int k, double px, double py) {
#pragma clang fp contract(off)
  const double x0 = pl.x[k], y0 = pl.y[k];
  const double dx = pl.x[k + 1] - x0, dy = pl.y[k + 1] - y0;
  const double L2 = dx * dx + dy * dy;
  double t = 0.0;
  if (L2 > 0.0) {
    const double u = (px - x0) * dx;
    const double v = (py - y0) * dy;
    t = (u + v) / L2;
    t = fmin(fmax(t, 0.0), 1.0);
  }
  const double qx = x0 + t * dx, qy = y0 + t * dy;
  const double ex = px - qx, ey = py - qy;
  const double d2 = ex * ex + ey * ey;
  return d2 <= pl.r * pl.r;
}

// Columns of row y that segment k's capsule can reach: the capsule's points
// in the row lie within r of segment points whose y is within r of the row,
// so [xl, xr] is the x range of the segment's part with |y(t) - y| <= r + 1,
// widened by r + 1 (the extra pixel absorbs rounding).  False: none.
__device__ __forceinline__ bool seg_row_span(const Polyline& pl, int k, double y, double* xl,
                                             double* xr) {
  const double r1 = pl.r + 1.0;
  const double x0 = pl.x[k], y0 = pl.y[k], x1 = pl.x[k + 1], y1 = pl.y[k + 1];
  const double dy = y1 - y0;
  double ta = 0.0, tb = 1.0;
  if (dy != 0.0) {
    double t0 = (y - r1 - y0) / dy, t1 = (y + r1 - y0) / dy;
    if (t0 > t1) {
      const double tmp = t0;
      t0 = t1;
      t1 = tmp;
    }
    ta = fmax(t0, 0.0);
    tb = fmin(t1, 1.0);
    if (ta > tb) return false;
  } else if (fabs(y - y0) > r1) {
    return false;
  }
  const double xa = x0 + ta * (x1 - x0), xb = x0 + tb * (x1 - x0);
  *xl = fmin(xa, xb) - r1;
  *xr = fmax(xa, xb) + r1;
  return true;
}

// A stroke of a paint launch: its polyline, image, whether the image has
// solved state to reflow (warm), and the stroke's first chunk among all the
// launch's (stroke, segment, chunk of rows) items.
struct PaintJob {
  Polyline pl;
  int img;
  int warm;
  int chunk0;
};

constexpr int kPaintJobsLds = 256;  // jobs whose first chunks a block stages in LDS

// Paints the strokes of `jobs` (chunk0 ascending): one wave per (stroke,
// segment, kPaintRows rows, piece of kPaintCols columns) a segment's capsule
// can reach, its lanes over the piece's candidate columns of each row in
// turn.  A pixel is painted by the first segment of its stroke whose capsule
// holds it, so every pixel of a stroke's union is edited once (strokes of
// one launch are of different images).
template <int T>
__global__ void __launch_bounds__(256)
k_paint(Geo g, CutDev s, const PaintJob* __restrict__ jobs, int njobs, int nchunks) {
  __shared__ int sc0[kPaintJobsLds];
  const bool lds = njobs <= kPaintJobsLds;
  if (lds)
    for (int j = threadIdx.x; j < njobs; j += blockDim.x) sc0[j] = jobs[j].chunk0;
  __syncthreads();
  // item, job and segment are wave-uniform: the stroke's fields are read
  // through the scalar cache
  const int lane = threadIdx.x % kWave, waves = blockDim.x / kWave;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  for (int item = blockIdx.x * waves + wave; item < nchunks; item += gridDim.x * waves) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((lds ? sc0[mid] : jobs[mid].chunk0) <= item) lo = mid;
      else hi = mid - 1;
    }
    const PaintJob& J = jobs[__builtin_amdgcn_readfirstlane(lo)];
    const Polyline& pl = J.pl;
    const int local = item - J.chunk0;
    // the segment: the last one whose first chunk is <= local (one load per
    // lane, then a ballot)
    const int nseg = pl.npts - 1;
    const unsigned long long le = __ballot(lane < nseg && pl.c0[lane] <= local);
    const int k = 63 - __clzll(le);
    const int piece = local - pl.c0[k];
    const int ya = pl.ya[k] + (piece / pl.xc[k]) * kPaintRows;
    const int yb = min(ya + kPaintRows, pl.ya[k] + pl.nr[k]);
    const int xpiece = piece % pl.xc[k];
    const ImgDesc d = g.img[J.img];
    long long dmin = 0, dpos = 0;
    // each row's candidate columns [rx0, rx1] in this piece (none: rx1 < rx0)
    int rx0[kPaintRows], rx1[kPaintRows], passes = 0;
#pragma unroll
    for (int ri = 0; ri < kPaintRows; ++ri) {
      const int y = ya + ri;
      int x0 = 0, x1 = -1;
      double xl, xr;
      if (y < yb && seg_row_span(pl, k, (double)y, &xl, &xr)) {
        xl = fmax(xl, 0.0);
        xr = fmin(xr, (double)(pl.W - 1));
        if (xl <= xr) {
          x0 = (int)floor(xl) + xpiece * kPaintCols;
          x1 = min((int)ceil(xr), x0 + kPaintCols - 1);
        }
      }
      rx0[ri] = __builtin_amdgcn_readfirstlane(x0);
      rx1[ri] = __builtin_amdgcn_readfirstlane(x1);
      passes = max(passes, (rx1[ri] - rx0[ri] + kWave) / kWave);
    }
    // a pass takes one column of every row: the loads of all its pixels are
    // in flight before the first store
    for (int p = 0; p < passes; ++p) {
      PixEdit E[kPaintRows];
      bool hit[kPaintRows];
#pragma unroll
      for (int ri = 0; ri < kPaintRows; ++ri) {
        const int x = rx0[ri] + p * kWave + lane, y = ya + ri;
        hit[ri] = x <= rx1[ri] && in_capsule(pl, k, (double)x, (double)y);
        for (int j = 0; j < k && hit[ri]; ++j) {
          const int4 b = pl.sbox[j];
          hit[ri] = !(x >= b.x && x <= b.y && y >= b.z && y <= b.w &&
                      in_capsule(pl, j, (double)x, (double)y));
        }
        if (hit[ri]) edit_load<T>(s, d, x, y, J.warm, E[ri]);
      }
#pragma unroll
      for (int ri = 0; ri < kPaintRows; ++ri)
        if (hit[ri]) edit_apply(s, E[ri], pl.cs, pl.ct, J.warm, dmin, dpos);
    }
    if (J.warm) {
      // the chunk's changes of the image's sums (one slotted atomic per wave)
      dmin = wave_sum64(dmin);
      dpos = wave_sum64(dpos);
      if (lane == 0) {
        const size_t slot = (size_t)J.img * SUM_ROWS * kSumSlots + blockIdx.x % kSumSlots;
        if (dmin) atomicAdd(&s.sums[slot + SUM_MIN * kSumSlots], (unsigned long long)dmin);
        if (dpos) atomicAdd(&s.sums[slot + SUM_POS0 * kSumSlots], (unsigned long long)dpos);
      }
    }
  }
}

}  // namespace gcmi

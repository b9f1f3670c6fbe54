// Device kernels of the tiled push-relabel grid cut (gfx950).
//
// HBM layout: the grid is padded to whole T x T tiles; every per-pixel array
// is tile-blocked -- tile t = ty*tw + tx occupies [t*T*T, (t+1)*T*T), row-major
// inside the tile -- so a tile's state is T*T contiguous int32 per field and
// every load/store of a tile is a run of full 128-byte lines.
//
// Per-pixel state (int32): e (excess; < 0 means remaining sink capacity),
// h (height / BFS distance), cu, cd, cl, cr (residual capacity of the arc to
// the up / down / left / right neighbour).  Flow pushed across a tile edge
// is not written into the neighbour tile (which may be resident in another
// work-group at the same time) but added atomically to a pending slot
// pend[(t*4 + side)*T + k] of the receiving tile, which folds it into e and
// the reverse capacity the next time it is loaded.
//
// Reference counterparts: push_up/down/left/right + add_border
// (graphcut.cl:114-536), relabel (:538-577), init_bfs / bfs_intratile /
// bfs_intertile (:582-1016), check_completion (:20-60).
#pragma once

#include "gcmi_common.h"

namespace gcmi {

enum { D_UP = 0, D_DOWN = 1, D_LEFT = 2, D_RIGHT = 3 };
enum { M_UP = 1, M_DOWN = 2, M_LEFT = 4, M_RIGHT = 8, M_SINK = 16, M_EXCESS = 32 };

// 64-bit sums are spread over kSumSlots addresses (by block) so that
// thousands of blocks do not serialise on one atomic; the host adds the
// slots.  Rows: sum of min(cs, ct); sum of max(e0, 0); sum of max(e, 0).
constexpr int kSumSlots = 64;
// Rows 3-4 tally tile passes of the BFS and flood rounds (added by block 0
// of each round launch) so the host learns them without a sync per round.
enum {
  SUM_MIN = 0,
  SUM_POS0 = 1,
  SUM_POSF = 2,
  TALLY_BFS = 3,
  TALLY_FLOOD = 4,
  SUM_ROWS_ALL = 5
};
constexpr size_t kSumBytes = (size_t)SUM_ROWS_ALL * kSumSlots * sizeof(unsigned long long);

struct Geo {
  int W, H;      // image
  int tw, th;    // tiles across / down
  int ntiles;
  int32_t maxh;  // MAX_HEIGHT: padded pixel count
};

template <int T>
__device__ __forceinline__ size_t tbase(int t) {
  return (size_t)t * (T * T);
}

// Block-wide sum of a 64-bit value; result valid in thread 0.  `red` is LDS
// scratch of NT/64 long longs.
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

// Device pointers of one cut's state, passed by value to the kernels.
struct CutDev {
  int32_t *e, *h, *cu, *cd, *cl, *cr;  // tile-blocked state
  uint8_t* mask;                       // residual bits + sink/excess bits
  int32_t* pend;                       // pending edge inflow, 4*T per tile
  int32_t* pflag;                      // tile has pending inflow
  int32_t* act;                        // tile flagged for the next push round
  long long* tpos;                     // per tile: sum of max(e, 0)
  unsigned long long* sums;            // slotted 64-bit sums / tallies
  const int32_t *cs, *ct, *right, *down;  // row-major inputs
};

__device__ __forceinline__ int pixel_mask(int e, int cu, int cd, int cl, int cr) {
  return (cu > 0 ? M_UP : 0) | (cd > 0 ? M_DOWN : 0) | (cl > 0 ? M_LEFT : 0) |
         (cr > 0 ? M_RIGHT : 0) | (e < 0 ? M_SINK : 0) | (e > 0 ? M_EXCESS : 0);
}

// ----------------------------------------------------------------- push
// One residency of each listed tile: load state (FIRST: build it from the
// row-major capacities), fold pending inflow, freeze the halo heights, run
// up to `sweeps` synchronous push/relabel sweeps in LDS, write back state,
// residual mask and the tile's positive-excess sum, and post edge outflow to
// the neighbours' pending slots.
//
// FIRST (the first round, over every tile): e = cs - ct, caps from the
// pairwise weights, h = 0 on sink pixels (e < 0) and 1 elsewhere -- a valid
// labelling, since every residual arc then satisfies h(p) <= h(q) + 1 --
// and sums[SUM_MIN] / sums[SUM_POS0] collect sum min(cs, ct), sum max(e0, 0).
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
template <int T, int NT, bool FIRST>
__global__ void __launch_bounds__(NT)
k_push(Geo g, CutDev s, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
       int sweeps) {
  constexpr int R = T * T / NT;
  constexpr int HP = T + 2;
  static_assert(4 * T <= NT, "halo loaders");
  __shared__ int32_t H0[HP * HP], H1[HP * HP];
  __shared__ int32_t F[4][T * T];
  __shared__ int32_t BO[4][T];
  __shared__ long long red[NT / kWave];
  __shared__ int sflag;
  const int lx = threadIdx.x % T, chain = threadIdx.x / T;
  const int32_t MAXH = g.maxh;
  const int n = FIRST ? g.ntiles : *count;
  long long smin = 0, spos = 0;
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const int t = FIRST ? li : list[li];
    const int tx = t % g.tw, ty = t / g.tw;
    const size_t base = tbase<T>(t);
    int e[R], h[R], cu[R], cd[R], cl[R], cr[R];
    if (FIRST) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int x = tx * T + lx, y = ty * T + chain * R + k;
        e[k] = 0; cu[k] = 0; cd[k] = 0; cl[k] = 0; cr[k] = 0;
        if (x < g.W && y < g.H) {
          const size_t p = (size_t)y * g.W + x;
          const int a = s.cs[p], b = s.ct[p];
          cr[k] = x < g.W - 1 ? s.right[p] : 0;
          cl[k] = x > 0 ? s.right[p - 1] : 0;
          cd[k] = y < g.H - 1 ? s.down[p] : 0;
          cu[k] = y > 0 ? s.down[p - g.W] : 0;
          e[k] = a - b;
          smin += min(a, b);
          spos += e[k] > 0 ? e[k] : 0;
        }
        h[k] = e[k] < 0 ? 0 : 1;
      }
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const size_t i = base + (chain * R + k) * T + lx;
        e[k] = s.e[i];
        h[k] = s.h[i];
        cu[k] = s.cu[i];
        cd[k] = s.cd[i];
        cl[k] = s.cl[i];
        cr[k] = s.cr[i];
      }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int c = (chain * R + k + 1) * HP + lx + 1;
      H0[c] = h[k];
      H1[c] = h[k];
    }
    if (threadIdx.x < 4 * T) {
      const int side = threadIdx.x / T, j = threadIdx.x % T;
      int nb = -1, nx = 0, ny = 0, idx;
      if (side == 0) { if (ty > 0) { nb = t - g.tw; nx = j; ny = T - 1; } idx = j + 1; }
      else if (side == 1) { if (ty < g.th - 1) { nb = t + g.tw; nx = j; ny = 0; } idx = (T + 1) * HP + j + 1; }
      else if (side == 2) { if (tx > 0) { nb = t - 1; nx = T - 1; ny = j; } idx = (j + 1) * HP; }
      else { if (tx < g.tw - 1) { nb = t + 1; nx = 0; ny = j; } idx = (j + 1) * HP + T + 1; }
      int v = MAXH;
      if (nb >= 0) {
        if (FIRST) {
          const int x = (nb % g.tw) * T + nx, y = (nb / g.tw) * T + ny;
          if (x < g.W && y < g.H) {
            const size_t p = (size_t)y * g.W + x;
            v = s.cs[p] - s.ct[p] < 0 ? 0 : 1;
          }
        } else {
          v = s.h[tbase<T>(nb) + ny * T + nx];
        }
      }
      H0[idx] = v;
      H1[idx] = v;
      BO[side][j] = 0;
    }
    if (threadIdx.x == 0) sflag = s.pflag[t];
    __syncthreads();
    if (sflag) {
      // issue every exchange first, then use the results (one wait)
      int vu[R], vd[R], vl[R], vr[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int ly = chain * R + k;
        vu[k] = ly == 0 ? atomicExch(&s.pend[((size_t)t * 4 + 0) * T + lx], 0) : 0;
        vd[k] = ly == T - 1 ? atomicExch(&s.pend[((size_t)t * 4 + 1) * T + lx], 0) : 0;
        vl[k] = lx == 0 ? atomicExch(&s.pend[((size_t)t * 4 + 2) * T + ly], 0) : 0;
        vr[k] = lx == T - 1 ? atomicExch(&s.pend[((size_t)t * 4 + 3) * T + ly], 0) : 0;
      }
#pragma unroll
      for (int k = 0; k < R; ++k) {
        e[k] += vu[k] + vd[k] + vl[k] + vr[k];
        cu[k] += vu[k];
        cd[k] += vd[k];
        cl[k] += vl[k];
        cr[k] += vr[k];
      }
    }
    int any0 = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) any0 |= e[k] > 0 && h[k] < MAXH;
    int active = __syncthreads_or(any0);

    for (int sw = 0; sw < sweeps && active; ++sw) {
      const int32_t* Hc = (sw & 1) ? H1 : H0;
      int32_t* Hn = (sw & 1) ? H0 : H1;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int ly = chain * R + k;
        const int i = ly * T + lx;
        const int c = (ly + 1) * HP + lx + 1;
        int fu = 0, fd = 0, fl = 0, fr = 0;
        const int hv = h[k];
        if (e[k] > 0 && hv < MAXH) {
          int rem = e[k];
          if (cd[k] > 0 && hv == Hc[c + HP] + 1) {
            fd = min(rem, cd[k]);
            rem -= fd;
            cd[k] -= fd;
          }
          if (rem > 0 && cu[k] > 0 && hv == Hc[c - HP] + 1) {
            fu = min(rem, cu[k]);
            rem -= fu;
            cu[k] -= fu;
          }
          if (rem > 0 && cr[k] > 0 && hv == Hc[c + 1] + 1) {
            fr = min(rem, cr[k]);
            rem -= fr;
            cr[k] -= fr;
          }
          if (rem > 0 && cl[k] > 0 && hv == Hc[c - 1] + 1) {
            fl = min(rem, cl[k]);
            rem -= fl;
            cl[k] -= fl;
          }
          e[k] = rem;
          if (ly == 0) BO[0][lx] += fu;
          if (ly == T - 1) BO[1][lx] += fd;
          if (lx == 0) BO[2][ly] += fl;
          if (lx == T - 1) BO[3][ly] += fr;
        }
        F[D_UP][i] = fu;
        F[D_DOWN][i] = fd;
        F[D_LEFT][i] = fl;
        F[D_RIGHT][i] = fr;
      }
      __syncthreads();
      int any = 0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int ly = chain * R + k;
        const int i = ly * T + lx;
        const int c = (ly + 1) * HP + lx + 1;
        if (ly > 0) {
          const int v = F[D_DOWN][i - T];
          e[k] += v;
          cu[k] += v;
        }
        if (ly < T - 1) {
          const int v = F[D_UP][i + T];
          e[k] += v;
          cd[k] += v;
        }
        if (lx > 0) {
          const int v = F[D_RIGHT][i - 1];
          e[k] += v;
          cl[k] += v;
        }
        if (lx < T - 1) {
          const int v = F[D_LEFT][i + 1];
          e[k] += v;
          cr[k] += v;
        }
        const int hv = h[k];
        if (e[k] > 0 && hv < MAXH) {
          const int hd = Hc[c + HP], hu = Hc[c - HP], hr = Hc[c + 1], hl = Hc[c - 1];
          const bool adm = (cd[k] > 0 && hv == hd + 1) || (cu[k] > 0 && hv == hu + 1) ||
                           (cr[k] > 0 && hv == hr + 1) || (cl[k] > 0 && hv == hl + 1);
          if (!adm) {
            int nh = MAXH;
            if (cd[k] > 0) nh = min(nh, hd + 1);
            if (cu[k] > 0) nh = min(nh, hu + 1);
            if (cr[k] > 0) nh = min(nh, hr + 1);
            if (cl[k] > 0) nh = min(nh, hl + 1);
            h[k] = nh;
          }
          any |= h[k] < MAXH;
        }
        Hn[c] = h[k];
      }
      active = __syncthreads_or(any);
    }

    long long tp = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const size_t i = base + (chain * R + k) * T + lx;
      s.e[i] = e[k];
      s.h[i] = h[k];
      s.cu[i] = cu[k];
      s.cd[i] = cd[k];
      s.cl[i] = cl[k];
      s.cr[i] = cr[k];
      s.mask[i] = (uint8_t)pixel_mask(e[k], cu[k], cd[k], cl[k], cr[k]);
      tp += e[k] > 0 ? e[k] : 0;
    }
    // edge outflow -> neighbour's pending slot (the receiving side is the
    // opposite one); a nonzero outflow implies the arc and so the neighbour
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      if (ly == 0 && ty > 0 && BO[0][lx]) {
        atomicAdd(&s.pend[((size_t)(t - g.tw) * 4 + 1) * T + lx], BO[0][lx]);
        s.pflag[t - g.tw] = 1;
        s.act[t - g.tw] = 1;
      }
      if (ly == T - 1 && ty < g.th - 1 && BO[1][lx]) {
        atomicAdd(&s.pend[((size_t)(t + g.tw) * 4 + 0) * T + lx], BO[1][lx]);
        s.pflag[t + g.tw] = 1;
        s.act[t + g.tw] = 1;
      }
      if (lx == 0 && tx > 0 && BO[2][ly]) {
        atomicAdd(&s.pend[((size_t)(t - 1) * 4 + 3) * T + ly], BO[2][ly]);
        s.pflag[t - 1] = 1;
        s.act[t - 1] = 1;
      }
      if (lx == T - 1 && tx < g.tw - 1 && BO[3][ly]) {
        atomicAdd(&s.pend[((size_t)(t + 1) * 4 + 2) * T + ly], BO[3][ly]);
        s.pflag[t + 1] = 1;
        s.act[t + 1] = 1;
      }
    }
    tp = block_sum64<NT>(tp, red);
    if (threadIdx.x == 0) {
      s.tpos[t] = tp;
      if (active) s.act[t] = 1;
    }
  }
  if (FIRST) {
    smin = block_sum64<NT>(smin, red);
    spos = block_sum64<NT>(spos, red);
    if (threadIdx.x == 0) {
      const int slot = blockIdx.x % kSumSlots;
      if (smin) atomicAdd(&s.sums[SUM_MIN * kSumSlots + slot], (unsigned long long)smin);
      if (spos) atomicAdd(&s.sums[SUM_POS0 * kSumSlots + slot], (unsigned long long)spos);
    }
  }
}

// Folds the pending inflow of every flagged tile into its state (between
// launches, when no tile is resident): e, the reverse capacity, the mask
// byte and the tile's positive-excess sum.  One block of kFoldNT threads per
// tile (unflagged tiles exit at once); one owner thread per edge pixel
// (threads < 4T - 4 walk the perimeter), so corners are not raced.
constexpr int kFoldNT = 256;

template <int T>
__global__ void __launch_bounds__(kFoldNT) k_fold(Geo g, CutDev s) {
  static_assert(4 * T - 4 <= kFoldNT, "perimeter owners");
  const int t = blockIdx.x;
  if (t >= g.ntiles || !s.pflag[t]) return;
  const int j = threadIdx.x;
  if (j == 0) s.pflag[t] = 0;
  if (j >= 4 * T - 4) return;
  int x, y;
  if (j < T) { x = j; y = 0; }
  else if (j < 2 * T) { x = j - T; y = T - 1; }
  else if (j < 3 * T - 2) { x = 0; y = j - 2 * T + 1; }
  else { x = T - 1; y = j - (3 * T - 2) + 1; }
  const size_t i = tbase<T>(t) + y * T + x;
  int32_t* pe = s.pend + (size_t)t * 4 * T;
  const int vu = y == 0 ? pe[0 * T + x] : 0;
  const int vd = y == T - 1 ? pe[1 * T + x] : 0;
  const int vl = x == 0 ? pe[2 * T + y] : 0;
  const int vr = x == T - 1 ? pe[3 * T + y] : 0;
  if ((vu | vd | vl | vr) == 0) return;
  const int e0 = s.e[i];
  const int e1 = e0 + vu + vd + vl + vr;
  const int cu = s.cu[i] + vu, cd = s.cd[i] + vd, cl = s.cl[i] + vl, cr = s.cr[i] + vr;
  s.e[i] = e1;
  s.cu[i] = cu;
  s.cd[i] = cd;
  s.cl[i] = cl;
  s.cr[i] = cr;
  s.mask[i] = (uint8_t)pixel_mask(e1, cu, cd, cl, cr);
  if (vu) pe[0 * T + x] = 0;
  if (vd) pe[1 * T + x] = 0;
  if (vl) pe[2 * T + y] = 0;
  if (vr) pe[3 * T + y] = 0;
  const long long delta = (long long)(e1 > 0 ? e1 : 0) - (e0 > 0 ? e0 : 0);
  if (delta) atomicAdd(reinterpret_cast<unsigned long long*>(&s.tpos[t]), (unsigned long long)delta);
}

// ------------------------------------------------------ global relabel
// Distance of every pixel to the sink set {e < 0} over residual arcs
// (MAX if unreachable): what init_bfs / bfs_intratile / bfs_intertile
// (graphcut.cl:582-1016) compute, without their 30-round cap.  Inside a
// tile the Bellman-Ford relaxation runs in LDS to its fixed point; across
// tiles, a tile whose edge distances fell flags its neighbours for another
// round, until no tile is flagged.

// Relaxes d[] (this thread's R pixels) to the in-tile fixed point given
// halo values already in D0/D1.  Returns after the last barrier.
template <int T, int NT>
__device__ __forceinline__ void bfs_relax(int (&d)[T * T / NT], const int (&m)[T * T / NT],
                                          int32_t* D0, int32_t* D1) {
  constexpr int R = T * T / NT;
  constexpr int HP = T + 2;
  const int lx = threadIdx.x % T, chain = threadIdx.x / T;
  int32_t* Dc = D0;
  int32_t* Dn = D1;
  for (;;) {
    int changed = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      const int c = (ly + 1) * HP + lx + 1;
      int nd = d[k];
      if (nd > 0) {
        if (m[k] & M_UP) nd = min(nd, Dc[c - HP] + 1);
        if (m[k] & M_DOWN) nd = min(nd, Dc[c + HP] + 1);
        if (m[k] & M_LEFT) nd = min(nd, Dc[c - 1] + 1);
        if (m[k] & M_RIGHT) nd = min(nd, Dc[c + 1] + 1);
        if (nd < d[k]) {
          d[k] = nd;
          changed = 1;
        }
      }
      Dn[c] = d[k];
    }
    if (!__syncthreads_or(changed)) break;
    int32_t* tmp = Dc;
    Dc = Dn;
    Dn = tmp;
  }
}

// Writes halo distances of tile t into D0/D1 (threads < 4T).  `init`: a
// neighbour pixel counts as 0 if it is a sink (mask bit), MAX otherwise;
// else its current distance from h_g.
template <int T>
__device__ __forceinline__ void load_halo_dist(const Geo& g, int t, const uint8_t* mask_g,
                                               const int32_t* h_g, bool init, int32_t* D0,
                                               int32_t* D1) {
  constexpr int HP = T + 2;
  if (threadIdx.x >= 4 * T) return;
  const int tx = t % g.tw, ty = t / g.tw;
  const int side = threadIdx.x / T, j = threadIdx.x % T;
  int nb = -1, nx = 0, ny = 0;  // neighbour tile and the pixel inside it
  int idx;
  if (side == 0) { if (ty > 0) { nb = t - g.tw; nx = j; ny = T - 1; } idx = j + 1; }
  else if (side == 1) { if (ty < g.th - 1) { nb = t + g.tw; nx = j; ny = 0; } idx = (T + 1) * HP + j + 1; }
  else if (side == 2) { if (tx > 0) { nb = t - 1; nx = T - 1; ny = j; } idx = (j + 1) * HP; }
  else { if (tx < g.tw - 1) { nb = t + 1; nx = 0; ny = j; } idx = (j + 1) * HP + T + 1; }
  int v = g.maxh;
  if (nb >= 0) {
    const size_t q = tbase<T>(nb) + ny * T + nx;
    v = init ? ((mask_g[q] & M_SINK) ? 0 : g.maxh) : h_g[q];
  }
  D0[idx] = v;
  D1[idx] = v;
}

// Flags the neighbour tiles across the edges this thread's pixel k lies on.
template <int T>
__device__ __forceinline__ void flag_edge_neighbours(const Geo& g, int t, int lx, int ly,
                                                     int32_t* flags) {
  const int tx = t % g.tw, ty = t / g.tw;
  if (ly == 0 && ty > 0) flags[t - g.tw] = 1;
  if (ly == T - 1 && ty < g.th - 1) flags[t + g.tw] = 1;
  if (lx == 0 && tx > 0) flags[t - 1] = 1;
  if (lx == T - 1 && tx < g.tw - 1) flags[t + 1] = 1;
}

// First pass over all tiles, from the residual masks (pending inflow has
// been folded): seeds, in-tile distances, and the completion test
// (check_completion, graphcut.cl:20-60) for the tile:
//   act[t]    = some pixel has e > 0 and a finite distance
//   haspos[t] = some pixel has e > 0
// Later rounds can only lower distances, so they can only switch act on.
template <int T, int NT>
__global__ void __launch_bounds__(NT)
k_bfs_init(Geo g, const uint8_t* __restrict__ mask_g, int32_t* h_g, int32_t* bflag,
           int32_t* haspos, int32_t* act) {
  constexpr int R = T * T / NT;
  constexpr int HP = T + 2;
  __shared__ int32_t D0[HP * HP], D1[HP * HP];
  const int lx = threadIdx.x % T, chain = threadIdx.x / T;
  for (int t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
    const size_t base = tbase<T>(t);
    int d[R], m[R];
    int any_pos = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) m[k] = mask_g[base + (chain * R + k) * T + lx];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      d[k] = (m[k] & M_SINK) ? 0 : g.maxh;
      D0[(ly + 1) * HP + lx + 1] = d[k];
      any_pos |= (m[k] & M_EXCESS) != 0;
    }
    load_halo_dist<T>(g, t, mask_g, h_g, true, D0, D1);
    __syncthreads();
    bfs_relax<T, NT>(d, m, D0, D1);
    int any_act = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      h_g[base + ly * T + lx] = d[k];
      any_act |= (m[k] & M_EXCESS) && d[k] < g.maxh;
      // neighbours assumed MAX for this non-sink pixel; a finite distance is news
      if (d[k] < g.maxh && !(m[k] & M_SINK)) flag_edge_neighbours<T>(g, t, lx, ly, bflag);
    }
    any_pos = __syncthreads_or(any_pos);
    any_act = __syncthreads_or(any_act);
    if (threadIdx.x == 0) {
      haspos[t] = any_pos;
      act[t] = any_act;
    }
  }
}

// Later passes over flagged tiles, from the stored distances.
template <int T, int NT>
__global__ void __launch_bounds__(NT)
k_bfs_round(Geo g, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
            int32_t* h_g, const uint8_t* __restrict__ mask_g, int32_t* bflag, int32_t* act,
            unsigned long long* sums) {
  constexpr int R = T * T / NT;
  constexpr int HP = T + 2;
  __shared__ int32_t D0[HP * HP], D1[HP * HP];
  const int lx = threadIdx.x % T, chain = threadIdx.x / T;
  const int n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0)
    atomicAdd(&sums[TALLY_BFS * kSumSlots], (unsigned long long)n);
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const int t = list[li];
    const size_t base = tbase<T>(t);
    int d[R], d0[R], m[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      const size_t i = base + ly * T + lx;
      d[k] = d0[k] = h_g[i];
      m[k] = mask_g[i];
      D0[(ly + 1) * HP + lx + 1] = d[k];
    }
    load_halo_dist<T>(g, t, nullptr, h_g, false, D0, D1);
    __syncthreads();
    bfs_relax<T, NT>(d, m, D0, D1);
    int any_act = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      if (d[k] < d0[k]) {
        h_g[base + ly * T + lx] = d[k];
        flag_edge_neighbours<T>(g, t, lx, ly, bflag);
        any_act |= (m[k] & M_EXCESS) != 0;
      }
    }
    any_act = __syncthreads_or(any_act);
    if (threadIdx.x == 0 && any_act) act[t] = 1;
  }
}

// ------------------------------------------------------ minimal source set
// reach = pixels reachable over residual arcs from {e > 0}: the minimal
// source set once no pixel with e > 0 can reach a sink.  Same tile scheme
// as the BFS: in-tile fixed point in LDS, flags across edges.

// Relaxes r[] to the in-tile fixed point.  RM holds (reach | mask << 1) of
// every tile pixel and halo pixel (halo: the neighbour's reach and mask).
template <int T, int NT>
__device__ __forceinline__ void flood_relax(int (&r)[T * T / NT], int32_t* RM) {
  constexpr int R = T * T / NT;
  constexpr int HP = T + 2;
  const int lx = threadIdx.x % T, chain = threadIdx.x / T;
  for (;;) {
    int changed = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (r[k]) continue;
      const int ly = chain * R + k;
      const int c = (ly + 1) * HP + lx + 1;
      // a neighbour q reaches p if q is reached and q's arc toward p is residual
      const int up = RM[c - HP], dn = RM[c + HP], lf = RM[c - 1], rt = RM[c + 1];
      if (((up & 1) && (up & (M_DOWN << 1))) || ((dn & 1) && (dn & (M_UP << 1))) ||
          ((lf & 1) && (lf & (M_RIGHT << 1))) || ((rt & 1) && (rt & (M_LEFT << 1)))) {
        r[k] = 1;
        changed = 1;
      }
    }
    // publish new reach bits (only ever set, so in-place updates are safe
    // across the barrier: a reader sees 0 or 1, both consistent)
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      const int c = (ly + 1) * HP + lx + 1;
      if (r[k]) RM[c] |= 1;
    }
    if (!__syncthreads_or(changed)) break;
  }
}

template <int T>
__device__ __forceinline__ void load_halo_flood(const Geo& g, int t, const uint8_t* reach_g,
                                                const uint8_t* mask_g, bool with_reach,
                                                int32_t* RM) {
  constexpr int HP = T + 2;
  if (threadIdx.x >= 4 * T) return;
  const int tx = t % g.tw, ty = t / g.tw;
  const int side = threadIdx.x / T, j = threadIdx.x % T;
  int nb = -1;
  size_t off = 0;
  int idx;
  if (side == 0) { if (ty > 0) { nb = t - g.tw; off = (T - 1) * T + j; } idx = j + 1; }
  else if (side == 1) { if (ty < g.th - 1) { nb = t + g.tw; off = j; } idx = (T + 1) * HP + j + 1; }
  else if (side == 2) { if (tx > 0) { nb = t - 1; off = j * T + T - 1; } idx = (j + 1) * HP; }
  else { if (tx < g.tw - 1) { nb = t + 1; off = j * T; } idx = (j + 1) * HP + T + 1; }
  int v = 0;
  if (nb >= 0) {
    const size_t q = tbase<T>(nb) + off;
    v = (int)mask_g[q] << 1;
    if (with_reach) v |= reach_g[q] ? 1 : 0;
  }
  RM[idx] = v;
}

// Does pixel (lx,ly) with mask m have a residual arc out of the tile?
template <int T>
__device__ __forceinline__ void flag_out_arcs(const Geo& g, int t, int lx, int ly, int m,
                                              int32_t* flags) {
  const int tx = t % g.tw, ty = t / g.tw;
  if (ly == 0 && ty > 0 && (m & M_UP)) flags[t - g.tw] = 1;
  if (ly == T - 1 && ty < g.th - 1 && (m & M_DOWN)) flags[t + g.tw] = 1;
  if (lx == 0 && tx > 0 && (m & M_LEFT)) flags[t - 1] = 1;
  if (lx == T - 1 && tx < g.tw - 1 && (m & M_RIGHT)) flags[t + 1] = 1;
}

// first = true: seeds from the masks' excess bit (tiles with excess);
// else: from the stored reach bytes of flagged tiles, with halo reach.
template <int T, int NT>
__global__ void __launch_bounds__(NT)
k_flood(Geo g, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
        const uint8_t* __restrict__ mask_g, uint8_t* reach_g, int32_t* fflag, int first,
        unsigned long long* sums, const long long* __restrict__ tpos) {
  constexpr int R = T * T / NT;
  constexpr int HP = T + 2;
  __shared__ int32_t RM[HP * HP];
  const int lx = threadIdx.x % T, chain = threadIdx.x / T;
  const int n = *count;
  if (!first && blockIdx.x == 0 && threadIdx.x == 0 && n > 0)
    atomicAdd(&sums[TALLY_FLOOD * kSumSlots], (unsigned long long)n);
  // first pass (over the tiles holding excess): the final sum of max(e, 0)
  if (first && threadIdx.x == 0) {
    long long p = 0;
    for (int li = blockIdx.x; li < n; li += gridDim.x) p += tpos[list[li]];
    if (p) atomicAdd(&sums[SUM_POSF * kSumSlots + blockIdx.x % kSumSlots], (unsigned long long)p);
  }
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const int t = list[li];
    const size_t base = tbase<T>(t);
    int r[R], r0[R], m[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      const size_t i = base + ly * T + lx;
      m[k] = mask_g[i];
      r0[k] = first ? 0 : (reach_g[i] ? 1 : 0);
      r[k] = r0[k] | ((m[k] & M_EXCESS) ? 1 : 0);
      RM[(ly + 1) * HP + lx + 1] = (m[k] << 1) | r[k];
    }
    load_halo_flood<T>(g, t, reach_g, mask_g, !first, RM);
    __syncthreads();
    flood_relax<T, NT>(r, RM);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int ly = chain * R + k;
      if (r[k] && !r0[k]) {
        reach_g[base + ly * T + lx] = 1;
        flag_out_arcs<T>(g, t, lx, ly, m[k], fflag);
      }
    }
    __syncthreads();
  }
}

// Row-major labels of the image from the tile-blocked reach bytes.
template <int T>
__global__ void k_labels(Geo g, const uint8_t* __restrict__ reach_g, uint8_t* out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= g.W) return;
  const int t = (y / T) * g.tw + x / T;
  out[(size_t)y * g.W + x] = reach_g[tbase<T>(t) + (y % T) * T + (x % T)];
}

// Tile-blocked heights -> row-major padded image (inspection only).
template <int T>
__global__ void k_untile(Geo g, const int32_t* __restrict__ src, int32_t* out) {
  const int PW = g.tw * T;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= PW) return;
  const int t = (y / T) * g.tw + x / T;
  out[(size_t)y * PW + x] = src[tbase<T>(t) + (y % T) * T + (x % T)];
}

}  // namespace gcmi

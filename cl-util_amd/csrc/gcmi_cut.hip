// Host side of the plain grid cut: the gcmi_cut handle, its device buffers
// and the driver loop (GraphCut.cut, GraphCut.pyx:348-418).
//
// Driver, per cut:
//   init                      row-major capacities -> tile-blocked state
//   repeat:
//     up to `bfs` push rounds  compact active tiles, one residency each
//     global relabel           fold pending flow, BFS from the sink set
//     completion check         tiles with e > 0 and h < MAX -> active list
//   until the active list is empty
//   minimal source set        flood from {e > 0}; labels; flow value
// Lists are built on the device; the host reads one list length per round
// to decide whether to continue (the reference reads one per tile-list build
// and waits after every launch, GraphCut.pyx:169-323).
#include <stdlib.h>
#include <string.h>

#include "gcmi_cut.h"

namespace gcmi {

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
    int* ncls = (int*)realloc(cls, sizeof(int) * ncap);
    if (!nev || !ncls) {
      if (nev) ev = nev;
      if (ncls) cls = ncls;
      on = false;
      return;
    }
    ev = nev;
    cls = ncls;
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

CutDev cut_dev(CutState* c) {
  CutDev d;
  d.e = c->e;
  d.h = c->h;
  d.cu = c->cu;
  d.cd = c->cd;
  d.cl = c->cl;
  d.cr = c->cr;
  d.mask = c->mask;
  d.pend = c->pend;
  d.pflag = c->pflag;
  d.act = c->act;
  d.tpos = c->tpos;
  d.sums = c->sums;
  d.cs = c->in[0];
  d.ct = c->in[1];
  d.right = c->in[2];
  d.down = c->in[3];
  return d;
}

template <int T, int NT>
struct CutKernels {
  static void push(CutState* c, int grid, int sweeps, bool first) {
    if (first)
      hipLaunchKernelGGL((k_push<T, NT, true>), dim3(grid), dim3(NT), 0, c->stream, c->g,
                         cut_dev(c), (const int32_t*)nullptr, (const int32_t*)nullptr, sweeps);
    else
      hipLaunchKernelGGL((k_push<T, NT, false>), dim3(grid), dim3(NT), 0, c->stream, c->g,
                         cut_dev(c), c->alist, c->counts + CNT_ACT, sweeps);
  }
  static void fold(CutState* c, int grid) {
    hipLaunchKernelGGL((k_fold<T>), dim3(grid), dim3(kFoldNT), 0, c->stream, c->g, cut_dev(c));
  }
  static void bfs_init(CutState* c, int grid) {
    hipLaunchKernelGGL((k_bfs_init<T, NT>), dim3(grid), dim3(NT), 0, c->stream, c->g, c->mask,
                       c->h, c->bflag, c->haspos, c->act);
  }
  static void bfs_round(CutState* c, int grid) {
    hipLaunchKernelGGL((k_bfs_round<T, NT>), dim3(grid), dim3(NT), 0, c->stream, c->g, c->list,
                       c->counts + CNT_BFS, c->h, c->mask, c->bflag, c->act, c->sums);
  }
  static void flood(CutState* c, int grid, bool first) {
    const int32_t* lst = first ? c->plist : c->list;
    const int32_t* cnt = c->counts + (first ? CNT_POS : CNT_FLOOD);
    hipLaunchKernelGGL((k_flood<T, NT>), dim3(grid), dim3(NT), 0, c->stream, c->g, lst, cnt,
                       c->mask, c->reach, c->fflag, first ? 1 : 0, c->sums,
                       (const long long*)c->tpos);
  }
  static void labels(CutState* c, uint8_t* out) {
    hipLaunchKernelGGL((k_labels<T>), dim3((c->g.W + 255) / 256, c->g.H), dim3(256), 0,
                       c->stream, c->g, c->reach, out);
  }
  static void untile(CutState* c, const int32_t* src, int32_t* out) {
    const int PW = c->g.tw * T;
    hipLaunchKernelGGL((k_untile<T>), dim3((PW + 255) / 256, c->g.th * T), dim3(256), 0,
                       c->stream, c->g, src, out);
  }
};

template <int T, int NT>
static const CutOps* ops_for() {
  static const CutOps ops = {T,
                             NT,
                             &CutKernels<T, NT>::push,
                             &CutKernels<T, NT>::fold,
                             &CutKernels<T, NT>::bfs_init,
                             &CutKernels<T, NT>::bfs_round,
                             &CutKernels<T, NT>::flood,
                             &CutKernels<T, NT>::labels,
                             &CutKernels<T, NT>::untile};
  return &ops;
}

const CutOps* cut_ops(int tile) {
  switch (tile) {
    case 16: return ops_for<16, 256>();
    case 32: return ops_for<32, 256>();
    case 64: return ops_for<64, 1024>();
    default: return nullptr;
  }
}

static int dev_alloc(void** p, size_t bytes, size_t* total) {
  if (hipMalloc(p, bytes ? bytes : 4) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of %zu bytes failed", bytes);
    return GCMI_ENOMEM;
  }
  *total += bytes;
  return GCMI_OK;
}

size_t cut_state_bytes(int W, int H, int tile) {
  const size_t tw = (W + tile - 1) / tile, th = (H + tile - 1) / tile;
  const size_t nt = tw * th, np = nt * tile * tile, n = (size_t)W * H;
  return np * (6 * 4 + 2) + nt * (4 * tile * 4 + 6 * 4) + n * (4 * 4 + 1) + 4096;
}

int cut_state_init(CutState* c, int W, int H, int tile, hipStream_t stream) {
  memset(c, 0, sizeof *c);
  c->ops = cut_ops(tile);
  if (!c->ops) {
    set_error("tile must be 16, 32 or 64 (got %d)", tile);
    return GCMI_EINVAL;
  }
  if (W < 1 || H < 1) {
    set_error("width and height must be >= 1 (got %d x %d)", W, H);
    return GCMI_EINVAL;
  }
  const int64_t tw = (W + tile - 1) / tile, th = (H + tile - 1) / tile;
  const int64_t np = tw * th * tile * tile;
  if (np >= (int64_t(1) << 30)) {
    set_error("padded grid of %lld pixels exceeds 2^30", (long long)np);
    return GCMI_EINVAL;
  }
  c->g.W = W;
  c->g.H = H;
  c->g.tw = (int)tw;
  c->g.th = (int)th;
  c->g.ntiles = (int)(tw * th);
  c->g.maxh = (int32_t)np;
  c->stream = stream;
  const size_t nt = c->g.ntiles;
  const size_t n = (size_t)W * H;
  size_t tot = 0;
  int rc = GCMI_OK;
#define ALLOC(ptr, bytes)                                           \
  if (rc == GCMI_OK) rc = dev_alloc((void**)&(ptr), (bytes), &tot)
  ALLOC(c->e, np * 4);
  ALLOC(c->h, np * 4);
  ALLOC(c->cu, np * 4);
  ALLOC(c->cd, np * 4);
  ALLOC(c->cl, np * 4);
  ALLOC(c->cr, np * 4);
  ALLOC(c->mask, np);
  ALLOC(c->reach, np);
  ALLOC(c->pend, nt * 4 * tile * 4);
  ALLOC(c->pflag, nt * 4);
  ALLOC(c->tpos, nt * 8);
  ALLOC(c->act, nt * 4);
  ALLOC(c->bflag, nt * 4);
  ALLOC(c->fflag, nt * 4);
  ALLOC(c->haspos, nt * 4);
  ALLOC(c->list, nt * 4);
  ALLOC(c->plist, nt * 4);
  ALLOC(c->alist, nt * 4);
  ALLOC(c->counts, CNT_N * 4);
  ALLOC(c->sums, kSumBytes);
  ALLOC(c->scratch, (size_t)compact_scratch_ints(c->g.ntiles) * 4);
  for (int k = 0; k < 4; ++k) ALLOC(c->in[k], n * 4);
  ALLOC(c->labels, n);
#undef ALLOC
  if (rc == GCMI_OK && (hipHostMalloc((void**)&c->host, 256, hipHostMallocDefault) != hipSuccess ||
                        hipHostMalloc((void**)&c->host_sums, kSumBytes, hipHostMallocDefault) !=
                            hipSuccess)) {
    (void)hipGetLastError();
    set_error("pinned host allocation failed");
    rc = GCMI_ENOMEM;
  }
  c->device_bytes = tot;
  if (rc != GCMI_OK) {
    cut_state_free(c);
    return rc;
  }
  GCMI_HIP(hipMemsetAsync(c->pend, 0, nt * 4 * tile * 4, stream));
  GCMI_HIP(hipMemsetAsync(c->act, 0, nt * 4, stream));
  GCMI_HIP(hipMemsetAsync(c->bflag, 0, nt * 4, stream));
  GCMI_HIP(hipMemsetAsync(c->fflag, 0, nt * 4, stream));
  GCMI_HIP(hipStreamSynchronize(stream));
  return GCMI_OK;
}

void cut_state_free(CutState* c) {
  void* ptrs[] = {c->pflag, c->tpos, c->e,     c->h,     c->cu,      c->cd,    c->cl,    c->cr,    c->mask,
                  c->reach, c->pend,  c->act,     c->bflag, c->fflag, c->haspos, c->list,
                  c->plist, c->alist, c->counts, c->sums,   c->scratch, c->in[0], c->in[1], c->in[2],
                  c->in[3], c->labels};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->host) (void)hipHostFree(c->host);
  if (c->host_sums) (void)hipHostFree(c->host_sums);
  const CutOps* ops = c->ops;
  Prof* prof = c->prof;
  memset(c, 0, sizeof *c);
  c->ops = ops;
  c->prof = prof;
}

// Reads the active count (and the BFS count) to the host; one sync.
static int read_counts(CutState* c, int32_t* nact, int32_t* nbfs) {
  int32_t* hc = (int32_t*)c->host;
  GCMI_HIP(hipMemcpyAsync(hc, c->counts, CNT_N * 4, hipMemcpyDeviceToHost, c->stream));
  GCMI_HIP(hipStreamSynchronize(c->stream));
  *nact = hc[CNT_ACT];
  if (nbfs) *nbfs = hc[CNT_BFS];
  c->stats[STAT_SYNCS]++;
  return GCMI_OK;
}

static int compact_flags(CutState* c, int32_t* flags, int32_t* list, int cnt, bool clear) {
  Pred p{};
  p.t1 = flags;
  p.kind = Pred::NONZERO;
  Prof* P = (c->prof && c->prof->on) ? c->prof : nullptr;
  if (P) P->begin(c->stream, K_COMPACT);
  int rc = compact_async(p, c->g.ntiles, list, c->counts + cnt, clear ? flags : nullptr,
                         c->scratch, c->stream);
  if (P) P->end(c->stream);
  return rc;
}

int cut_validate_inputs(const int32_t* cs, const int32_t* ct, const int32_t* right,
                        const int32_t* down, size_t n) {
  int64_t mt = 0, mw = 0;
  for (size_t i = 0; i < n; ++i) {
    if (cs[i] < 0 || ct[i] < 0 || right[i] < 0 || down[i] < 0) {
      set_error("capacities must be >= 0 (pixel %zu)", i);
      return GCMI_EINVAL;
    }
    mt = cs[i] > mt ? cs[i] : mt;
    mt = ct[i] > mt ? ct[i] : mt;
    mw = right[i] > mw ? right[i] : mw;
    mw = down[i] > mw ? down[i] : mw;
  }
  if (mt + 4 * mw >= (int64_t(1) << 31)) {
    set_error("max(source,sink) + 4*max(pairwise) = %lld must be < 2^31",
              (long long)(mt + 4 * mw));
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

int cut_run(CutState* c, int sweeps, int bfs, int64_t max_rounds, int64_t* flow_out) {
  if (!c->uploaded) {
    set_error("cut run before upload");
    return GCMI_EINVAL;
  }
  if (sweeps <= 0) sweeps = kDefaultSweeps;
  if (bfs <= 0) bfs = kDefaultBfs;
  const CutOps* ops = c->ops;
  const int T = ops->tile;
  Prof* P = (c->prof && c->prof->on) ? c->prof : nullptr;
  auto B = [&](int k) {
    if (P) P->begin(c->stream, k);
  };
  auto E = [&]() {
    if (P) P->end(c->stream);
  };
  if (P) P->reset();
  for (int k = 0; k < STAT_N; ++k) c->stats[k] = 0;
  c->stats[STAT_TILES] = c->g.ntiles;
  c->solved = false;
  int dev = 0, ncu = 0;
  GCMI_HIP(hipGetDevice(&dev));
  GCMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid_cap = ncu > 0 ? ncu * 8 : 2048;
  auto grid_for = [&](int n) { return n < grid_cap ? (n > 0 ? n : 1) : grid_cap; };
  const int nt = c->g.ntiles;
  int32_t* hc = (int32_t*)c->host;  // pinned count readback

  GCMI_HIP(hipMemsetAsync(c->sums, 0, kSumBytes, c->stream));
  GCMI_HIP(hipMemsetAsync(c->pend, 0, (size_t)nt * 4 * T * 4, c->stream));
  GCMI_HIP(hipMemsetAsync(c->pflag, 0, (size_t)nt * 4, c->stream));
  GCMI_HIP(hipMemsetAsync(c->act, 0, (size_t)nt * 4, c->stream));

  // round 1 over every tile, building the state from the row-major inputs
  B(K_PUSH_FIRST);
  ops->push(c, nt, sweeps, true);
  E();
  GCMI_HIP(hipGetLastError());
  int64_t rounds = 1;
  c->stats[STAT_RESIDENCIES] += nt;
  int32_t nact = 0;
  for (;;) {
    // further push rounds before the global relabel
    for (int r = 1; r < bfs; ++r) {
      GCMI_TRY(compact_flags(c, c->act, c->alist, CNT_ACT, true));
      GCMI_TRY(read_counts(c, &nact, nullptr));
      if (nact == 0) break;
      if (max_rounds > 0 && rounds >= max_rounds) {
        set_error("cut did not terminate within %lld push rounds", (long long)max_rounds);
        return GCMI_ENOTCONVERGED;
      }
      B(K_PUSH);
      ops->push(c, grid_for(nact), sweeps, false);
      E();
      GCMI_HIP(hipGetLastError());
      ++rounds;
      c->stats[STAT_RESIDENCIES] += nact;
    }
    // fold pending flow, then global relabel + completion test
    B(K_FOLD);
    ops->fold(c, nt);
    E();
    B(K_BFS_INIT);
    ops->bfs_init(c, nt);
    E();
    GCMI_HIP(hipGetLastError());
    c->stats[STAT_GLOBAL_RELABELS]++;
    // later BFS rounds, kRoundBatch per host check; a round whose list is
    // empty exits at once
    bool list_ready = false;
    for (;;) {
      for (int b = 0; b < kRoundBatch; ++b) {
        if (!list_ready) GCMI_TRY(compact_flags(c, c->bflag, c->list, CNT_BFS, true));
        list_ready = false;
        B(K_BFS_ROUND);
        ops->bfs_round(c, grid_cap);
        E();
        GCMI_HIP(hipGetLastError());
      }
      GCMI_TRY(compact_flags(c, c->bflag, c->list, CNT_BFS, true));
      GCMI_TRY(compact_flags(c, c->act, c->alist, CNT_ACT, false));
      int32_t nb = 0;
      GCMI_TRY(read_counts(c, &nact, &nb));
      if (nb == 0) break;
      list_ready = true;
    }
    if (nact == 0) break;
    // the active list was built without clearing the flags (more BFS rounds
    // could still have added tiles); clear them for the push kernel
    GCMI_HIP(hipMemsetAsync(c->act, 0, (size_t)nt * 4, c->stream));
    if (max_rounds > 0 && rounds >= max_rounds) {
      set_error("cut did not terminate within %lld push rounds", (long long)max_rounds);
      return GCMI_ENOTCONVERGED;
    }
    B(K_PUSH);
    ops->push(c, grid_for(nact), sweeps, false);
    E();
    GCMI_HIP(hipGetLastError());
    ++rounds;
    c->stats[STAT_RESIDENCIES] += nact;
  }
  GCMI_TRY(compact_flags(c, c->haspos, c->plist, CNT_POS, false));
  c->stats[STAT_ROUNDS] = rounds;

  // minimal source set
  GCMI_HIP(hipMemsetAsync(c->reach, 0, (size_t)nt * T * T, c->stream));
  B(K_FLOOD);
  ops->flood(c, grid_cap, true);
  E();
  GCMI_HIP(hipGetLastError());
  for (;;) {
    for (int b = 0; b < kRoundBatch; ++b) {
      GCMI_TRY(compact_flags(c, c->fflag, c->list, CNT_FLOOD, true));
      B(K_FLOOD);
      ops->flood(c, grid_cap, false);
      E();
      GCMI_HIP(hipGetLastError());
    }
    // anything flagged by the last round?
    GCMI_TRY(compact_flags(c, c->fflag, c->list, CNT_FLOOD, false));
    GCMI_HIP(hipMemcpyAsync(hc, c->counts + CNT_FLOOD, 4, hipMemcpyDeviceToHost, c->stream));
    B(K_LABELS);
    ops->labels(c, c->labels);
    E();
    GCMI_HIP(hipGetLastError());
    GCMI_HIP(hipMemcpyAsync(c->host_sums, c->sums, kSumBytes, hipMemcpyDeviceToHost,
                            c->stream));
    GCMI_HIP(hipStreamSynchronize(c->stream));
    c->stats[STAT_SYNCS]++;
    if (hc[0] == 0) break;
  }
  const unsigned long long* hs = c->host_sums;
  int64_t row[SUM_ROWS_ALL] = {0};
  for (int r = 0; r < SUM_ROWS_ALL; ++r)
    for (int k = 0; k < kSumSlots; ++k) row[r] += (int64_t)hs[r * kSumSlots + k];
  c->flow = row[SUM_MIN] + row[SUM_POS0] - row[SUM_POSF];
  c->stats[STAT_BFS_PASSES] = c->stats[STAT_GLOBAL_RELABELS] * nt + row[TALLY_BFS];
  c->stats[STAT_FLOOD_PASSES] = row[TALLY_FLOOD];
  if (P) P->collect();
  // algorithmic bytes: first round 16 per image pixel read + 25 per pixel
  // written (state + mask); later residencies 49 (24 read, 25 written);
  // relabel first pass 5 (mask read, h written); later BFS passes 9 (mask +
  // h read, h written); flood passes 2 (mask read, reach written); labels 2.
  const int64_t tp = (int64_t)T * T;
  c->stats[STAT_ALG_BYTES] = (int64_t)c->g.W * c->g.H * 16 + (int64_t)nt * tp * 25 +
                             (c->stats[STAT_RESIDENCIES] - nt) * tp * 49 +
                             c->stats[STAT_GLOBAL_RELABELS] * nt * tp * 5 +
                             row[TALLY_BFS] * tp * 9 + row[TALLY_FLOOD] * tp * 2 +
                             (int64_t)c->g.W * c->g.H * 2;
  c->solved = true;
  if (flow_out) *flow_out = c->flow;
  return GCMI_OK;
}

}  // namespace gcmi

using namespace gcmi;

struct gcmi_cut {
  CutState s;
  Prof prof;
};

extern "C" {

int gcmi_cut_create(int width, int height, int tile, gcmi_cut** out) {
  if (!out) { set_error("null out"); return GCMI_EINVAL; }
  *out = nullptr;
  if (!cut_ops(tile)) {
    set_error("tile must be 16, 32 or 64 (got %d)", tile);
    return GCMI_EINVAL;
  }
  if (width < 1 || height < 1) {
    set_error("width and height must be >= 1 (got %d x %d)", width, height);
    return GCMI_EINVAL;
  }
  GCMI_TRY(require_device());
  gcmi_cut* c = (gcmi_cut*)calloc(1, sizeof(gcmi_cut));
  if (!c) { set_error("host allocation failed"); return GCMI_ENOMEM; }
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    free(c);
    set_error("hipStreamCreate failed");
    return GCMI_EHIP;
  }
  int rc = cut_state_init(&c->s, width, height, tile, st);
  if (rc != GCMI_OK) {
    (void)hipStreamDestroy(st);
    free(c);
    return rc;
  }
  c->s.prof = &c->prof;
  *out = c;
  return GCMI_OK;
}

int gcmi_cut_destroy(gcmi_cut* c) {
  if (!c) return GCMI_OK;
  hipStream_t st = c->s.stream;
  if (st) (void)hipStreamSynchronize(st);
  cut_state_free(&c->s);
  c->prof.release();
  if (st) (void)hipStreamDestroy(st);
  free(c);
  return GCMI_OK;
}

int gcmi_cut_profile(gcmi_cut* c, int enable) {
  if (!c) { set_error("null handle"); return GCMI_EINVAL; }
  c->prof.on = enable != 0;
  c->prof.reset();
  return GCMI_OK;
}

int gcmi_cut_kernel_times(gcmi_cut* c, double* ms_out, int64_t* launches_out, int n) {
  if (!c || n < 0) { set_error("bad argument"); return GCMI_EINVAL; }
  for (int k = 0; k < n && k < K_N; ++k) {
    if (ms_out) ms_out[k] = c->prof.ms[k];
    if (launches_out) launches_out[k] = c->prof.launches[k];
  }
  return GCMI_OK;
}

int gcmi_cut_upload(gcmi_cut* c, const int32_t* h_source, const int32_t* h_sink,
                    const int32_t* h_right, const int32_t* h_down) {
  if (!c || !h_source || !h_sink || !h_right || !h_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  const size_t n = (size_t)c->s.g.W * c->s.g.H;
  GCMI_TRY(cut_validate_inputs(h_source, h_sink, h_right, h_down, n));
  const int32_t* src[4] = {h_source, h_sink, h_right, h_down};
  for (int k = 0; k < 4; ++k)
    GCMI_HIP(hipMemcpyAsync(c->s.in[k], src[k], n * 4, hipMemcpyHostToDevice, c->s.stream));
  GCMI_HIP(hipStreamSynchronize(c->s.stream));
  c->s.uploaded = true;
  c->s.solved = false;
  return GCMI_OK;
}

int gcmi_cut_upload_device(gcmi_cut* c, const int32_t* d_source, const int32_t* d_sink,
                           const int32_t* d_right, const int32_t* d_down) {
  if (!c || !d_source || !d_sink || !d_right || !d_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  const size_t n = (size_t)c->s.g.W * c->s.g.H;
  const int32_t* src[4] = {d_source, d_sink, d_right, d_down};
  for (int k = 0; k < 4; ++k)
    GCMI_HIP(hipMemcpyAsync(c->s.in[k], src[k], n * 4, hipMemcpyDeviceToDevice, c->s.stream));
  GCMI_HIP(hipStreamSynchronize(c->s.stream));
  c->s.uploaded = true;
  c->s.solved = false;
  return GCMI_OK;
}

int gcmi_cut_run(gcmi_cut* c, int sweeps, int bfs, int64_t max_rounds, int64_t* flow_out) {
  if (!c) { set_error("null handle"); return GCMI_EINVAL; }
  return cut_run(&c->s, sweeps, bfs, max_rounds, flow_out);
}

int gcmi_cut_labels(gcmi_cut* c, uint8_t* h_out) {
  if (!c || !h_out) { set_error("null argument"); return GCMI_EINVAL; }
  if (!c->s.solved) { set_error("no completed cut"); return GCMI_EINVAL; }
  const size_t n = (size_t)c->s.g.W * c->s.g.H;
  GCMI_HIP(hipMemcpyAsync(h_out, c->s.labels, n, hipMemcpyDeviceToHost, c->s.stream));
  GCMI_HIP(hipStreamSynchronize(c->s.stream));
  return GCMI_OK;
}

int gcmi_cut_labels_device(gcmi_cut* c, uint8_t* d_out) {
  if (!c || !d_out) { set_error("null argument"); return GCMI_EINVAL; }
  if (!c->s.solved) { set_error("no completed cut"); return GCMI_EINVAL; }
  const size_t n = (size_t)c->s.g.W * c->s.g.H;
  GCMI_HIP(hipMemcpyAsync(d_out, c->s.labels, n, hipMemcpyDeviceToDevice, c->s.stream));
  GCMI_HIP(hipStreamSynchronize(c->s.stream));
  return GCMI_OK;
}

int gcmi_cut_stats(gcmi_cut* c, int64_t* out, int n) {
  if (!c || !out || n < 0) { set_error("bad argument"); return GCMI_EINVAL; }
  for (int k = 0; k < n && k < STAT_N; ++k) out[k] = c->s.stats[k];
  return GCMI_OK;
}

int gcmi_cut_heights(gcmi_cut* c, int32_t* h_out) {
  if (!c || !h_out) { set_error("null argument"); return GCMI_EINVAL; }
  const int T = c->s.ops->tile;
  const size_t np = (size_t)c->s.g.ntiles * T * T;
  int32_t* tmp = nullptr;
  if (hipMalloc(&tmp, np * 4) != hipSuccess) {
    (void)hipGetLastError();
    set_error("allocation failed");
    return GCMI_ENOMEM;
  }
  c->s.ops->untile(&c->s, c->s.h, tmp);
  hipError_t e1 = hipMemcpyAsync(h_out, tmp, np * 4, hipMemcpyDeviceToHost, c->s.stream);
  hipError_t e2 = hipStreamSynchronize(c->s.stream);
  (void)hipFree(tmp);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    set_error("height readback failed");
    return GCMI_EHIP;
  }
  return GCMI_OK;
}

int gcmi_cut_geometry(gcmi_cut* c, int* out3) {
  if (!c || !out3) { set_error("null argument"); return GCMI_EINVAL; }
  const int T = c->s.ops->tile;
  out3[0] = c->s.g.tw * T;
  out3[1] = c->s.g.th * T;
  out3[2] = T;
  return GCMI_OK;
}

}  // extern "C"

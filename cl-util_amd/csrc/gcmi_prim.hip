// Tile-list primitives: prefix sum, stream compaction, IncrementalTileList
// flag / logical / increment kernels.
//
// The reference runs a recursive Blelloch scan (prefixsum.cl:114-212) with a
// blocking wait per level, a copy + scan + in-place scatter for compaction
// (StreamCompact.pyx:66-82) and a separate flag kernel before it
// (IncrementalTileList.pyx:119-136).  Here the predicate is evaluated inside
// the compaction kernel (no flag buffer), a wave's share of the output comes
// from one 64-bit ballot and a popcount, and sizes up to kSingleMax run as a
// single work-group launch; larger inputs use count / scan / write passes.
// The entry points only enqueue their kernels: a result's length stays in
// device memory, and reading it back is what synchronises (as the
// reference's lazy IncrementalTileList.length, IncrementalTileList.pyx:
// 106-112).
#include "gcmi_common.h"

namespace gcmi {

constexpr int kSingleNT = 1024;            // threads of the one-block kernels
constexpr int kSingleMax = 1 << 14;        // one-block path up to this n (one pass)
constexpr int kMultiNT = 256;              // threads of the multi-block kernels
constexpr int kMultiIT = 16;                    // elements per thread of a multi-block pass
constexpr int kMultiChunk = kMultiNT * kMultiIT; // elements per block (4096): one step each

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int lane = lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// ------------------------------------------------------------------ scan

// In-place exclusive scan of a[lo, hi), starting from `carry`; returns the
// final carry (the same value in every thread).  4 elements per thread per
// step, so one step covers 4*NT elements.
template <int NT>
__device__ int block_scan_range(int32_t* a, int lo, int hi, int carry, int* wsum) {
  for (int base = lo; base < hi; base += 4 * NT) {
    const int i0 = base + 4 * threadIdx.x;
    int v[4];
    int s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = (i0 + j < hi) ? a[i0 + j] : 0;
      s += v[j];
    }
    int total = 0;
    int pre = block_excl_scan<NT>(s, wsum, &total) + carry;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (i0 + j < hi) a[i0 + j] = pre;
      pre += v[j];
    }
    carry += total;
  }
  return carry;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_scan_single(int32_t* a, int32_t* total, int n) {
  __shared__ int wsum[NT / kWave];
  int c = block_scan_range<NT>(a, 0, n, 0, wsum);
  if (threadIdx.x == 0 && total) *total = c;
}

// Sum of each block's chunk into part[blockIdx.x].
__global__ void __launch_bounds__(kMultiNT) k_scan_reduce(const int32_t* a, int n, int32_t* part) {
  __shared__ int wsum[kMultiNT / kWave];
  const int lo = blockIdx.x * kMultiChunk;
  const int hi = min(n, lo + kMultiChunk);
  int s = 0;
  for (int i = lo + threadIdx.x; i < hi; i += kMultiNT) s += a[i];
  s = wave_sum(s);
  if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kMultiNT / kWave; ++w) t += wsum[w];
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kMultiNT) k_scan_apply(int32_t* a, int n, const int32_t* part) {
  __shared__ int wsum[kMultiNT / kWave];
  const int lo = blockIdx.x * kMultiChunk;
  const int hi = min(n, lo + kMultiChunk);
  block_scan_range<kMultiNT>(a, lo, hi, part[blockIdx.x], wsum);
}

// ------------------------------------------------------------ compaction

// Compacts indices of [lo, hi) satisfying pred into list[carry...]; returns
// the final carry.  Each thread owns IT consecutive elements per step and
// keeps their predicate bits in one register; the block's offsets come from
// one exclusive scan of the per-thread counts, whose wave part is a 64-lane
// shuffle scan.
template <int NT, int IT>
__device__ int block_compact_range(const Pred& pred, int lo, int hi, int carry,
                                   int32_t* list, int32_t* clear, int* wsum) {
  static_assert(IT % 4 == 0 && IT <= 32, "IT: whole int4 groups, one 32-bit mask");
  if (pred.ids) {
    // gather form: flags of listed ids
    for (int base = lo; base < hi; base += NT * IT) {
      const int i0 = base + IT * threadIdx.x;
      int id[IT], v[IT];
      if (i0 + IT <= hi && (reinterpret_cast<uintptr_t>(pred.ids + i0) & 15) == 0) {
#pragma unroll
        for (int j = 0; j < IT; j += 4) {
          const int4 a4 = *reinterpret_cast<const int4*>(pred.ids + i0 + j);
          id[j] = a4.x; id[j + 1] = a4.y; id[j + 2] = a4.z; id[j + 3] = a4.w;
        }
      } else {
#pragma unroll
        for (int j = 0; j < IT; ++j) id[j] = i0 + j < hi ? pred.ids[i0 + j] : -1;
      }
      // all gathers issued before any is used
#pragma unroll
      for (int j = 0; j < IT; ++j) v[j] = pred.t1[id[j] >= 0 ? id[j] : 0];
      unsigned bits = 0;
#pragma unroll
      for (int j = 0; j < IT; ++j)
        if (id[j] >= 0 && v[j] != 0) bits |= 1u << j;
      int total = 0;
      int off = block_excl_scan<NT>(__popc(bits), wsum, &total) + carry;
      for (unsigned b = bits; b; b &= b - 1) list[off++] = pred.ids[i0 + __ffs(b) - 1];
      if (clear) {
#pragma unroll
        for (int j = 0; j < IT; ++j)
          if (bits & (1u << j)) clear[id[j]] = 0;
      }
      carry += total;
    }
    return carry;
  }
  for (int base = lo; base < hi; base += NT * IT) {
    const int i0 = base + IT * threadIdx.x;
    // all loads first (independent, so they overlap), then the predicate
    int v1[IT], v2[IT];
    const bool logical = pred.kind == Pred::LOGICAL;
    if (i0 + IT <= hi && ((reinterpret_cast<uintptr_t>(pred.t1 + i0) & 15) == 0) &&
        (!logical || (reinterpret_cast<uintptr_t>(pred.t2 + i0) & 15) == 0)) {
#pragma unroll
      for (int j = 0; j < IT; j += 4) {
        const int4 a4 = *reinterpret_cast<const int4*>(pred.t1 + i0 + j);
        v1[j] = a4.x; v1[j + 1] = a4.y; v1[j + 2] = a4.z; v1[j + 3] = a4.w;
        if (logical) {
          const int4 b4 = *reinterpret_cast<const int4*>(pred.t2 + i0 + j);
          v2[j] = b4.x; v2[j + 1] = b4.y; v2[j + 2] = b4.z; v2[j + 3] = b4.w;
        } else {
          v2[j] = v2[j + 1] = v2[j + 2] = v2[j + 3] = 0;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const bool in = i0 + j < hi;
        v1[j] = in ? pred.t1[i0 + j] : 0;
        v2[j] = (in && logical) ? pred.t2[i0 + j] : 0;
      }
    }
    unsigned bits = 0;
#pragma unroll
    for (int j = 0; j < IT; ++j)
      if (i0 + j < hi && pred.eval_values(v1[j], v2[j])) bits |= 1u << j;
    int total = 0;
    int off = block_excl_scan<NT>(__popc(bits), wsum, &total) + carry;
    while (bits) {
      const int j = __ffs(bits) - 1;
      bits &= bits - 1;
      list[off++] = i0 + j;
    }
    if (clear) {
      if (i0 + IT <= hi && (reinterpret_cast<uintptr_t>(clear + i0) & 15) == 0) {
#pragma unroll
        for (int j = 0; j < IT; j += 4)
          *reinterpret_cast<int4*>(clear + i0 + j) = make_int4(0, 0, 0, 0);
      } else {
#pragma unroll
        for (int j = 0; j < IT; ++j)
          if (i0 + j < hi) clear[i0 + j] = 0;
      }
    }
    carry += total;
  }
  return carry;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_compact_single(Pred pred, int n, int32_t* list,
                                                      int32_t* count, int32_t* clear) {
  __shared__ int wcnt[NT / kWave];
  int c = block_compact_range<NT, 16>(pred, 0, n, 0, list, clear, wcnt);
  if (threadIdx.x == 0) *count = c;
}

__global__ void __launch_bounds__(kMultiNT) k_compact_count(Pred pred, int n, int32_t* part) {
  __shared__ int wsum[kMultiNT / kWave];
  const int lo = blockIdx.x * kMultiChunk;
  const int hi = min(n, lo + kMultiChunk);
  // every predicate of the chunk issued before any is used (the gathers of
  // a selection-list predicate overlap)
  bool f[kMultiIT];
#pragma unroll
  for (int j = 0; j < kMultiIT; ++j) {
    const int i = lo + j * kMultiNT + threadIdx.x;
    f[j] = i < hi && pred.eval(i);
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < kMultiIT; ++j) s += __popcll(__ballot(f[j]));
  if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kMultiNT / kWave; ++w) t += wsum[w];
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kMultiNT) k_compact_write(Pred pred, int n, const int32_t* part,
                                                           int32_t* list, int32_t* clear) {
  __shared__ int wcnt[kMultiNT / kWave];
  const int lo = blockIdx.x * kMultiChunk;
  const int hi = min(n, lo + kMultiChunk);
  block_compact_range<kMultiNT, kMultiIT>(pred, lo, hi, part[blockIdx.x], list, clear, wcnt);
}

// ------------------------------------------------------------ tile kernels

__global__ void k_fill(int32_t* t, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = v;
}

__global__ void k_flag(Pred pred, int n, int32_t* flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = pred.eval(i) ? 1 : 0;
}

// increment_logical: in place on t1 (the predicate reads t1[i] first).
__global__ void k_increment(Pred pred, int n, int32_t* t1, int32_t iteration) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && pred.eval(i)) t1[i] = iteration;
}

// --------------------------------------------------------------- host side

static int nblocks_multi(int n) { return (n + kMultiChunk - 1) / kMultiChunk; }

int compact_scratch_ints(int n) { return n > kSingleMax ? nblocks_multi(n) + 1 : 1; }

int compact_async(const Pred& pred, int n, int32_t* list, int32_t* d_count,
                  int32_t* clear, int32_t* scratch, hipStream_t stream) {
  if (n <= kSingleMax) {
    hipLaunchKernelGGL(k_compact_single<kSingleNT>, dim3(1), dim3(kSingleNT), 0, stream,
                       pred, n, list, d_count, clear);
  } else {
    const int nb = nblocks_multi(n);
    hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(kMultiNT), 0, stream, pred, n, scratch);
    hipLaunchKernelGGL(k_scan_single<kSingleNT>, dim3(1), dim3(kSingleNT), 0, stream,
                       scratch, d_count, nb);
    hipLaunchKernelGGL(k_compact_write, dim3(nb), dim3(kMultiNT), 0, stream, pred, n,
                       (const int32_t*)scratch, list, clear);
  }
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int scan_async(int32_t* a, int32_t* d_total, int n, int32_t* scratch, hipStream_t stream) {
  if (n <= kSingleMax) {
    hipLaunchKernelGGL(k_scan_single<kSingleNT>, dim3(1), dim3(kSingleNT), 0, stream, a,
                       d_total, n);
  } else {
    const int nb = nblocks_multi(n);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kMultiNT), 0, stream, (const int32_t*)a, n,
                       scratch);
    hipLaunchKernelGGL(k_scan_single<kSingleNT>, dim3(1), dim3(kSingleNT), 0, stream, scratch,
                       d_total, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kMultiNT), 0, stream, a, n,
                       (const int32_t*)scratch);
  }
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

}  // namespace gcmi

// Scratch of gcmi_tiles_build_noalloc: the per-block counts of its
// multi-block compaction (one per 4096 tiles) and their scan.  Each device
// holds its own copy until the process ends.  Host code never uses this
// name as an address: it takes the device address from hipGetSymbolAddress
// (static_scratch below).
constexpr int kNoAllocScratchInts = GCMI_NOALLOC_MAX_TILES / gcmi::kMultiChunk + 64;
static __device__ __align__(256) int32_t scratch[kNoAllocScratchInts];

using namespace gcmi;

namespace {

// Operator and logical values are not rejected: as in the reference's
// switch statements (incremental_tile_list.cl:49-55, 100-103) an unknown
// operator flags nothing and an unknown logical keeps the first flag.
int check_op(int) { return GCMI_OK; }
int check_logical(int) { return GCMI_OK; }

int check_n(int n) {
  if (n < 0) {
    set_error("negative length %d", n);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

Pred cmp_pred(const int32_t* t, int op, int32_t operand) {
  Pred p{};
  p.t1 = t;
  p.kind = Pred::CMP;
  p.op1 = op;
  p.a = operand;
  return p;
}

Pred logical_pred(const int32_t* t1, const int32_t* t2, int op1, int op2, int32_t a,
                  int32_t b, int logical) {
  Pred p{};
  p.t1 = t1;
  p.t2 = t2;
  p.kind = Pred::LOGICAL;
  p.op1 = op1;
  p.op2 = op2;
  p.a = a;
  p.b = b;
  p.logical = logical;
  return p;
}

int grid1d(int n) { return (n + 255) / 256; }

int run_compact(const Pred& p, int n, int32_t* list, int32_t* len, void* stream) {
  hipStream_t s = pick_stream(stream);
  ScratchLease lease;
  int32_t* scr = nullptr;
  GCMI_TRY(lease.acquire(s, compact_scratch_ints(n), &scr));
  return compact_async(p, n, list, len, nullptr, scr, s);
}

// Device address of this file's `scratch` array, per device: taken once
// per device (with that device current) by hipGetSymbolAddress.
constexpr int kStaticDevices = 64;
std::mutex g_static_mu;
int32_t* g_static_ptr[kStaticDevices] = {};

int static_scratch(int32_t** out) {
  int dev = 0;
  GCMI_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kStaticDevices) {
    set_error("device %d beyond the %d the library tracks", dev, kStaticDevices);
    return GCMI_ENODEVICE;
  }
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return GCMI_EHIP;
    }
    g_static_ptr[dev] = static_cast<int32_t*>(ptr);
  }
  *out = g_static_ptr[dev];
  return GCMI_OK;
}

}  // namespace

extern "C" {

int gcmi_memset_i32(int32_t* d_dst, int32_t value, size_t count, void* stream) {
  GCMI_TRY(require_device());
  if (!count) return GCMI_OK;
  if (!d_dst || count > 0x7fffffff) { set_error("bad memset arguments"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  hipLaunchKernelGGL(k_fill, dim3(grid1d((int)count)), dim3(256), 0, s, d_dst, (int)count, value);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int gcmi_prefix_sum(int32_t* d_array, int32_t* d_total, int n, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  if (!d_total || (n > 0 && !d_array)) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  ScratchLease lease;
  int32_t* scr = nullptr;
  GCMI_TRY(lease.acquire(s, compact_scratch_ints(n), &scr));
  return scan_async(d_array, d_total, n, scr, s);
}

int gcmi_prefix_sum_timed(int32_t* d_array, int32_t* d_total, int n, void* stream,
                          int64_t* ns_out) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  if (!d_total || !ns_out || (n > 0 && !d_array)) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (hipEvent_t& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      if (ev[0]) (void)hipEventDestroy(ev[0]);
      set_error("event creation failed");
      return GCMI_EHIP;
    }
  int rc = GCMI_OK;
  float ms = 0;
  if (hipEventRecord(ev[0], s) != hipSuccess) rc = GCMI_EHIP;
  if (rc == GCMI_OK) {
    ScratchLease lease;
    int32_t* scr = nullptr;
    rc = lease.acquire(s, compact_scratch_ints(n), &scr);
    if (rc == GCMI_OK) rc = scan_async(d_array, d_total, n, scr, s);
  }
  if (rc == GCMI_OK && (hipEventRecord(ev[1], s) != hipSuccess ||
                        hipEventSynchronize(ev[1]) != hipSuccess ||
                        hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess)) {
    (void)hipGetLastError();
    set_error("scan timing failed");
    rc = GCMI_EHIP;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc == GCMI_OK) *ns_out = (int64_t)((double)ms * 1e6);
  return rc;
}

int gcmi_stream_compact(const int32_t* d_flags, int32_t* d_list, int32_t* d_length, int n,
                        void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  if (!d_length || (n > 0 && (!d_flags || !d_list))) { set_error("null pointer"); return GCMI_EINVAL; }
  Pred p{};
  p.t1 = d_flags;
  p.kind = Pred::NONZERO;
  return run_compact(p, n, d_list, d_length, stream);
}

int gcmi_tiles_fill(int32_t* d_tiles, int n, int32_t value, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  if (n == 0) return GCMI_OK;
  if (!d_tiles) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  hipLaunchKernelGGL(k_fill, dim3(grid1d(n)), dim3(256), 0, s, d_tiles, n, value);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int gcmi_tiles_flag(const int32_t* d_tiles, int32_t* d_flags, int n, int op, int32_t operand,
                    void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  GCMI_TRY(check_op(op));
  if (n == 0) return GCMI_OK;
  if (!d_tiles || !d_flags) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  hipLaunchKernelGGL(k_flag, dim3(grid1d(n)), dim3(256), 0, s, cmp_pred(d_tiles, op, operand), n,
                     d_flags);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int gcmi_tiles_flag_logical(const int32_t* d_t1, const int32_t* d_t2, int32_t* d_flags, int n,
                            int op1, int op2, int32_t a, int32_t b, int logical, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  GCMI_TRY(check_op(op1));
  GCMI_TRY(check_op(op2));
  GCMI_TRY(check_logical(logical));
  if (n == 0) return GCMI_OK;
  if (!d_t1 || !d_t2 || !d_flags) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  hipLaunchKernelGGL(k_flag, dim3(grid1d(n)), dim3(256), 0, s,
                     logical_pred(d_t1, d_t2, op1, op2, a, b, logical), n, d_flags);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int gcmi_tiles_increment_logical(int32_t* d_t1, const int32_t* d_t2, int n, int op1, int op2,
                                 int32_t a, int32_t b, int logical, int32_t iteration,
                                 void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  GCMI_TRY(check_op(op1));
  GCMI_TRY(check_op(op2));
  GCMI_TRY(check_logical(logical));
  if (n == 0) return GCMI_OK;
  if (!d_t1 || !d_t2) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = pick_stream(stream);
  hipLaunchKernelGGL(k_increment, dim3(grid1d(n)), dim3(256), 0, s,
                     logical_pred(d_t1, d_t2, op1, op2, a, b, logical), n, d_t1, iteration);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int gcmi_tiles_build(const int32_t* d_tiles, int n, int op, int32_t operand, int32_t* d_list,
                     int32_t* d_length, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  GCMI_TRY(check_op(op));
  if (!d_length || (n > 0 && (!d_tiles || !d_list))) { set_error("null pointer"); return GCMI_EINVAL; }
  return run_compact(cmp_pred(d_tiles, op, operand), n, d_list, d_length, stream);
}

int gcmi_tiles_build_noalloc(const int32_t* d_tiles, int n, int op, int32_t operand,
                             int32_t* d_list, int32_t* d_length, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  GCMI_TRY(check_op(op));
  if (!d_length || (n > 0 && (!d_tiles || !d_list))) { set_error("null pointer"); return GCMI_EINVAL; }
  if (n > GCMI_NOALLOC_MAX_TILES) {
    set_error("gcmi_tiles_build_noalloc: %d tiles exceed its maximum of %d", n,
              GCMI_NOALLOC_MAX_TILES);
    return GCMI_EINVAL;
  }
  int32_t* scr = nullptr;
  GCMI_TRY(static_scratch(&scr));
  // NULL: HIP's null stream (the library creates no stream here)
  return compact_async(cmp_pred(d_tiles, op, operand), n, d_list, d_length, nullptr, scr,
                       reinterpret_cast<hipStream_t>(stream));
}

int gcmi_tiles_build_logical(const int32_t* d_t1, const int32_t* d_t2, int n, int op1, int op2,
                             int32_t a, int32_t b, int logical, int32_t* d_list,
                             int32_t* d_length, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_n(n));
  GCMI_TRY(check_op(op1));
  GCMI_TRY(check_op(op2));
  GCMI_TRY(check_logical(logical));
  if (!d_length || (n > 0 && (!d_t1 || !d_t2 || !d_list))) { set_error("null pointer"); return GCMI_EINVAL; }
  return run_compact(logical_pred(d_t1, d_t2, op1, op2, a, b, logical), n, d_list, d_length,
                     stream);
}

}  // extern "C"

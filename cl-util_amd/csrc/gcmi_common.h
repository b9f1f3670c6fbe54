// Shared host/device helpers for libgcmi (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <string>

#include "../../include/gcmi.h"

namespace gcmi {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
void clear_error();

#define GCMI_HIP(call)                                                        \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::gcmi::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                  \
      return GCMI_EHIP;                                                       \
    }                                                                         \
  } while (0)

#define GCMI_TRY(expr)          \
  do {                          \
    int _rc = (expr);           \
    if (_rc != GCMI_OK) return _rc; \
  } while (0)

// Fails with GCMI_ENODEVICE when no HIP device is visible.
int require_device();
// The library's own stream for the current device.
hipStream_t default_stream();
inline hipStream_t pick_stream(void* s) {
  return s ? reinterpret_cast<hipStream_t>(s) : default_stream();
}

// ---------------------------------------------------------------- device
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive scan of v across the 64 lanes of a wave.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int t = __shfl_up(v, off, kWave);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Exclusive scan across a block of NT threads; *total receives the block sum.
// `wsum` is LDS scratch of at least NT/64 ints.  Contains barriers: every
// thread of the block must call it.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int* total) {
  constexpr int NW = NT / kWave;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  int incl = wave_incl_scan(v);
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    int s = wsum[w];
    before += (w < wid) ? s : 0;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Comparison of IncrementalTileList (incremental_tile_list.cl:49-55); an
// unknown operator flags nothing, as the reference's switch does.
__device__ __forceinline__ bool tile_cmp(int32_t v, int op, int32_t operand) {
  switch (op) {
    case GCMI_OP_EQUAL: return v == operand;
    case GCMI_OP_GT: return v > operand;
    case GCMI_OP_LT: return v < operand;
    case GCMI_OP_GTE: return v >= operand;
    case GCMI_OP_LTE: return v <= operand;
    default: return false;
  }
}

// ------------------------------------------------------- compaction API
// Predicate over i < n used by the fused compaction kernels.
struct Pred {
  enum Kind { NONZERO = 0, CMP = 1, LOGICAL = 2 };
  const int32_t* t1;
  const int32_t* t2;
  int kind;
  int op1, op2;
  int32_t a, b;
  int logical;
  // Optional indirection: element i stands for id ids[i] -- the predicate
  // reads t1[ids[i]] (NONZERO only), the output is ids[i] and clearing
  // zeroes t1[ids[i]].
  const int32_t* ids;

  __device__ __forceinline__ bool eval(int i) const {
    if (ids) return t1[ids[i]] != 0;
    return eval_values(t1[i], kind == LOGICAL ? t2[i] : 0);
  }

  // The predicate on already-loaded values (v2 is ignored unless LOGICAL).
  __device__ __forceinline__ bool eval_values(int32_t v1, int32_t v2) const {
    if (kind == NONZERO) return v1 != 0;
    bool f1 = tile_cmp(v1, op1, a);
    if (kind == CMP) return f1;
    bool f2 = tile_cmp(v2, op2, b);
    if (logical == GCMI_LOGICAL_AND) return f1 && f2;
    if (logical == GCMI_LOGICAL_OR) return f1 || f2;
    return f1;
  }
};

// Scratch for the C-ABI primitives' multi-block passes: one buffer per
// (device, stream), grown on demand.  acquire() takes the device's scratch
// lock and keeps it until the lease ends, so enqueue every kernel that uses
// the scratch before then: a later call on the same stream is ordered after
// them, one on another stream gets another buffer, and a buffer is grown
// (freed and reallocated) only after its stream has drained.
class ScratchLease {
 public:
  int acquire(hipStream_t stream, int count, int32_t** out);

 private:
  std::unique_lock<std::mutex> lk_;
};

// Ints of scratch that compact_async / scan_async need for n elements.
int compact_scratch_ints(int n);

// Ascending indices i < n with pred(i) into list[], count into *d_count.
// If `clear` is non-null, clear[i] is zeroed for every i < n (the
// double-buffered flag arrays of the cut use this).  `scratch` holds at
// least compact_scratch_ints(n) ints.  Asynchronous on `stream`.
int compact_async(const Pred& pred, int n, int32_t* list, int32_t* d_count,
                  int32_t* clear, int32_t* scratch, hipStream_t stream);

// In-place exclusive scan of a[0..n) with the total in *d_total.
int scan_async(int32_t* a, int32_t* d_total, int n, int32_t* scratch,
               hipStream_t stream);

}  // namespace gcmi

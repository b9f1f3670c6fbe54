// Runtime plumbing of libgcmi: error state, device discovery, per-device
// default streams and scratch, and the device-memory entry points.
#include <stdarg.h>
#include <stdio.h>

#include <mutex>
#include <vector>

#include "gcmi_common.h"

namespace gcmi {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

namespace {
struct DeviceState {
  hipStream_t stream = nullptr;
  int32_t* scratch = nullptr;
  int scratch_count = 0;
};
std::mutex g_mu;
std::vector<DeviceState> g_dev;

int current_state(DeviceState** out) {
  int dev = 0;
  GCMI_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  *out = &g_dev[dev];
  return GCMI_OK;
}
}  // namespace

int require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    set_error("no HIP device available");
    return GCMI_ENODEVICE;
  }
  return GCMI_OK;
}

hipStream_t default_stream() {
  DeviceState* st = nullptr;
  if (current_state(&st) != GCMI_OK) return nullptr;
  if (!st->stream) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess)
      st->stream = s;
  }
  return st->stream;
}

int scratch_i32(int count, int32_t** out) {
  DeviceState* st = nullptr;
  GCMI_TRY(current_state(&st));
  if (st->scratch_count < count) {
    if (st->scratch) {
      GCMI_HIP(hipDeviceSynchronize());
      GCMI_HIP(hipFree(st->scratch));
      st->scratch = nullptr;
      st->scratch_count = 0;
    }
    int want = count < 4096 ? 4096 : count;
    if (hipMalloc(&st->scratch, sizeof(int32_t) * (size_t)want) != hipSuccess) {
      set_error("scratch allocation of %d ints failed", want);
      return GCMI_ENOMEM;
    }
    st->scratch_count = want;
  }
  *out = st->scratch;
  return GCMI_OK;
}

}  // namespace gcmi

using namespace gcmi;

extern "C" {

const char* gcmi_last_error(void) { return g_err; }

const char* gcmi_version(void) { return "gcmi 0.1 gfx950"; }

int gcmi_device_count(int* out) {
  if (!out) { set_error("null out"); return GCMI_EINVAL; }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *out = n;
  return GCMI_OK;
}

int gcmi_malloc(void** d_out, size_t bytes) {
  if (!d_out) { set_error("null out"); return GCMI_EINVAL; }
  GCMI_TRY(require_device());
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipMalloc(%zu) failed", bytes);
    return GCMI_ENOMEM;
  }
  *d_out = p;
  return GCMI_OK;
}

int gcmi_free(void* d_ptr) {
  if (!d_ptr) return GCMI_OK;
  GCMI_TRY(require_device());
  GCMI_HIP(hipFree(d_ptr));
  return GCMI_OK;
}

int gcmi_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes) {
  GCMI_TRY(require_device());
  if (!bytes) return GCMI_OK;
  if (!d_dst || !h_src) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = default_stream();
  GCMI_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
  GCMI_HIP(hipStreamSynchronize(s));
  return GCMI_OK;
}

int gcmi_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes) {
  GCMI_TRY(require_device());
  if (!bytes) return GCMI_OK;
  if (!h_dst || !d_src) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = default_stream();
  GCMI_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
  GCMI_HIP(hipStreamSynchronize(s));
  return GCMI_OK;
}

int gcmi_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes) {
  GCMI_TRY(require_device());
  if (!bytes) return GCMI_OK;
  if (!d_dst || !d_src) { set_error("null pointer"); return GCMI_EINVAL; }
  hipStream_t s = default_stream();
  GCMI_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, s));
  GCMI_HIP(hipStreamSynchronize(s));
  return GCMI_OK;
}

int gcmi_stream_sync(void* stream) {
  GCMI_TRY(require_device());
  GCMI_HIP(hipStreamSynchronize(pick_stream(stream)));
  return GCMI_OK;
}

}  // extern "C"

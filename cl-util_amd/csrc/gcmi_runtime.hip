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
// Per device: the library's own stream and the primitives' scratch, one
// buffer per stream the primitives were called on.  States live in a fixed
// array, so a pointer to one stays valid while other devices are added.
constexpr int kMaxDevices = 64;
struct StreamScratch {
  hipStream_t stream;
  int32_t* buf;
  int count;
};
struct DeviceState {
  std::mutex mu;
  hipStream_t stream = nullptr;
  std::vector<StreamScratch> scratch;
};
DeviceState g_dev[kMaxDevices];

int current_state(DeviceState** out) {
  int dev = 0;
  GCMI_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d beyond the %d the library tracks", dev, kMaxDevices);
    return GCMI_ENODEVICE;
  }
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

// The library's stream if it was created already, else null (no stream is
// created: every extra stream takes a share of the device's few hardware
// queues, which concurrent solvers in the same process need).
static hipStream_t existing_default_stream() {
  DeviceState* st = nullptr;
  if (current_state(&st) != GCMI_OK) return nullptr;
  std::lock_guard<std::mutex> lk(st->mu);
  return st->stream;
}

hipStream_t default_stream() {
  DeviceState* st = nullptr;
  if (current_state(&st) != GCMI_OK) return nullptr;
  std::lock_guard<std::mutex> lk(st->mu);
  if (!st->stream) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess)
      st->stream = s;
  }
  return st->stream;
}

int ScratchLease::acquire(hipStream_t stream, int count, int32_t** out) {
  DeviceState* st = nullptr;
  GCMI_TRY(current_state(&st));
  lk_ = std::unique_lock<std::mutex>(st->mu);
  StreamScratch* e = nullptr;
  for (StreamScratch& x : st->scratch)
    if (x.stream == stream) e = &x;
  if (!e) {
    st->scratch.push_back(StreamScratch{stream, nullptr, 0});
    e = &st->scratch.back();
  }
  if (e->count < count) {
    if (e->buf) {
      // kernels enqueued on this stream earlier may still read the buffer
      GCMI_HIP(hipStreamSynchronize(stream));
      GCMI_HIP(hipFree(e->buf));
      e->buf = nullptr;
      e->count = 0;
    }
    const int want = count < 4096 ? 4096 : count;
    if (hipMalloc(&e->buf, sizeof(int32_t) * (size_t)want) != hipSuccess) {
      (void)hipGetLastError();
      e->buf = nullptr;
      set_error("scratch allocation of %d ints failed", want);
      return GCMI_ENOMEM;
    }
    e->count = want;
  }
  *out = e->buf;
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

int gcmi_host_alloc(void** h_out, size_t bytes) {
  if (!h_out) { set_error("null out"); return GCMI_EINVAL; }
  GCMI_TRY(require_device());
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipHostMalloc(%zu) failed", bytes);
    return GCMI_ENOMEM;
  }
  *h_out = p;
  return GCMI_OK;
}

int gcmi_host_free(void* h_ptr) {
  if (!h_ptr) return GCMI_OK;
  GCMI_TRY(require_device());
  GCMI_HIP(hipHostFree(h_ptr));
  return GCMI_OK;
}

int gcmi_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes) {
  GCMI_TRY(require_device());
  if (!bytes) return GCMI_OK;
  if (!d_dst || !h_src) { set_error("null pointer"); return GCMI_EINVAL; }
  // ordered after the primitives enqueued on the library's stream, if any;
  // a synchronous copy otherwise
  hipStream_t s = existing_default_stream();
  if (s) {
    GCMI_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
    GCMI_HIP(hipStreamSynchronize(s));
  } else {
    GCMI_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  }
  return GCMI_OK;
}

int gcmi_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes) {
  GCMI_TRY(require_device());
  if (!bytes) return GCMI_OK;
  if (!h_dst || !d_src) { set_error("null pointer"); return GCMI_EINVAL; }
  // ordered after the primitives enqueued on the library's stream, if any;
  // a synchronous copy otherwise
  hipStream_t s = existing_default_stream();
  if (s) {
    GCMI_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
    GCMI_HIP(hipStreamSynchronize(s));
  } else {
    GCMI_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  }
  return GCMI_OK;
}

int gcmi_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes) {
  GCMI_TRY(require_device());
  if (!bytes) return GCMI_OK;
  if (!d_dst || !d_src) { set_error("null pointer"); return GCMI_EINVAL; }
  // ordered after the primitives enqueued on the library's stream, if any;
  // a synchronous copy otherwise
  hipStream_t s = existing_default_stream();
  if (s) {
    GCMI_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, s));
    GCMI_HIP(hipStreamSynchronize(s));
  } else {
    GCMI_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
  }
  return GCMI_OK;
}

int gcmi_device_sync(void) {
  GCMI_TRY(require_device());
  GCMI_HIP(hipDeviceSynchronize());
  return GCMI_OK;
}

int gcmi_stream_sync(void* stream) {
  GCMI_TRY(require_device());
  // NULL: the library's stream, if it exists (nothing was enqueued on it
  // otherwise)
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : existing_default_stream();
  if (s) GCMI_HIP(hipStreamSynchronize(s));
  return GCMI_OK;
}

}  // extern "C"

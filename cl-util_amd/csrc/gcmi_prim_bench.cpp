// Native micro-benchmark of the tile-list primitives through the C ABI:
// exclusive scan, stream compaction and the fused tile-list build over the
// tile counts of a sweep of image sizes and tile sizes (the sweep of the
// reference's PrefixSum.cpp:230-335: 11 image sizes x tiles 16x16, 32x32,
// scanning a list of ones).  Every result is checked on the host.
//
//   gcmi_prim_bench [iterations]      (default 100)
//
// Prints CSV: img dim, mp, tile dim, num tiles, then per primitive the mean
// wall time per call in ms, call and completion (the calls only enqueue;
// each is followed by a wait for its stream inside the timed region, as the
// reference's .wait()-ed launches).  Exit status 0 if every check
// passed, 1 on a wrong result, 3 without a usable GPU.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "../../include/gcmi.h"

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

#define CHECK(call)                                                           \
  do {                                                                        \
    int _rc = (call);                                                         \
    if (_rc != GCMI_OK) {                                                     \
      fprintf(stderr, "%s failed (%d): %s\n", #call, _rc, gcmi_last_error()); \
      return _rc == GCMI_ENODEVICE ? 3 : 1;                                   \
    }                                                                         \
  } while (0)

struct Dim {
  int w, h;
};

int run_case(Dim img, Dim tile, int iters) {
  const int n = (img.w / tile.w) * (img.h / tile.h);
  std::vector<int32_t> ones(n, 1), stamps(n), host(n);
  for (int i = 0; i < n; ++i) stamps[i] = (int)((i * 2654435761u) % 20u);  // stamps in [0, 20)
  int32_t *d_a = nullptr, *d_total = nullptr, *d_flags = nullptr, *d_list = nullptr;
  CHECK(gcmi_malloc((void**)&d_a, (size_t)n * 4));
  CHECK(gcmi_malloc((void**)&d_total, 4));
  CHECK(gcmi_malloc((void**)&d_flags, (size_t)n * 4));
  CHECK(gcmi_malloc((void**)&d_list, (size_t)n * 4));
  CHECK(gcmi_memcpy_h2d(d_flags, stamps.data(), (size_t)n * 4));

  // scan of ones, re-uploaded before each call (the scan is in place)
  double t_scan = 0;
  for (int it = 0; it < iters; ++it) {
    CHECK(gcmi_memcpy_h2d(d_a, ones.data(), (size_t)n * 4));
    const double t0 = now_ms();
    CHECK(gcmi_prefix_sum(d_a, d_total, n, nullptr));
    CHECK(gcmi_stream_sync(nullptr));
    t_scan += now_ms() - t0;
  }
  int32_t total = -1;
  CHECK(gcmi_memcpy_d2h(&total, d_total, 4));
  CHECK(gcmi_memcpy_d2h(host.data(), d_a, (size_t)n * 4));
  bool ok = total == n;
  for (int i = 0; i < n && ok; ++i) ok = host[i] == i;

  // compaction of nonzero stamps
  double t_compact = 0;
  int32_t len = -1;
  for (int it = 0; it < iters; ++it) {
    const double t0 = now_ms();
    CHECK(gcmi_stream_compact(d_flags, d_list, d_total, n, nullptr));
    CHECK(gcmi_stream_sync(nullptr));
    t_compact += now_ms() - t0;
  }
  CHECK(gcmi_memcpy_d2h(&len, d_total, 4));
  CHECK(gcmi_memcpy_d2h(host.data(), d_list, (size_t)n * 4));
  int k = 0;
  for (int i = 0; i < n && ok; ++i)
    if (stamps[i] != 0) ok = k < len && host[k++] == i;
  ok = ok && k == len;

  // tile-list build: stamps >= 10 (the reference test's GTE 10)
  double t_build = 0;
  for (int it = 0; it < iters; ++it) {
    const double t0 = now_ms();
    CHECK(gcmi_tiles_build(d_flags, n, GCMI_OP_GTE, 10, d_list, d_total, nullptr));
    CHECK(gcmi_stream_sync(nullptr));
    t_build += now_ms() - t0;
  }
  CHECK(gcmi_memcpy_d2h(&len, d_total, 4));
  CHECK(gcmi_memcpy_d2h(host.data(), d_list, (size_t)n * 4));
  k = 0;
  for (int i = 0; i < n && ok; ++i)
    if (stamps[i] >= 10) ok = k < len && host[k++] == i;
  ok = ok && k == len;

  gcmi_free(d_a);
  gcmi_free(d_total);
  gcmi_free(d_flags);
  gcmi_free(d_list);
  printf("(%dx%d),%.3f,(%dx%d),%d,%.4f,%.4f,%.4f,%s\n", img.w, img.h,
         img.w * (double)img.h / 1e6, tile.w, tile.h, n, t_scan / iters, t_compact / iters,
         t_build / iters, ok ? "ok" : "WRONG");
  return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 100;
  if (iters < 1) {
    fprintf(stderr, "usage: %s [iterations >= 1]\n", argv[0]);
    return 2;
  }
  int ndev = 0;
  if (gcmi_device_count(&ndev) != GCMI_OK || ndev < 1) {
    fprintf(stderr, "no HIP device: %s\n", gcmi_last_error());
    return 3;
  }
  const Dim imgs[] = {{400, 300},   {640, 480},   {800, 600},   {1024, 768},
                      {1600, 1200}, {1920, 1080}, {2560, 1440}, {2048, 2048},
                      {3600, 2400}, {4096, 4096}, {8192, 8192}};
  const Dim tiles[] = {{16, 16}, {32, 32}};
  printf("img dim,mp,tile dim,num tiles,scan ms,compact ms,build ms,check\n");
  int rc = 0;
  for (const Dim& im : imgs)
    for (const Dim& t : tiles) {
      const int r = run_case(im, t, iters);
      if (r == 3) return 3;
      if (r) rc = 1;
    }
  return rc;
}

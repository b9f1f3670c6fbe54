// GrowCut (Vezhnevets & Konouchine's cellular automaton) on the tile-list
// runtime, and the tile-list visualisation filter.
//
// The reference (GrowCut.py:26-153, growcut.cl:28-179) keeps a label and a
// strength per pixel, double-buffered, and an IncrementalTileList of 16x16
// tiles stamped with the iteration in which they (or a neighbour's edge)
// changed.  An evolve step lists the tiles stamped with the current
// iteration, bumps the iteration and updates those tiles synchronously:
// each pixel is attacked by its four neighbours (up, down, left, right, in
// that order), an attack g(|c_p - c_q|) * strength_q that beats the current
// defence takes over the neighbour's label and strength; a tile in which
// something changed stamps itself, and a changed edge pixel stamps the tile
// across the edge.  A tile not listed holds the same state in both buffers
// (it was not changed by the last step it ran in, or it was written since),
// so updating only listed tiles is the synchronous update of the whole
// image.  Here a block is one listed tile, its halo is staged in LDS, and
// the list is built by the library's fused flag + compaction; the host
// checks the list's length every few steps instead of after every one.
//
// Arithmetic is float32 in a fixed order (contraction off): colours c / 255
// per channel, distance sqrt(((dr^2 + dg^2) + db^2) + da^2), g(X) one of
// the weights below with X / sqrt(3), attack = g * strength.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>

#include "gcmi_common.h"

#pragma clang fp contract(off)

namespace gcmi {

constexpr int kGcT = 16;  // tile edge (TILEW, TILEH of GrowCut.py:18-19)
constexpr float kSqrt3 = 1.7320508075688772f;

__device__ __forceinline__ float gc_weight(int kind, float X) {
  const float x = X / kSqrt3;
  switch (kind) {
    case GCMI_GROWCUT_WEIGHT_POW2: return 1.0f - x * x;
    case GCMI_GROWCUT_WEIGHT_POW3: return 1.0f - (x * x) * x;
    case GCMI_GROWCUT_WEIGHT_POW1_5: return 1.0f - x * sqrtf(x);
    case GCMI_GROWCUT_WEIGHT_SQRT: return 1.0f - sqrtf(x);
    default: return 1.0f - x;  // GCMI_GROWCUT_WEIGHT_LINEAR
  }
}

__device__ __forceinline__ float4 gc_colour(const uchar4* img, int i) {
  const uchar4 c = img[i];
  return make_float4((float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f,
                     (float)c.w / 255.0f);
}

__device__ __forceinline__ float gc_distance(float4 a, float4 b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z, dw = a.w - b.w;
  return sqrtf(((dx * dx + dy * dy) + dz * dz) + dw * dw);
}

// One step over the listed tiles (*count of them): labels / strengths from
// the `in` buffers to the `out` buffers; tiles that change, and tiles
// across a changed edge pixel, get stamp `iteration`.
__global__ void __launch_bounds__(kGcT * kGcT)
k_growcut_step(const int32_t* __restrict__ list, const int32_t* __restrict__ count, int W, int H,
               int tw, const uchar4* __restrict__ img, const uint8_t* __restrict__ lab_in,
               uint8_t* lab_out, const float* __restrict__ str_in, float* str_out,
               int32_t* tiles, int32_t iteration, int weight) {
  constexpr int SW = kGcT + 2;
  __shared__ float s_str[SW * SW];
  __shared__ uint8_t s_lab[SW * SW];
  __shared__ int s_flag[5];
  const int n = *count;
  const int lx = threadIdx.x % kGcT, ly = threadIdx.x / kGcT;
  for (int pos = blockIdx.x; pos < n; pos += gridDim.x) {
    const int t = list[pos];
    const int tx = t % tw, ty = t / tw;
    const int ix = tx * kGcT + lx, iy = ty * kGcT + ly;
    const bool in = ix < W && iy < H;
    const int i = iy * W + ix;
    const int sxy = (ly + 1) * SW + lx + 1;
    if (threadIdx.x < 5) s_flag[threadIdx.x] = 0;
    // the tile and its halo; pixels outside the image have strength 0, so
    // their attack never wins
    s_str[sxy] = in ? str_in[i] : 0.0f;
    s_lab[sxy] = in ? lab_in[i] : 0;
    if (ly == 0) {
      const bool ok = in && iy > 0;
      s_str[sxy - SW] = ok ? str_in[i - W] : 0.0f;
      s_lab[sxy - SW] = ok ? lab_in[i - W] : 0;
    }
    if (ly == kGcT - 1) {
      const bool ok = in && iy + 1 < H;
      s_str[sxy + SW] = ok ? str_in[i + W] : 0.0f;
      s_lab[sxy + SW] = ok ? lab_in[i + W] : 0;
    }
    if (lx == 0) {
      const bool ok = in && ix > 0;
      s_str[sxy - 1] = ok ? str_in[i - 1] : 0.0f;
      s_lab[sxy - 1] = ok ? lab_in[i - 1] : 0;
    }
    if (lx == kGcT - 1) {
      const bool ok = in && ix + 1 < W;
      s_str[sxy + 1] = ok ? str_in[i + 1] : 0.0f;
      s_lab[sxy + 1] = ok ? lab_in[i + 1] : 0;
    }
    __syncthreads();
    if (in) {
      const float4 c = gc_colour(img, i);
      uint8_t label = s_lab[sxy];
      float defence = s_str[sxy];
      const float d0 = defence;
      const uint8_t l0 = label;
      // up, down, left, right (growcut.cl:151-177)
      const int nb[4] = {iy > 0 ? i - W : -1, iy + 1 < H ? i + W : -1, ix > 0 ? i - 1 : -1,
                         ix + 1 < W ? i + 1 : -1};
      const int so[4] = {-SW, SW, -1, 1};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (nb[k] < 0) continue;
        const float a = gc_weight(weight, gc_distance(c, gc_colour(img, nb[k]))) * s_str[sxy + so[k]];
        if (a > defence) {
          defence = a;
          label = s_lab[sxy + so[k]];
        }
      }
      str_out[i] = defence;
      lab_out[i] = label;
      if (defence != d0 || label != l0) {
        s_flag[0] = 1;
        if (ly == 0 && iy > 0) s_flag[1] = 1;
        if (ly == kGcT - 1 && iy + 1 < H) s_flag[2] = 1;
        if (lx == 0 && ix > 0) s_flag[3] = 1;
        if (lx == kGcT - 1 && ix + 1 < W) s_flag[4] = 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (s_flag[0]) tiles[t] = iteration;
      if (s_flag[1]) tiles[t - tw] = iteration;
      if (s_flag[2]) tiles[t + tw] = iteration;
      if (s_flag[3]) tiles[t - 1] = iteration;
      if (s_flag[4]) tiles[t + 1] = iteration;
    }
    __syncthreads();  // s_flag and the LDS tile are rewritten for the next tile
  }
}

// GrowCut.label (growcut.cl:28-48): seed pixels get the label and strength
// 1 in both buffers, and their tiles the current iteration -- and, unlike
// the reference, so do the tiles across an edge a seed lies on, since the
// seed can attack there in the next step (without them the neighbour tile
// would lag a step behind the synchronous update).
__global__ void k_growcut_label(const int32_t* __restrict__ pts, int n, uint8_t label, int W,
                                int H, int tw, uint8_t* lab0, uint8_t* lab1, float* str0,
                                float* str1, int32_t* tiles, int32_t iteration) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = pts[k];
  lab0[i] = label;
  lab1[i] = label;
  str0[i] = 1.0f;
  str1[i] = 1.0f;
  const int x = i % W, y = i / W;
  const int t = (y / kGcT) * tw + x / kGcT;
  tiles[t] = iteration;
  if (y % kGcT == 0 && y > 0) tiles[t - tw] = iteration;
  if (y % kGcT == kGcT - 1 && y + 1 < H) tiles[t + tw] = iteration;
  if (x % kGcT == 0 && x > 0) tiles[t - 1] = iteration;
  if (x % kGcT == kGcT - 1 && x + 1 < W) tiles[t + 1] = iteration;
}

__global__ void k_growcut_rgba(const uint8_t* __restrict__ src, int ch, long long n,
                               uchar4* dst) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const uint8_t* p = src + i * ch;
    dst[i] = make_uchar4(p[0], p[1], p[2], ch == 4 ? p[3] : 0);
  }
}

// IncrementalTileList's filter (incremental_tile_list.cl:156-175): opaque
// red (0xFF0000FF) where the pixel's tile holds `iteration`, opaque black
// elsewhere.
__global__ void k_tiles_filter(const int32_t* __restrict__ tiles, int tiles_w, int tiles_h,
                               int tile_w, int tile_h, int W, int H, int32_t iteration,
                               uint32_t* out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= W || y >= H) return;
  const int tx = x / tile_w, ty = y / tile_h;
  // (pixels beyond the last whole tile have no tile: black)
  const bool on = tx < tiles_w && ty < tiles_h && tiles[ty * tiles_w + tx] == iteration;
  out[(size_t)y * W + x] = on ? 0xFF0000FFu : 0xFF000000u;
}

}  // namespace gcmi

using namespace gcmi;

struct gcmi_growcut {
  int W = 0, H = 0, tw = 0, th = 0, ntiles = 0;
  int weight = GCMI_GROWCUT_WEIGHT_LINEAR;
  int cur = 0;            // buffer holding the current state
  int32_t iteration = 0;  // the IncrementalTileList's iteration
  hipStream_t stream = nullptr;
  uchar4* img = nullptr;
  uint8_t* lab[2] = {nullptr, nullptr};
  float* str[2] = {nullptr, nullptr};
  int32_t *tiles = nullptr, *list = nullptr, *count = nullptr, *scratch = nullptr;
  int32_t* h_count = nullptr;  // pinned
  int32_t* d_pts = nullptr;
  int pts_cap = 0;
  int64_t steps = 0;
};

namespace {

void gc_free(gcmi_growcut* g) {
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  void* d[] = {g->img, g->lab[0], g->lab[1], g->str[0], g->str[1], g->tiles, g->list,
               g->count, g->scratch, g->d_pts};
  for (void* p : d)
    if (p) (void)hipFree(p);
  if (g->h_count) (void)hipHostFree(g->h_count);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

// Lists the tiles stamped with the current iteration (count on the device).
int gc_build(gcmi_growcut* g) {
  Pred p{};
  p.t1 = g->tiles;
  p.kind = Pred::CMP;
  p.op1 = GCMI_OP_EQUAL;
  p.a = g->iteration;
  return compact_async(p, g->ntiles, g->list, g->count, nullptr, g->scratch, g->stream);
}

}  // namespace

extern "C" {

int gcmi_growcut_create(int width, int height, const uint8_t* h_rgb, int channels, int weight,
                        gcmi_growcut** out) {
  if (!out || width < 1 || height < 1 || !h_rgb || (channels != 3 && channels != 4) ||
      weight < GCMI_GROWCUT_WEIGHT_LINEAR || weight > GCMI_GROWCUT_WEIGHT_SQRT ||
      (long long)width * height >= (1LL << 31)) {
    set_error("bad GrowCut arguments");
    return GCMI_EINVAL;
  }
  *out = nullptr;
  GCMI_TRY(require_device());
  gcmi_growcut* g = new (std::nothrow) gcmi_growcut();
  if (!g) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  g->W = width;
  g->H = height;
  g->tw = (width + kGcT - 1) / kGcT;
  g->th = (height + kGcT - 1) / kGcT;
  g->ntiles = g->tw * g->th;
  g->weight = weight;
  const size_t npx = (size_t)width * height;
  uint8_t* d_src = nullptr;
  bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc((void**)&g->img, npx * 4) == hipSuccess &&
            hipMalloc((void**)&g->lab[0], npx) == hipSuccess &&
            hipMalloc((void**)&g->lab[1], npx) == hipSuccess &&
            hipMalloc((void**)&g->str[0], npx * 4) == hipSuccess &&
            hipMalloc((void**)&g->str[1], npx * 4) == hipSuccess &&
            hipMalloc((void**)&g->tiles, (size_t)g->ntiles * 4) == hipSuccess &&
            hipMalloc((void**)&g->list, (size_t)g->ntiles * 4) == hipSuccess &&
            hipMalloc((void**)&g->count, 4) == hipSuccess &&
            hipMalloc((void**)&g->scratch, (size_t)compact_scratch_ints(g->ntiles) * 4) ==
                hipSuccess &&
            hipHostMalloc((void**)&g->h_count, 4, hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void**)&d_src, npx * channels) == hipSuccess;
  if (ok) {
    // labels 0, strengths 0, stamps -1 (IncrementalTileList's init_iteration)
    ok = hipMemcpyAsync(d_src, h_rgb, npx * channels, hipMemcpyHostToDevice, g->stream) ==
             hipSuccess &&
         hipMemsetAsync(g->lab[0], 0, npx, g->stream) == hipSuccess &&
         hipMemsetAsync(g->lab[1], 0, npx, g->stream) == hipSuccess &&
         hipMemsetAsync(g->str[0], 0, npx * 4, g->stream) == hipSuccess &&
         hipMemsetAsync(g->str[1], 0, npx * 4, g->stream) == hipSuccess &&
         hipMemsetAsync(g->tiles, 0xFF, (size_t)g->ntiles * 4, g->stream) == hipSuccess;
    if (ok) {
      const long long n = (long long)npx;
      hipLaunchKernelGGL(k_growcut_rgba, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)),
                         dim3(256), 0, g->stream, d_src, channels, n, g->img);
      ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(g->stream) == hipSuccess;
    }
  }
  if (d_src) (void)hipFree(d_src);
  if (!ok) {
    (void)hipGetLastError();
    gc_free(g);
    set_error("GrowCut allocation or upload failed");
    return GCMI_ENOMEM;
  }
  *out = g;
  return GCMI_OK;
}

int gcmi_growcut_destroy(gcmi_growcut* g) {
  if (g) gc_free(g);
  return GCMI_OK;
}

int gcmi_growcut_label(gcmi_growcut* g, int n, const int32_t* h_points, int label) {
  if (!g || n < 0 || (n > 0 && !h_points) || label < 0 || label > 255) {
    set_error("bad GrowCut label arguments");
    return GCMI_EINVAL;
  }
  const long long npx = (long long)g->W * g->H;
  for (int k = 0; k < n; ++k)
    if (h_points[k] < 0 || h_points[k] >= npx) {
      set_error("point %d (pixel %d) outside the image", k, h_points[k]);
      return GCMI_EINVAL;
    }
  if (n == 0) return GCMI_OK;
  if (g->pts_cap < n) {
    GCMI_HIP(hipStreamSynchronize(g->stream));
    if (g->d_pts) (void)hipFree(g->d_pts);
    g->d_pts = nullptr;
    g->pts_cap = 0;
    if (hipMalloc((void**)&g->d_pts, (size_t)n * 4) != hipSuccess) {
      (void)hipGetLastError();
      set_error("allocation of %d points failed", n);
      return GCMI_ENOMEM;
    }
    g->pts_cap = n;
  }
  GCMI_HIP(hipMemcpyAsync(g->d_pts, h_points, (size_t)n * 4, hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(k_growcut_label, dim3((n + 255) / 256), dim3(256), 0, g->stream,
                     (const int32_t*)g->d_pts, n, (uint8_t)label, g->W, g->H, g->tw, g->lab[0],
                     g->lab[1], g->str[0], g->str[1], g->tiles, g->iteration);
  GCMI_HIP(hipGetLastError());
  GCMI_HIP(hipStreamSynchronize(g->stream));
  return GCMI_OK;
}

int gcmi_growcut_evolve(gcmi_growcut* g, int iterations, int* complete_out, int* steps_out) {
  if (!g || iterations < 0) {
    set_error("bad GrowCut evolve arguments");
    return GCMI_EINVAL;
  }
  constexpr int kCheckEvery = 8;  // steps between host reads of the list's length
  int done = 0, complete = 0;
  while (done < iterations && !complete) {
    const int batch = std::min(kCheckEvery, iterations - done);
    for (int b = 0; b < batch; ++b) {
      GCMI_TRY(gc_build(g));
      ++g->iteration;
      const int nxt = 1 - g->cur;
      hipLaunchKernelGGL(k_growcut_step, dim3((unsigned)std::min(g->ntiles, 8192)),
                         dim3(kGcT * kGcT), 0, g->stream, (const int32_t*)g->list,
                         (const int32_t*)g->count, g->W, g->H, g->tw, (const uchar4*)g->img,
                         (const uint8_t*)g->lab[g->cur], g->lab[nxt],
                         (const float*)g->str[g->cur], g->str[nxt], g->tiles, g->iteration,
                         g->weight);
      GCMI_HIP(hipGetLastError());
      g->cur = nxt;
    }
    // the last step's list length: 0 means that step changed nothing, and
    // so did nothing (GrowCut.evolve's isComplete)
    GCMI_HIP(hipMemcpyAsync(g->h_count, g->count, 4, hipMemcpyDeviceToHost, g->stream));
    GCMI_HIP(hipStreamSynchronize(g->stream));
    done += batch;
    if (*g->h_count == 0) complete = 1;
  }
  // a finished run's next list is empty too
  if (!complete && iterations == 0) {
    GCMI_TRY(gc_build(g));
    GCMI_HIP(hipMemcpyAsync(g->h_count, g->count, 4, hipMemcpyDeviceToHost, g->stream));
    GCMI_HIP(hipStreamSynchronize(g->stream));
    complete = *g->h_count == 0;
  }
  g->steps += done;
  if (complete_out) *complete_out = complete;
  if (steps_out) *steps_out = done;
  return GCMI_OK;
}

int gcmi_growcut_read(gcmi_growcut* g, uint8_t* h_labels, float* h_strength, int32_t* h_tiles,
                      int32_t* iteration_out) {
  if (!g) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  const size_t npx = (size_t)g->W * g->H;
  if (h_labels)
    GCMI_HIP(hipMemcpyAsync(h_labels, g->lab[g->cur], npx, hipMemcpyDeviceToHost, g->stream));
  if (h_strength)
    GCMI_HIP(hipMemcpyAsync(h_strength, g->str[g->cur], npx * 4, hipMemcpyDeviceToHost,
                            g->stream));
  if (h_tiles)
    GCMI_HIP(hipMemcpyAsync(h_tiles, g->tiles, (size_t)g->ntiles * 4, hipMemcpyDeviceToHost,
                            g->stream));
  GCMI_HIP(hipStreamSynchronize(g->stream));
  if (iteration_out) *iteration_out = g->iteration;
  return GCMI_OK;
}

int gcmi_tiles_filter(const int32_t* d_tiles, int tiles_w, int tiles_h, int tile_w, int tile_h,
                      int width, int height, int32_t iteration, uint32_t* d_out, void* stream) {
  GCMI_TRY(require_device());
  if (!d_tiles || !d_out || tiles_w < 1 || tiles_h < 1 || tile_w < 1 || tile_h < 1 || width < 0 ||
      height < 0) {
    set_error("bad tile filter arguments");
    return GCMI_EINVAL;
  }
  if (width == 0 || height == 0) return GCMI_OK;
  const dim3 block(16, 16);
  const dim3 grid((width + 15) / 16, (height + 15) / 16);
  hipLaunchKernelGGL(k_tiles_filter, grid, block, 0, pick_stream(stream), d_tiles, tiles_w,
                     tiles_h, tile_w, tile_h, width, height, iteration, d_out);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

}  // extern "C"

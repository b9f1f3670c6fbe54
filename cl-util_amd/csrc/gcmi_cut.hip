// C ABI of the plain cut (a pool of one image) and of batches (a pool of
// many images solved together).
#include <stdlib.h>
#include <string.h>

#include <new>

#include "gcmi_pool.h"

using namespace gcmi;

struct gcmi_cut {
  Pool pool;
  int img = -1;
};

struct gcmi_batch {
  Pool pool;
  int count = 0;
  int* ids = nullptr;  // image ids, in creation order
};

namespace {

int new_cut(int W, int H, int tile, gcmi_cut** out) {
  if (!out) {
    set_error("null out");
    return GCMI_EINVAL;
  }
  *out = nullptr;
  if (!pool_ops(tile)) {
    set_error("tile must be 16, 32 or 64 (got %d)", tile);
    return GCMI_EINVAL;
  }
  int tw = 0, th = 0;
  GCMI_TRY(tiles_for(W, H, tile, &tw, &th));
  GCMI_TRY(require_device());
  gcmi_cut* c = new (std::nothrow) gcmi_cut();
  if (!c) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  int rc = pool_init(&c->pool, tile, tw * th, (long long)W * H, 1, nullptr);
  if (rc == GCMI_OK) rc = pool_add_image(&c->pool, W, H, &c->img);
  if (rc != GCMI_OK) {
    pool_free(&c->pool);
    delete c;
    return rc;
  }
  *out = c;
  return GCMI_OK;
}

int check_batch_img(gcmi_batch* b, int i) {
  if (!b || i < 0 || i >= b->count) {
    set_error("no image %d in batch", i);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ cut

int gcmi_cut_create(int width, int height, int tile, gcmi_cut** out) {
  return new_cut(width, height, tile, out);
}

int gcmi_cut_destroy(gcmi_cut* c) {
  if (!c) return GCMI_OK;
  pool_free(&c->pool);
  delete c;
  return GCMI_OK;
}

int gcmi_cut_upload(gcmi_cut* c, const int32_t* h_source, const int32_t* h_sink,
                    const int32_t* h_right, const int32_t* h_down) {
  if (!c || !h_source || !h_sink || !h_right || !h_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_upload(&c->pool, c->img, h_source, h_sink, h_right, h_down, false);
}

void gcmi_rgb_model_default(gcmi_rgb_model* m) {
  if (!m) return;
  *m = gcmi_rgb_model{};
  const double bg[3] = {60.0, 80.0, 100.0}, fg[3] = {180.0, 150.0, 90.0};
  for (int k = 0; k < 3; ++k) {
    m->mu_bg[k] = bg[k];
    m->mu_fg[k] = fg[k];
  }
  m->unary_scale = 10.0;
  m->unary_div = 100.0;
  m->unary_cap = 100.0;
  m->pair_kind = GCMI_PAIR_EXP;
  m->pair_scale = 50.0;
  m->inv_beta = 0.1;
  m->inv_eps = 0.05;
}

int gcmi_cut_upload_rgb(gcmi_cut* c, const uint8_t* h_rgb, int channels,
                        const gcmi_rgb_model* model) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_upload_rgb_host(&c->pool, c->img, h_rgb, channels, model);
}

int gcmi_cut_upload_rgb_gmm(gcmi_cut* c, const uint8_t* h_rgb, int channels,
                            const gcmi_rgb_model* pairwise, const gcmi_gmm* fg,
                            const gcmi_gmm* bg, double unary_scale, double unary_cap) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_upload_rgb_gmm_host(&c->pool, c->img, h_rgb, channels, pairwise, fg, bg,
                                  unary_scale, unary_cap);
}

int gcmi_cut_capacities(gcmi_cut* c, int32_t* h_source, int32_t* h_sink, int32_t* h_right,
                        int32_t* h_down) {
  if (!c || !h_source || !h_sink || !h_right || !h_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_capacities(&c->pool, c->img, h_source, h_sink, h_right, h_down);
}

int gcmi_cut_upload_device(gcmi_cut* c, const int32_t* d_source, const int32_t* d_sink,
                           const int32_t* d_right, const int32_t* d_down) {
  if (!c || !d_source || !d_sink || !d_right || !d_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_upload(&c->pool, c->img, d_source, d_sink, d_right, d_down, true);
}

int gcmi_cut_edit(gcmi_cut* c, int n, const int32_t* h_index, const int32_t* h_source,
                  const int32_t* h_sink) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_edit(&c->pool, c->img, n, h_index, h_source, h_sink);
}

int gcmi_cut_paint(gcmi_cut* c, int npts, const double* h_xy, double width, int32_t source,
                   int32_t sink) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_paint(&c->pool, c->img, npts, h_xy, width, source, sink);
}

int gcmi_cut_labels_packed(gcmi_cut* c, uint32_t* h_bits) {
  if (!c || !h_bits) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_labels_packed(&c->pool, c->img, h_bits);
}

int gcmi_cut_forget(gcmi_cut* c) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_forget(&c->pool, c->img);
}

int gcmi_cut_run(gcmi_cut* c, int sweeps, int bfs, int64_t max_rounds, int64_t* flow_out) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  int64_t f = 0;
  GCMI_TRY(pool_solve(&c->pool, &c->img, 1, sweeps, bfs, max_rounds, &f));
  if (flow_out) *flow_out = f;
  return GCMI_OK;
}

int gcmi_cut_push(gcmi_cut* c, int sweeps) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_step_push(&c->pool, c->img, sweeps);
}

int gcmi_cut_relabel(gcmi_cut* c, int32_t* active_out) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_step_relabel(&c->pool, c->img, false, active_out);
}

int gcmi_cut_start_bfs(gcmi_cut* c, int32_t* active_out) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_step_relabel(&c->pool, c->img, true, active_out);
}

int gcmi_cut_is_completed(gcmi_cut* c, int* done_out, int64_t* flow_out) {
  if (!c || !done_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_step_completed(&c->pool, c->img, done_out, flow_out);
}

int gcmi_cut_tile_flags(gcmi_cut* c, int which, int32_t* d_out) {
  if (!c || !d_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  Pool* p = &c->pool;
  const int32_t* src = nullptr;
  switch (which) {
    case GCMI_TILES_ACTIVE: src = p->dev.act; break;
    case GCMI_TILES_BORDER: src = p->dev.pflag; break;
    case GCMI_TILES_CHANGED: src = p->dev.dirty; break;
    case GCMI_TILES_EXCESS: src = p->dev.haspos; break;
    default:
      set_error("unknown tile set %d", which);
      return GCMI_EINVAL;
  }
  const PoolImage& im = p->imgs[c->img];
  GCMI_HIP(hipMemcpyAsync(d_out, src + im.tile0, (size_t)im.ntiles * 4, hipMemcpyDeviceToDevice,
                          p->stream));
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

int gcmi_cut_labels(gcmi_cut* c, uint8_t* h_out) {
  if (!c || !h_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_labels(&c->pool, c->img, h_out, false);
}

int gcmi_cut_labels_device(gcmi_cut* c, uint8_t* d_out) {
  if (!c || !d_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_labels(&c->pool, c->img, d_out, true);
}

int gcmi_cut_stats(gcmi_cut* c, int64_t* out, int n) {
  if (!c || !out || n < 0) {
    set_error("bad argument");
    return GCMI_EINVAL;
  }
  for (int k = 0; k < n && k < STAT_N; ++k) out[k] = c->pool.stats[k];
  return GCMI_OK;
}

int gcmi_cut_heights(gcmi_cut* c, int32_t* h_out) {
  if (!c || !h_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_heights(&c->pool, c->img, h_out);
}

int gcmi_cut_field(gcmi_cut* c, int which, int32_t* h_out) {
  if (!c || !h_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_field(&c->pool, c->img, which, h_out);
}

int gcmi_cut_geometry(gcmi_cut* c, int* out3) {
  if (!c || !out3) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  const PoolImage& im = c->pool.imgs[c->img];
  out3[0] = im.tw * c->pool.T;
  out3[1] = im.th * c->pool.T;
  out3[2] = c->pool.T;
  return GCMI_OK;
}

int gcmi_cut_profile(gcmi_cut* c, int enable) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  c->pool.prof.on = enable != 0;
  c->pool.prof.reset();
  return GCMI_OK;
}

int gcmi_cut_set_option(gcmi_cut* c, int option, int value) {
  if (!c) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_set_option(&c->pool, option, value);
}

int gcmi_cut_kernel_times(gcmi_cut* c, double* ms_out, int64_t* launches_out, int n) {
  if (!c || n < 0) {
    set_error("bad argument");
    return GCMI_EINVAL;
  }
  for (int k = 0; k < n && k < K_N; ++k) {
    if (ms_out) ms_out[k] = c->pool.prof.ms[k];
    if (launches_out) launches_out[k] = c->pool.prof.launches[k];
  }
  return GCMI_OK;
}

// ---------------------------------------------------------------- batch

int gcmi_batch_create(int count, const int* widths, const int* heights, int tile,
                      gcmi_batch** out) {
  if (!out || count < 1 || !widths || !heights) {
    set_error("bad batch arguments");
    return GCMI_EINVAL;
  }
  *out = nullptr;
  if (!pool_ops(tile)) {
    set_error("tile must be 16, 32 or 64 (got %d)", tile);
    return GCMI_EINVAL;
  }
  long long tiles = 0, pix = 0;
  for (int i = 0; i < count; ++i) {
    int tw = 0, th = 0;
    GCMI_TRY(tiles_for(widths[i], heights[i], tile, &tw, &th));
    tiles += (long long)tw * th;
    pix += (long long)widths[i] * heights[i];
  }
  if (tiles * tile * tile >= (1LL << 30)) {
    set_error("batch of %lld padded pixels exceeds 2^30", tiles * tile * tile);
    return GCMI_EINVAL;
  }
  GCMI_TRY(require_device());
  gcmi_batch* b = new (std::nothrow) gcmi_batch();
  if (!b || !(b->ids = (int*)malloc(sizeof(int) * count))) {
    delete b;
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  b->count = count;
  int rc = pool_init(&b->pool, tile, (int)tiles, pix, count, nullptr);
  for (int i = 0; i < count && rc == GCMI_OK; ++i)
    rc = pool_add_image(&b->pool, widths[i], heights[i], &b->ids[i]);
  if (rc != GCMI_OK) {
    pool_free(&b->pool);
    free(b->ids);
    delete b;
    return rc;
  }
  *out = b;
  return GCMI_OK;
}

int gcmi_batch_destroy(gcmi_batch* b) {
  if (!b) return GCMI_OK;
  pool_free(&b->pool);
  free(b->ids);
  delete b;
  return GCMI_OK;
}

int gcmi_batch_upload(gcmi_batch* b, int i, const int32_t* h_source, const int32_t* h_sink,
                      const int32_t* h_right, const int32_t* h_down) {
  GCMI_TRY(check_batch_img(b, i));
  if (!h_source || !h_sink || !h_right || !h_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_upload(&b->pool, b->ids[i], h_source, h_sink, h_right, h_down, false);
}

int gcmi_batch_upload_rgb(gcmi_batch* b, int i, const uint8_t* h_rgb, int channels,
                          const gcmi_rgb_model* model) {
  GCMI_TRY(check_batch_img(b, i));
  return pool_upload_rgb_host(&b->pool, b->ids[i], h_rgb, channels, model);
}

int gcmi_batch_upload_rgb_gmm(gcmi_batch* b, int i, const uint8_t* h_rgb, int channels,
                              const gcmi_rgb_model* pairwise, const gcmi_gmm* fg,
                              const gcmi_gmm* bg, double unary_scale, double unary_cap) {
  GCMI_TRY(check_batch_img(b, i));
  return pool_upload_rgb_gmm_host(&b->pool, b->ids[i], h_rgb, channels, pairwise, fg, bg,
                                  unary_scale, unary_cap);
}

int gcmi_batch_upload_device(gcmi_batch* b, int i, const int32_t* d_source,
                             const int32_t* d_sink, const int32_t* d_right,
                             const int32_t* d_down) {
  GCMI_TRY(check_batch_img(b, i));
  if (!d_source || !d_sink || !d_right || !d_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_upload(&b->pool, b->ids[i], d_source, d_sink, d_right, d_down, true);
}

int gcmi_batch_edit(gcmi_batch* b, int i, int n, const int32_t* h_index,
                    const int32_t* h_source, const int32_t* h_sink) {
  GCMI_TRY(check_batch_img(b, i));
  return pool_edit(&b->pool, b->ids[i], n, h_index, h_source, h_sink);
}

int gcmi_batch_paint(gcmi_batch* b, int i, int npts, const double* h_xy, double width,
                     int32_t source, int32_t sink) {
  GCMI_TRY(check_batch_img(b, i));
  return pool_paint(&b->pool, b->ids[i], npts, h_xy, width, source, sink);
}

int gcmi_batch_run(gcmi_batch* b, int sweeps, int bfs, int64_t max_rounds, int64_t* flows_out) {
  if (!b) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_solve(&b->pool, b->ids, b->count, sweeps, bfs, max_rounds, flows_out);
}

int gcmi_batch_run_subset(gcmi_batch* b, int n, const int* which, int sweeps, int bfs,
                          int64_t max_rounds, int64_t* flows_out) {
  if (!b || n < 1 || !which) {
    set_error("bad subset");
    return GCMI_EINVAL;
  }
  int* ids = (int*)malloc(sizeof(int) * n);
  if (!ids) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  int rc = GCMI_OK;
  for (int k = 0; k < n && rc == GCMI_OK; ++k) {
    rc = check_batch_img(b, which[k]);
    if (rc == GCMI_OK) ids[k] = b->ids[which[k]];
  }
  if (rc == GCMI_OK) rc = pool_solve(&b->pool, ids, n, sweeps, bfs, max_rounds, flows_out);
  free(ids);
  return rc;
}

int gcmi_batch_labels(gcmi_batch* b, int i, uint8_t* h_out) {
  GCMI_TRY(check_batch_img(b, i));
  if (!h_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  return pool_labels(&b->pool, b->ids[i], h_out, false);
}

int gcmi_batch_stats(gcmi_batch* b, int64_t* out, int n) {
  if (!b || !out || n < 0) {
    set_error("bad argument");
    return GCMI_EINVAL;
  }
  for (int k = 0; k < n && k < STAT_N; ++k) out[k] = b->pool.stats[k];
  return GCMI_OK;
}

int gcmi_batch_profile(gcmi_batch* b, int enable) {
  if (!b) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  b->pool.prof.on = enable != 0;
  b->pool.prof.reset();
  return GCMI_OK;
}

int gcmi_batch_set_option(gcmi_batch* b, int option, int value) {
  if (!b) {
    set_error("null handle");
    return GCMI_EINVAL;
  }
  return pool_set_option(&b->pool, option, value);
}

int gcmi_batch_kernel_times(gcmi_batch* b, double* ms_out, int64_t* launches_out, int n) {
  if (!b || n < 0) {
    set_error("bad argument");
    return GCMI_EINVAL;
  }
  for (int k = 0; k < n && k < K_N; ++k) {
    if (ms_out) ms_out[k] = b->pool.prof.ms[k];
    if (launches_out) launches_out[k] = b->pool.prof.launches[k];
  }
  return GCMI_OK;
}

}  // extern "C"

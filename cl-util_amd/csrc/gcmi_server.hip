// Serving loop: resident sessions in one image pool, a request queue and a
// worker thread that batches continuously.
//
// A session is one image of the pool: its capacities, residual graph,
// heights and tile state stay in HBM between requests.  Requests are either
// a new image (full upload; the next solve starts from it) or a brush
// stroke (new unaries for some pixels; the next solve reflows warm from the
// session's residual graph).  The worker takes every queued request it can
// -- at most one per session and `max_batch` in all, oldest first --
// applies them, solves those sessions together in one pool solve, packs
// each mask to bits into the session's pinned host buffer and completes the
// requests.  Requests that arrive while a solve runs wait for the next one.
//
// New images reach the device while the worker solves: the submitting
// thread copies the capacities into a device staging range on a copy stream
// of its own, and the worker's upload is then a device-to-device copy.
//
// Admission: a session opens only if the pool, sized from the HBM budget,
// has a slot and room for its tiles and pixels.
//
// In plain mode the worker serves one request at a time and cuts it from
// scratch: a stroke rewrites the unaries on the device and the image is cut
// from its capacities, as a client without resident solver state would.
// Cold mode batches like the warm mode but cuts every session from scratch.
//
// No reference counterpart: the reference is a single-user tool
// (test_GraphCut.py:74-75 re-uploads the excess image per stroke).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gcmi_pool.h"

using namespace gcmi;

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

enum {
  REQ_IMAGE = GCMI_REQ_IMAGE,
  REQ_STROKE = GCMI_REQ_STROKE,
  REQ_POLYLINE = GCMI_REQ_POLYLINE,
  REQ_RGB = GCMI_REQ_RGB
};
enum { ST_QUEUED = 0, ST_DONE = 1, ST_FAILED = 2 };

constexpr long long kStageAlign = 64;  // staging offsets in int32 elements

struct Request {
  int64_t id = 0;
  int session = 0;
  int kind = REQ_STROKE;
  int n = 0;                    // stroke pixels
  std::vector<int32_t> stroke;  // idx, cs, ct
  long long stage = -1;         // image: offset of cs, ct, right, down in staging
  long long stage_len = 0;
  int32_t max_pair = 0;         // image: largest pairwise weight
  std::vector<double> xy;       // polyline: points
  double width = 0;             // polyline
  int32_t pcs = 0, pct = 0;     // polyline: unaries painted
  gcmi_rgb_model model{};       // colour image: the capacity model (n: channels)
  int64_t t_arrive = 0, t_start = 0, t_done = 0;
  int64_t t_queued = 0;  // when it entered the queue (after staging)
  int64_t flow = 0;
  int status = ST_QUEUED;
  int batch = 0;  // sessions solved together with it
};

struct Session {
  bool open = false;
  bool closing = false;      // close() waits for its requests; no new ones
  bool busy = false;         // a request of this session is being served
  int img = -1;
  int W = 0, H = 0;
  int32_t max_pair = -1;     // largest pairwise weight of the submitted image
  uint32_t* bits = nullptr;  // pinned packed mask of the last completed request
  int64_t last_done = 0;     // id of the last completed request
  int submits = 0;           // submissions past their session check, not yet queued
  int lane = 0;              // the lane (pool) holding the session
};

}  // namespace

// One mask copy of a solved batch: the session's packed bits on the device
// to its host buffer.
struct MaskCopy {
  void* dst;
  const void* src;
  size_t bytes;
};

// A solved batch whose masks are still to be copied out.
struct Pending {
  std::vector<Request*> batch;
  std::vector<MaskCopy> copies;
  hipEvent_t solved = nullptr;  // recorded on the lane's stream after the solve
  hipEvent_t done = nullptr;    // recorded on out_stream after the last copy
};

// A lane: one pool of sessions with its own stream and worker thread.
// Lanes solve their batches concurrently, so one batch's chain of small
// dependent launches overlaps another's on the device.  A session lives in
// one lane for its whole life.
constexpr int kMaxLanes = 4;
struct Lane {
  Pool pool;
  std::mutex pool_mu;  // held by the lane's worker while it serves a batch
  std::thread worker;
  int64_t ns_apply = 0, ns_solve = 0, ns_out = 0;  // worker time per phase (under mu)
};

struct gcmi_server {
  Lane lanes[kMaxLanes];
  int nlanes = 1;
  int device = 0;
  int mode = 0;  // GCMI_SERVE_WARM, GCMI_SERVE_PLAIN or GCMI_SERVE_COLD
  int max_batch = 64;
  int sweeps = 0, bfs = 0;
  gcmi_rgb_model rgb_model{};  // of later colour-image requests
  std::vector<Session> sessions;
  int32_t* d_stage = nullptr;  // device staging for new images
  RangeAlloc stage_alloc;
  size_t stage_bytes = 0;
  hipStream_t copy_stream = nullptr;
  // Masks are copied out on out_stream, after the solve that produced them:
  // the finisher thread issues the copies of every solved batch and
  // completes each batch's requests when its copies are done, while the
  // lanes already run their next batches.
  hipStream_t out_stream = nullptr;
  std::deque<Pending> outq;
  std::vector<hipEvent_t> free_events;
  bool stop_out = false;
  std::thread finisher;
  // Lock order: the lanes' pool_mu (in lane order), then copy_mu, then mu.
  // A lane's worker holds its pool_mu (not mu) while it serves a batch;
  // open/close hold pool_mu while they change a pool; submitters hold
  // copy_mu while they stage an image.
  std::mutex copy_mu, mu;
  std::condition_variable cv_work, cv_done, cv_out;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, Request*> reqs;
  int64_t next_id = 1;
  bool stop = false;
  int64_t iterations = 0, solved_requests = 0;
  char err[256] = "";
};

namespace {

// Applies the batch's requests to the lane's pool, solves, records `solved`
// on the lane's stream after the solve and lists the masks' copies for the
// finisher.  Runs on the lane's worker thread with its pool_mu held.  The
// batch's sessions stay busy until the finisher has completed them, so no
// later solve rewrites their labels before the copies have read them.
// ns[0..2] receive the time spent applying, solving and listing.
int serve_batch(gcmi_server* s, Lane& L, std::vector<Request*>& batch, hipEvent_t solved,
                std::vector<MaskCopy>& copies, int64_t* ns) {
  const int64_t t0 = now_ns();
  Pool* p = &L.pool;
  std::vector<int> imgs;
  imgs.reserve(batch.size());
  std::vector<EditReq> edits;
  std::vector<PaintReq> paints;
  std::vector<Request*> images;
  for (Request* r : batch) {
    const Session& ss = s->sessions[r->session];
    PoolImage& im = p->imgs[ss.img];
    if (r->kind == REQ_IMAGE || r->kind == REQ_RGB) {
      images.push_back(r);
    } else {
      // plain and cold modes cut every request from scratch: the edit only
      // rewrites the unaries
      if (s->mode != GCMI_SERVE_WARM) im.fresh = true;
      if (r->kind == REQ_POLYLINE) {
        paints.push_back(PaintReq{ss.img, r->n, r->xy.data(), r->width, r->pcs, r->pct});
      } else {
        const int32_t* d = r->stroke.data();
        edits.push_back(EditReq{ss.img, r->n, d, d + r->n, d + 2 * r->n});
      }
    }
    imgs.push_back(ss.img);
  }
  // strokes first (their staging waits for the stream, idle until now), one
  // launch for all polylines; then the new images (one request per session,
  // so no image is both edited and replaced)
  if (!paints.empty()) GCMI_TRY(pool_paint_many(p, (int)paints.size(), paints.data()));
  if (!edits.empty()) GCMI_TRY(pool_edit_many(p, (int)edits.size(), edits.data()));
  std::vector<RgbReq> rgbs;
  for (Request* r : images) {
    const Session& ss = s->sessions[r->session];
    const size_t npx = (size_t)ss.W * ss.H;
    if (r->kind == REQ_IMAGE) {
      const int32_t* d = s->d_stage + r->stage;
      GCMI_TRY(pool_upload(p, ss.img, d, d + npx, d + 2 * npx, d + 3 * npx, true));
      p->imgs[ss.img].max_pair = r->max_pair;
    } else {
      rgbs.push_back(RgbReq{ss.img, (const uint8_t*)(s->d_stage + r->stage), r->n, &r->model});
    }
  }
  // every colour image of the batch in one set of launches
  if (!rgbs.empty()) GCMI_TRY(pool_upload_rgb_many(p, (int)rgbs.size(), rgbs.data()));
  const int64_t t1 = now_ns();
  GCMI_TRY(pool_solve(p, imgs.data(), (int)imgs.size(), s->sweeps, s->bfs, 0, nullptr));
  const int64_t t2 = now_ns();
  // (the solve has synchronised: the flows are on the host) the masks,
  // kept packed on the device by the solve, are copied out after it by the
  // finisher
  GCMI_HIP(hipEventRecord(solved, p->stream));
  copies.clear();
  for (Request* r : batch) {
    const Session& ss = s->sessions[r->session];
    const long long words = ((long long)ss.W * ss.H + 31) / 32;
    copies.push_back(MaskCopy{ss.bits, packed_mask(p, ss.img), (size_t)words * 4});
  }
  for (Request* r : batch) r->flow = p->imgs[s->sessions[r->session].img].flow;
  ns[0] = t1 - t0;
  ns[1] = t2 - t1;
  ns[2] = now_ns() - t2;
  return GCMI_OK;
}

// Completes the requests of a batch (called with mu held).
void complete(gcmi_server* s, std::vector<Request*>& batch, bool ok, int64_t t) {
  for (Request* r : batch) {
    r->t_done = t;
    r->status = ok ? ST_DONE : ST_FAILED;
    r->batch = (int)batch.size();
    r->stroke.clear();
    r->stroke.shrink_to_fit();
    Session& ss = s->sessions[r->session];
    ss.busy = false;
    if (ok) ss.last_done = r->id;
  }
  s->iterations++;
  s->solved_requests += (int64_t)batch.size();
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

// Takes every solved batch queued so far, issues all their mask copies on
// out_stream (each batch behind its solve's event), then completes the
// batches in order as their copies finish.
void finisher_main(gcmi_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_out.wait(lk, [&] { return s->stop_out || !s->outq.empty(); });
    if (s->outq.empty()) break;  // stopping, and nothing left to finish
    std::vector<Pending> work;
    while (!s->outq.empty()) {
      work.push_back(std::move(s->outq.front()));
      s->outq.pop_front();
    }
    lk.unlock();
    std::vector<hipError_t> errs(work.size(), hipSuccess);
    for (size_t k = 0; k < work.size(); ++k) {
      const Pending& pd = work[k];
      hipError_t e = hipStreamWaitEvent(s->out_stream, pd.solved, 0);
      for (const MaskCopy& c : pd.copies)
        if (e == hipSuccess)
          e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s->out_stream);
      if (e == hipSuccess) e = hipEventRecord(pd.done, s->out_stream);
      errs[k] = e;
    }
    for (size_t k = 0; k < work.size(); ++k) {
      Pending& pd = work[k];
      hipError_t e = errs[k];
      if (e == hipSuccess) e = hipEventSynchronize(pd.done);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s->out_stream);
      }
      const int64_t t = now_ns();
      lk.lock();
      if (e != hipSuccess)
        snprintf(s->err, sizeof s->err, "mask copy failed: %s", hipGetErrorString(e));
      s->free_events.push_back(pd.solved);
      s->free_events.push_back(pd.done);
      complete(s, pd.batch, e == hipSuccess, t);
      lk.unlock();
    }
    lk.lock();
  }
}

// Whether some queued request of the lane has a free session (called with
// mu held).
bool work_ready(const gcmi_server* s, int lane) {
  for (const Request* r : s->queue) {
    const Session& ss = s->sessions[r->session];
    if (ss.lane == lane && !ss.busy) return true;
  }
  return false;
}

void worker_main(gcmi_server* s, int lane) {
  (void)hipSetDevice(s->device);
  Lane& L = s->lanes[lane];
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] { return s->stop || work_ready(s, lane); });
    if (s->stop) break;
    // oldest first, one request per session, at most max_batch, of this
    // lane's sessions.  In the warm mode a batch holds either new images
    // (cut from scratch) or strokes (reflowed warm), the class of the
    // oldest request it can take, so strokes are not held to the longer
    // solve of a new image; a session whose next request is of the other
    // class waits for a later batch.
    const int cap = s->mode == GCMI_SERVE_PLAIN ? 1 : s->max_batch;
    std::vector<Request*> batch;
    int cls = -1;  // 1: new images, 0: strokes
    std::vector<char> held;
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)batch.size() < cap;) {
      Session& ss = s->sessions[(*it)->session];
      if (ss.lane != lane || ss.busy || (!held.empty() && held[(*it)->session])) {
        ++it;
        continue;
      }
      const int c = ((*it)->kind == REQ_IMAGE || (*it)->kind == REQ_RGB) ? 1 : 0;
      if (s->mode == GCMI_SERVE_WARM) {
        if (cls < 0) cls = c;
        if (c != cls) {
          if (held.empty()) held.assign(s->sessions.size(), 0);
          held[(*it)->session] = 1;
          ++it;
          continue;
        }
      }
      ss.busy = true;
      batch.push_back(*it);
      it = s->queue.erase(it);
    }
    if (batch.empty()) continue;
    const int64_t t0 = now_ns();
    for (Request* r : batch) r->t_start = t0;
    hipEvent_t ev[2] = {nullptr, nullptr};  // solved, done
    for (hipEvent_t& e : ev)
      if (!s->free_events.empty()) {
        e = s->free_events.back();
        s->free_events.pop_back();
      }
    lk.unlock();
    int rc = GCMI_OK;
    char err[sizeof s->err] = "";
    for (hipEvent_t& e : ev)
      if (rc == GCMI_OK && !e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        e = nullptr;
        rc = GCMI_EHIP;
        snprintf(err, sizeof err, "event creation failed");
      }
    std::vector<MaskCopy> copies;
    int64_t ns[3] = {0, 0, 0};
    if (rc == GCMI_OK) {
      std::lock_guard<std::mutex> plk(L.pool_mu);
      rc = serve_batch(s, L, batch, ev[0], copies, ns);
      if (rc != GCMI_OK) snprintf(err, sizeof err, "%s", gcmi_last_error());
    }
    lk.lock();
    L.ns_apply += ns[0];
    L.ns_solve += ns[1];
    L.ns_out += ns[2];
    // the images' staging ranges were read by the solve, which has finished
    for (Request* r : batch)
      if (r->stage >= 0) {
        s->stage_alloc.release(r->stage, r->stage_len);
        r->stage = -1;
      }
    if (rc != GCMI_OK) {
      memcpy(s->err, err, sizeof err);
      for (hipEvent_t e : ev)
        if (e) s->free_events.push_back(e);
      complete(s, batch, false, now_ns());
    } else {
      s->outq.push_back(Pending{std::move(batch), std::move(copies), ev[0], ev[1]});
      s->cv_out.notify_one();
    }
    s->cv_done.notify_all();
  }
}

// Called with mu held.
int check_session(gcmi_server* s, int session) {
  if (session < 0 || session >= (int)s->sessions.size() || !s->sessions[session].open ||
      s->sessions[session].closing) {
    set_error("no open session %d", session);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

// A submission between its session check and its enqueue.  begin() checks
// the session and counts the submission in the session's `submits` (all
// under mu); gcmi_server_close waits for that count to reach zero, so the
// session cannot be closed, removed or reopened while a submission that
// passed its check is still staging its data.  A submission that fails
// after begin() drops its count in the destructor.
struct Submit {
  gcmi_server* s;
  int session = -1;
  explicit Submit(gcmi_server* srv) : s(srv) {}
  // Called with mu held.
  int begin(int sess) {
    GCMI_TRY(check_session(s, sess));
    session = sess;
    s->sessions[sess].submits++;
    return GCMI_OK;
  }
  // Called with mu held.
  void end_locked() {
    if (session < 0) return;
    s->sessions[session].submits--;
    session = -1;
    s->cv_done.notify_all();
  }
  ~Submit() {
    if (session < 0) return;
    std::lock_guard<std::mutex> lk(s->mu);
    end_locked();
  }
};

// Queues r and ends its submission; the caller has set r->t_arrive.
void enqueue(gcmi_server* s, Submit& sub, Request* r, int64_t* id_out) {
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  r->t_queued = now_ns();
  s->reqs[r->id] = r;
  s->queue.push_back(r);
  if (id_out) *id_out = r->id;
  sub.end_locked();
  s->cv_work.notify_all();  // (each lane's worker waits for its own sessions)
}

// Locks every lane's pool_mu, in lane order.
struct AllPools {
  std::unique_lock<std::mutex> lk[kMaxLanes];
  explicit AllPools(gcmi_server* s) {
    for (int i = 0; i < s->nlanes; ++i) lk[i] = std::unique_lock<std::mutex>(s->lanes[i].pool_mu);
  }
};

}  // namespace

extern "C" {

int gcmi_server_create_lanes(int tile, size_t hbm_budget, int max_sessions, int max_batch,
                             int mode, int lanes, gcmi_server** out) {
  if (!out || max_sessions < 1 || max_batch < 1 || lanes < 0 || lanes > kMaxLanes ||
      (mode != GCMI_SERVE_WARM && mode != GCMI_SERVE_PLAIN && mode != GCMI_SERVE_COLD)) {
    set_error("bad server arguments");
    return GCMI_EINVAL;
  }
  *out = nullptr;
  if (!pool_ops(tile)) {
    set_error("tile must be 16, 32 or 64 (got %d)", tile);
    return GCMI_EINVAL;
  }
  // plain mode serves one request at a time: one lane.  Otherwise three by
  // default: with the staging and mask-copy streams that is five streams,
  // and more lanes measured slower (the device's hardware queues are shared
  // between more streams).
  if (lanes == 0) lanes = mode == GCMI_SERVE_PLAIN ? 1 : 3;
  if (mode == GCMI_SERVE_PLAIN) lanes = 1;
  GCMI_TRY(require_device());
  // a quarter of the budget (at most 1 GiB) stages new images; the lanes'
  // pools (their packed masks included) share the rest equally
  const size_t stage_bytes = std::min<size_t>(hbm_budget / 4, size_t(1) << 30) & ~size_t(255);
  const long long T2 = (long long)tile * tile;
  const int probe = 1 << 12;
  const size_t b1 = pool_bytes(tile, probe, probe * T2, max_sessions);
  const size_t b2 = pool_bytes(tile, 2 * probe, 2 * probe * T2, max_sessions);
  const double per_tile = (double)(b2 - b1) / probe;
  const double fixed = (double)b1 - probe * per_tile + 4.0 * max_sessions;
  const double rest = ((double)hbm_budget - (double)stage_bytes) / lanes - fixed;
  long long cap_tiles = rest > 0 ? (long long)(rest / per_tile) : 0;
  cap_tiles = std::min<long long>(cap_tiles, ((1LL << 30) - (1LL << 20)) / T2);
  if (cap_tiles < 1 || stage_bytes < 1024) {
    set_error("HBM budget %zu too small", hbm_budget);
    return GCMI_EBUDGET;
  }
  gcmi_server* s = new (std::nothrow) gcmi_server();
  if (!s) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  s->mode = mode;
  s->max_batch = max_batch;
  s->nlanes = lanes;
  s->stage_bytes = stage_bytes;
  gcmi_rgb_model_default(&s->rgb_model);
  (void)hipGetDevice(&s->device);
  const long long cap_pix = cap_tiles * T2;
  int rc = GCMI_OK;
  for (int i = 0; i < lanes && rc == GCMI_OK; ++i)
    rc = pool_init(&s->lanes[i].pool, tile, (int)cap_tiles, cap_pix, max_sessions, nullptr);
  if (rc == GCMI_OK) {
    if (hipMalloc((void**)&s->d_stage, stage_bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->out_stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      set_error("server staging allocation failed");
      rc = GCMI_ENOMEM;
    }
  }
  if (rc != GCMI_OK) {
    if (s->d_stage) (void)hipFree(s->d_stage);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->out_stream) (void)hipStreamDestroy(s->out_stream);
    for (int i = 0; i < lanes; ++i) pool_free(&s->lanes[i].pool);
    delete s;
    return rc;
  }
  s->stage_alloc.init((long long)(stage_bytes / 4));
  s->sessions.assign(max_sessions, Session());
  for (int i = 0; i < lanes; ++i) s->lanes[i].worker = std::thread(worker_main, s, i);
  s->finisher = std::thread(finisher_main, s);
  *out = s;
  return GCMI_OK;
}

int gcmi_server_create(int tile, size_t hbm_budget, int max_sessions, int max_batch, int mode,
                       gcmi_server** out) {
  return gcmi_server_create_lanes(tile, hbm_budget, max_sessions, max_batch, mode, 0, out);
}

int gcmi_server_destroy(gcmi_server* s) {
  if (!s) return GCMI_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  for (int i = 0; i < s->nlanes; ++i)
    if (s->lanes[i].worker.joinable()) s->lanes[i].worker.join();
  {
    // the finisher completes what is queued, then stops
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_out = true;
    s->cv_out.notify_all();
  }
  if (s->finisher.joinable()) s->finisher.join();
  for (auto& kv : s->reqs) delete kv.second;
  for (Session& ss : s->sessions)
    if (ss.bits) (void)hipHostFree(ss.bits);
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  if (s->out_stream) (void)hipStreamDestroy(s->out_stream);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  if (s->d_stage) (void)hipFree(s->d_stage);
  for (int i = 0; i < s->nlanes; ++i) pool_free(&s->lanes[i].pool);
  delete s;
  return GCMI_OK;
}

int gcmi_server_set_params(gcmi_server* s, int sweeps, int bfs) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  AllPools locks(s);
  s->sweeps = sweeps;
  s->bfs = bfs;
  return GCMI_OK;
}

int gcmi_server_set_option(gcmi_server* s, int option, int value) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  AllPools locks(s);
  for (int i = 0; i < s->nlanes; ++i) GCMI_TRY(pool_set_option(&s->lanes[i].pool, option, value));
  return GCMI_OK;
}

int gcmi_server_open(gcmi_server* s, int width, int height, int* session_out) {
  if (!s || !session_out) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  AllPools locks(s);
  std::lock_guard<std::mutex> lk(s->mu);
  int slot = -1;
  for (int i = 0; i < (int)s->sessions.size(); ++i)
    if (!s->sessions[i].open) {
      slot = i;
      break;
    }
  if (slot < 0) {
    set_error("all %zu session slots are open", s->sessions.size());
    return GCMI_EBUDGET;
  }
  // the lane with the most free room first
  int order[kMaxLanes];
  for (int i = 0; i < s->nlanes; ++i) order[i] = i;
  std::sort(order, order + s->nlanes, [&](int a, int b) {
    const long long fa = s->lanes[a].pool.tiles_alloc.largest();
    const long long fb = s->lanes[b].pool.tiles_alloc.largest();
    return fa != fb ? fa > fb : a < b;
  });
  int img = -1, lane = -1, rc = GCMI_EBUDGET;
  for (int k = 0; k < s->nlanes && lane < 0; ++k) {
    rc = pool_add_image(&s->lanes[order[k]].pool, width, height, &img);
    if (rc == GCMI_OK) lane = order[k];
    else if (rc != GCMI_EBUDGET) return rc;
  }
  if (lane < 0) return rc;
  Session& ss = s->sessions[slot];
  ss = Session();
  const size_t words = ((size_t)width * height + 31) / 32;
  if (hipHostMalloc((void**)&ss.bits, words * 4, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    ss.bits = nullptr;
    pool_remove_image(&s->lanes[lane].pool, img);
    set_error("pinned mask buffer allocation failed");
    return GCMI_ENOMEM;
  }
  memset(ss.bits, 0, words * 4);
  ss.open = true;
  ss.img = img;
  ss.lane = lane;
  ss.W = width;
  ss.H = height;
  *session_out = slot;
  return GCMI_OK;
}

int gcmi_server_close(gcmi_server* s, int session) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  int lane;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    GCMI_TRY(check_session(s, session));
    s->sessions[session].closing = true;
    s->cv_done.wait(lk, [&] {
      if (s->sessions[session].busy || s->sessions[session].submits) return false;
      for (Request* r : s->queue)
        if (r->session == session) return false;
      return true;
    });
    lane = s->sessions[session].lane;
  }
  std::lock_guard<std::mutex> plk(s->lanes[lane].pool_mu);
  std::lock_guard<std::mutex> lk(s->mu);
  Session& ss = s->sessions[session];
  pool_remove_image(&s->lanes[lane].pool, ss.img);
  if (ss.bits) (void)hipHostFree(ss.bits);
  ss = Session();
  return GCMI_OK;
}

}  // extern "C"

namespace {

int submit_image_at(gcmi_server* s, int session, const int32_t* h_source, const int32_t* h_sink,
                    const int32_t* h_right, const int32_t* h_down, int64_t t_arrive,
                    int64_t* req_out) {
  if (!s || !h_source || !h_sink || !h_right || !h_down) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  size_t npx;
  Submit sub(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    GCMI_TRY(sub.begin(session));
    npx = (size_t)s->sessions[session].W * s->sessions[session].H;
  }
  int32_t maxw = 0;
  GCMI_TRY(validate_caps(h_source, h_sink, h_right, h_down, npx, &maxw));
  const long long len = ((long long)(4 * npx) + kStageAlign - 1) / kStageAlign * kStageAlign;
  if (len > s->stage_alloc.cap) {
    set_error("image of %zu pixels exceeds the staging area", npx);
    return GCMI_EBUDGET;
  }
  Request* r = new (std::nothrow) Request();
  if (!r) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  r->session = session;
  r->kind = REQ_IMAGE;
  r->max_pair = maxw;
  r->t_arrive = t_arrive;
  r->stage_len = len;
  {
    // wait for staging room: the worker frees ranges as batches complete
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return (r->stage = s->stage_alloc.alloc(len)) >= 0; });
  }
  {
    std::lock_guard<std::mutex> clk(s->copy_mu);
    int32_t* d = s->d_stage + r->stage;
    const int32_t* src[4] = {h_source, h_sink, h_right, h_down};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 4 && e == hipSuccess; ++k)
      e = hipMemcpyAsync(d + k * npx, src[k], npx * 4, hipMemcpyHostToDevice, s->copy_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->copy_stream);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("image staging copy failed: %s", hipGetErrorString(e));
      std::lock_guard<std::mutex> lk(s->mu);
      s->stage_alloc.release(r->stage, len);
      s->cv_done.notify_all();
      delete r;
      return GCMI_EHIP;
    }
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->sessions[session].max_pair = maxw;
  }
  enqueue(s, sub, r, req_out);
  return GCMI_OK;
}

int submit_rgb_at(gcmi_server* s, int session, const uint8_t* h_rgb, int channels,
                  int64_t t_arrive, int64_t* req_out) {
  if (!s || !h_rgb) {
    set_error("null argument");
    return GCMI_EINVAL;
  }
  if (channels != 3 && channels != 4) {
    set_error("a colour image has 3 or 4 channels (got %d)", channels);
    return GCMI_EINVAL;
  }
  size_t npx;
  gcmi_rgb_model model;
  Submit sub(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    GCMI_TRY(sub.begin(session));
    npx = (size_t)s->sessions[session].W * s->sessions[session].H;
    model = s->rgb_model;
  }
  int32_t wmax = 0;
  GCMI_TRY(check_rgb_model(&model, &wmax));
  const long long bytes = (long long)npx * channels;
  const long long len = ((bytes + 3) / 4 + kStageAlign - 1) / kStageAlign * kStageAlign;
  if (len > s->stage_alloc.cap) {
    set_error("image of %zu pixels exceeds the staging area", npx);
    return GCMI_EBUDGET;
  }
  Request* r = new (std::nothrow) Request();
  if (!r) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  r->session = session;
  r->kind = REQ_RGB;
  r->n = channels;
  r->model = model;
  r->max_pair = wmax;
  r->t_arrive = t_arrive;
  r->stage_len = len;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return (r->stage = s->stage_alloc.alloc(len)) >= 0; });
  }
  {
    std::lock_guard<std::mutex> clk(s->copy_mu);
    hipError_t e = hipMemcpyAsync(s->d_stage + r->stage, h_rgb, (size_t)bytes,
                                  hipMemcpyHostToDevice, s->copy_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->copy_stream);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("image staging copy failed: %s", hipGetErrorString(e));
      std::lock_guard<std::mutex> lk(s->mu);
      s->stage_alloc.release(r->stage, len);
      s->cv_done.notify_all();
      delete r;
      return GCMI_EHIP;
    }
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->sessions[session].max_pair = wmax;
  }
  enqueue(s, sub, r, req_out);
  return GCMI_OK;
}

int submit_stroke_at(gcmi_server* s, int session, int n, const int32_t* h_index,
                     const int32_t* h_source, const int32_t* h_sink, int64_t t_arrive,
                     int64_t* req_out) {
  if (!s || n < 0 || (n > 0 && (!h_index || !h_source || !h_sink))) {
    set_error("bad stroke arguments");
    return GCMI_EINVAL;
  }
  int32_t wmax;
  long long npx;
  Submit sub(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    GCMI_TRY(sub.begin(session));
    wmax = s->sessions[session].max_pair;
    npx = (long long)s->sessions[session].W * s->sessions[session].H;
  }
  // the rules of a warm edit, checked here so that both modes accept and
  // refuse the same strokes
  if (wmax < 0) {
    set_error("stroke before the session's first image");
    return GCMI_EINVAL;
  }
  static thread_local std::vector<uint64_t> mark;
  GCMI_TRY(check_edit(npx, n, h_index, h_source, h_sink, wmax, mark));
  Request* r = new (std::nothrow) Request();
  if (!r) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  r->session = session;
  r->kind = REQ_STROKE;
  r->n = n;
  r->t_arrive = t_arrive;
  r->stroke.resize(3 * (size_t)n);
  if (n) {
    memcpy(r->stroke.data(), h_index, (size_t)n * 4);
    memcpy(r->stroke.data() + n, h_source, (size_t)n * 4);
    memcpy(r->stroke.data() + 2 * n, h_sink, (size_t)n * 4);
  }
  enqueue(s, sub, r, req_out);
  return GCMI_OK;
}

int submit_polyline_at(gcmi_server* s, int session, int npts, const double* h_xy, double width,
                       int32_t source, int32_t sink, int64_t t_arrive, int64_t* req_out) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  int32_t wmax;
  int W, H;
  Submit sub(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    GCMI_TRY(sub.begin(session));
    wmax = s->sessions[session].max_pair;
    W = s->sessions[session].W;
    H = s->sessions[session].H;
  }
  if (wmax < 0) {
    set_error("stroke before the session's first image");
    return GCMI_EINVAL;
  }
  Polyline pl;
  GCMI_TRY(make_polyline(W, H, npts, h_xy, width, source, sink, wmax, &pl));
  Request* r = new (std::nothrow) Request();
  if (!r) {
    set_error("host allocation failed");
    return GCMI_ENOMEM;
  }
  r->session = session;
  r->kind = REQ_POLYLINE;
  r->n = npts;
  r->t_arrive = t_arrive;
  r->xy.assign(h_xy, h_xy + 2 * (size_t)npts);
  r->width = width;
  r->pcs = source;
  r->pct = sink;
  enqueue(s, sub, r, req_out);
  return GCMI_OK;
}

}  // namespace

extern "C" {

int gcmi_server_submit_image(gcmi_server* s, int session, const int32_t* h_source,
                             const int32_t* h_sink, const int32_t* h_right,
                             const int32_t* h_down, int64_t* req_out) {
  return submit_image_at(s, session, h_source, h_sink, h_right, h_down, now_ns(), req_out);
}

int gcmi_server_set_rgb_model(gcmi_server* s, const gcmi_rgb_model* model) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  GCMI_TRY(check_rgb_model(model, nullptr));
  std::lock_guard<std::mutex> lk(s->mu);
  s->rgb_model = *model;
  return GCMI_OK;
}

int gcmi_server_submit_rgb(gcmi_server* s, int session, const uint8_t* h_rgb, int channels,
                           int64_t* req_out) {
  return submit_rgb_at(s, session, h_rgb, channels, now_ns(), req_out);
}

int gcmi_server_submit_stroke(gcmi_server* s, int session, int n, const int32_t* h_index,
                              const int32_t* h_source, const int32_t* h_sink,
                              int64_t* req_out) {
  return submit_stroke_at(s, session, n, h_index, h_source, h_sink, now_ns(), req_out);
}

int gcmi_server_submit_polyline(gcmi_server* s, int session, int npts, const double* h_xy,
                                double width, int32_t source, int32_t sink, int64_t* req_out) {
  return submit_polyline_at(s, session, npts, h_xy, width, source, sink, now_ns(), req_out);
}

int gcmi_server_replay(gcmi_server* s, int count, const gcmi_request* reqs, int64_t* req_out) {
  if (!s || count < 0 || (count > 0 && !reqs)) {
    set_error("bad replay arguments");
    return GCMI_EINVAL;
  }
  for (int i = 0; i < count; ++i) {
    const gcmi_request& q = reqs[i];
    int64_t t = q.t_ns;
    if (t > 0) {
      std::this_thread::sleep_until(
          std::chrono::steady_clock::time_point(std::chrono::nanoseconds(t)));
    } else {
      t = now_ns();
    }
    int64_t id = 0;
    int rc;
    switch (q.kind) {
      case GCMI_REQ_IMAGE:
        rc = submit_image_at(s, q.session, (const int32_t*)q.p0, (const int32_t*)q.p1,
                             (const int32_t*)q.p2, (const int32_t*)q.p3, t, &id);
        break;
      case GCMI_REQ_STROKE:
        rc = submit_stroke_at(s, q.session, q.n, (const int32_t*)q.p0, (const int32_t*)q.p1,
                              (const int32_t*)q.p2, t, &id);
        break;
      case GCMI_REQ_POLYLINE:
        rc = submit_polyline_at(s, q.session, q.n, (const double*)q.p0, q.width, q.source, q.sink,
                                t, &id);
        break;
      case GCMI_REQ_RGB:
        rc = submit_rgb_at(s, q.session, (const uint8_t*)q.p0, q.n, t, &id);
        break;
      default:
        set_error("request %d: unknown kind %d", i, q.kind);
        rc = GCMI_EINVAL;
    }
    if (rc != GCMI_OK) return rc;
    if (req_out) req_out[i] = id;
  }
  return GCMI_OK;
}

int gcmi_server_wait(gcmi_server* s, int64_t req, int timeout_ms, int64_t* info_out) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) {
    set_error("unknown request %lld", (long long)req);
    return GCMI_EINVAL;
  }
  Request* r = it->second;
  auto done = [&] { return r->status != ST_QUEUED; };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, done);
  else
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
  if (info_out) {
    info_out[0] = r->status;
    info_out[1] = r->flow;
    info_out[2] = r->t_arrive;
    info_out[3] = r->t_start;
    info_out[4] = r->t_done;
    info_out[5] = r->batch;
    info_out[6] = r->t_queued;
  }
  if (r->status == ST_QUEUED) {
    set_error("request %lld still in flight", (long long)req);
    return GCMI_EBUSY;
  }
  if (r->status == ST_FAILED) {
    set_error("request %lld failed: %s", (long long)req, s->err);
    return GCMI_EHIP;
  }
  return GCMI_OK;
}

int gcmi_server_mask(gcmi_server* s, int session, uint32_t* h_bits, int64_t* req_out) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  GCMI_TRY(check_session(s, session));
  const Session& ss = s->sessions[session];
  if (ss.busy) {
    set_error("session %d is being served", session);
    return GCMI_EBUSY;
  }
  if (h_bits) memcpy(h_bits, ss.bits, (((size_t)ss.W * ss.H + 31) / 32) * 4);
  if (req_out) *req_out = ss.last_done;
  return GCMI_OK;
}

int gcmi_server_stats(gcmi_server* s, int64_t* out, int n) {
  if (!s || !out || n < 0) {
    set_error("bad argument");
    return GCMI_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  const int T = s->lanes[0].pool.T;
  int64_t open = 0, used = 0, cap = 0, bytes = (int64_t)s->stage_bytes;
  int64_t ns_apply = 0, ns_solve = 0, ns_out = 0;
  for (const Session& ss : s->sessions)
    if (ss.open) {
      ++open;
      used += (int64_t)((ss.W + T - 1) / T) * ((ss.H + T - 1) / T);
    }
  for (int i = 0; i < s->nlanes; ++i) {
    const Lane& L = s->lanes[i];
    cap += L.pool.cap_tiles;
    bytes += (int64_t)L.pool.device_bytes;
    ns_apply += L.ns_apply;
    ns_solve += L.ns_solve;
    ns_out += L.ns_out;
  }
  int64_t ns_wait = 0;
  for (int i = 0; i < s->nlanes; ++i) ns_wait += s->lanes[i].pool.ns_wait;
  const int64_t v[12] = {s->iterations, s->solved_requests, (int64_t)s->queue.size(), open, used,
                         cap, bytes, ns_apply, ns_solve, ns_out, s->nlanes, ns_wait};
  for (int k = 0; k < n && k < 12; ++k) out[k] = v[k];
  return GCMI_OK;
}

int gcmi_server_forget(gcmi_server* s, int64_t req) {
  if (!s) {
    set_error("null server");
    return GCMI_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) return GCMI_OK;
  if (it->second->status == ST_QUEUED) {
    set_error("request %lld still in flight", (long long)req);
    return GCMI_EBUSY;
  }
  delete it->second;
  s->reqs.erase(it);
  return GCMI_OK;
}

}  // extern "C"

// server.hip -- the serving driver of the wavefront loop: progressive-render requests on
// resident scenes, admitted while the loop runs (continuous batching).
//
// The reference's viewer calls Pathtracer::process() again and again on one scene and shows
// pixels / iteration (main.cpp:104-124, 158-173); each iteration's filter adds into the same
// pixels (TentFilter.cpp:57-60). Here one thread owns a long-lived wavefront loop over every
// resident scene and replays its graph-captured waves while requests come and go:
//   admission  FIFO; a request waits for a free job slot and for ranges of the per-sample and
//              framebuffer arenas (first fit, sized by the memory budget; no hipMalloc on the
//              way), then its framebuffer is cleared and its first iteration joins the loop;
//   iterations up to `iteration_window` iterations of a request are in the loop at once, each a
//              job with its own slot and sample range, topped up as resources free (newly
//              arrived requests first); a finished iteration is filtered and added to the
//              framebuffer strictly in iteration order (a later one that finishes first waits
//              with its samples), so the sum is the one a lone run makes;
//   delivery   after the last iteration the framebuffer is copied to pinned host memory; when
//              that copy lands, the mean over iterations is written to the caller's buffer.
// All device work is ordered on the loop's stream, so a slot or arena range can be handed to
// the next request as soon as the copies that read it are enqueued.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "range_alloc.h"
#include "wavefront.h"

namespace swr {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// One iteration of a request in the loop: a job with its own slot and sample range.
struct Iter {
  int index = 0;
  uint32_t slot = 0;
  uint64_t sample_off = 0;
  uint32_t seq = 0;  // admission number of its job
  bool done = false;
  FrameParams p{};
};

struct Request {
  uint64_t id = 0;
  int scene = 0;
  swr_render_desc desc{};
  int iterations = 1;
  // Column bands per iteration (1 unless the iteration's per-sample state exceeds the arena):
  // band b filters columns [b * band_cols, (b + 1) * band_cols); each band is a job of its own.
  int bands = 1;
  uint32_t band_cols = 0;
  float* dst = nullptr;
  Clock::time_point t_submit, t_admit, t_done;
  int status = SWR_OK;
  bool finished = false;
  // on the device (server thread only)
  uint64_t sample_n = 0, pixel_off = 0, pixel_n = 0;
  int started = 0, accumulated = 0;
  std::deque<Iter> iters;  // in flight or waiting to be accumulated, in iteration order
  float* pinned = nullptr;
  size_t pinned_bytes = 0;
  hipEvent_t ev = nullptr;
};

}  // namespace
}  // namespace swr

struct swr_server {
  swr_context* ctx = nullptr;
  swr_server_config cfg{};
  std::vector<const swr_scene*> scenes;
  swr::LoopBuffers bufs;
  std::unique_ptr<swr::WaveLoop> loop;
  swr::RangeAlloc sample_space, pixel_space;
  std::thread thread;
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  bool stop = false;
  int waiters = 0;
  int error = SWR_OK;
  std::string error_msg;
  std::unordered_map<uint64_t, std::unique_ptr<swr::Request>> reqs;
  std::deque<swr::Request*> pending;
  uint64_t next_id = 1;
  uint64_t last_cursor = 0;  // paths generated, as of the latest snapshot
  size_t rr = 0;             // round-robin position in `active`
  bool busy = false;         // some request on the device since busy_since
  std::chrono::steady_clock::time_point busy_since{};
  swr_server_totals stats{};
  // server thread only
  std::vector<uint32_t> free_slots;
  std::vector<swr::Request*> active, completing;
  std::vector<std::pair<float*, size_t>> pinned_pool;
  std::vector<hipEvent_t> event_pool;
};

namespace swr {
namespace {

float* take_pinned(swr_server* s, size_t bytes, size_t* got) {
  for (size_t i = 0; i < s->pinned_pool.size(); ++i) {
    if (s->pinned_pool[i].second >= bytes) {
      auto e = s->pinned_pool[i];
      s->pinned_pool.erase(s->pinned_pool.begin() + (long)i);
      *got = e.second;
      return e.first;
    }
  }
  float* p = nullptr;
  SWR_HIP(hipHostMalloc((void**)&p, bytes, hipHostMallocDefault));
  *got = bytes;
  return p;
}

hipEvent_t take_event(swr_server* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  SWR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return e;
}

// Column band width for a request: the whole frame when its per-sample state fits the arena,
// else bands of at most half the arena (two in flight), or the whole arena, holding their columns
// and the tent filter's neighbour columns; 0 when not even one column fits.
uint32_t plan_band_cols(const swr_render_desc& d, uint64_t arena_samples) {
  const uint64_t per_col = (uint64_t)d.height * spp_of(d);
  const bool box = d.filter == SWR_FILTER_BOX;
  if ((uint64_t)d.width * per_col <= arena_samples) return (uint32_t)d.width;
  const uint64_t half = band_width_for((uint64_t)d.width, per_col, box, arena_samples / 2);
  return (uint32_t)(half ? half : band_width_for((uint64_t)d.width, per_col, box, arena_samples));
}

// Samples one band job of the request needs in the arena (an interior band is the largest).
uint64_t band_job_samples(const Request& r) {
  const uint64_t per_col = (uint64_t)r.desc.height * spp_of(r.desc);
  if (r.bands == 1) return (uint64_t)r.desc.width * per_col;
  return band_max_samples((uint64_t)r.desc.width, per_col, r.desc.filter == SWR_FILTER_BOX, r.band_cols);
}

// Starts the next band job (iteration r->started / bands, band r->started % bands) of an
// admitted request if a slot and a sample range are free (caller holds the lock for the
// allocation; the device calls run without it).
bool start_iteration(swr_server* s, Request* r, std::unique_lock<std::mutex>& lk) {
  if (s->free_slots.empty()) return false;
  uint64_t off = 0;
  if (!s->sample_space.alloc(r->sample_n, &off)) return false;
  Iter it;
  it.index = r->started / r->bands;
  const uint32_t band = (uint32_t)(r->started % r->bands);
  it.slot = s->free_slots.back();
  s->free_slots.pop_back();
  it.sample_off = off;
  lk.unlock();
  it.p = s->loop->params((uint32_t)r->scene, r->desc, it.sample_off, r->pixel_off);
  it.p.iteration = r->desc.iteration + (uint32_t)it.index;
  if (r->bands > 1)
    set_band(&it.p, band * r->band_cols, std::min<uint32_t>(it.p.width, (band + 1) * r->band_cols));
  it.seq = s->loop->admit(it.slot, it.p);
  lk.lock();
  r->started += 1;
  r->iters.push_back(it);
  return true;
}

// Admission. Paths are generated in admission order, so what is admitted is what runs next:
// the server keeps the paths admitted but not yet generated (the backlog, as of the latest
// snapshot) under `backlog_waves` waves and fills it round robin -- a newly arrived request
// (first come, first served) gets its first iteration at once, then the admitted requests take
// turns, one iteration each. Every admitted request advances one iteration per round, like
// processor sharing at iteration granularity: short requests pass long ones, and long ones are
// never starved by a stream of short ones.
void admit(swr_server* s) {
  const int window = (int)(s->cfg.iteration_window ? s->cfg.iteration_window : 16);
  const uint64_t limit = (uint64_t)(s->cfg.backlog_waves ? s->cfg.backlog_waves : 2) * s->loop->capacity();
  std::unique_lock<std::mutex> lk(s->mu);
  while (s->loop->paths_admitted() - s->last_cursor < limit && !s->free_slots.empty()) {
    if (!s->pending.empty()) {
      Request* r = s->pending.front();
      const uint64_t ns = band_job_samples(*r), np = WaveLoop::job_pixels(r->desc);
      if (s->sample_space.fits(ns) && s->pixel_space.fits(np)) {
        s->pending.pop_front();
        s->pixel_space.alloc(np, &r->pixel_off);
        r->sample_n = ns;
        r->pixel_n = np;
        r->t_admit = Clock::now();
        lk.unlock();
        FrameParams p = s->loop->params((uint32_t)r->scene, r->desc, 0, r->pixel_off);
        s->loop->clear_pixels(p);
        lk.lock();
        start_iteration(s, r, lk);  // fits: checked above under the same lock
        s->active.push_back(r);
        continue;
      }
      // The head waits for memory (first come, first served); admitted requests go on.
    }
    bool started = false;
    const size_t n = s->active.size();
    for (size_t k = 0; k < n && !started; ++k) {
      Request* r = s->active[(s->rr + k) % n];
      if (r->started < r->iterations * r->bands && (int)r->iters.size() < window && start_iteration(s, r, lk)) {
        started = true;
        s->rr = (s->rr + k + 1) % n;
      }
    }
    if (!started) break;
  }
}

// Retires iterations finished as of a snapshot: frees their slots, filter-accumulates them in
// iteration order (freeing their sample ranges), and starts the delivery copy of requests whose
// every iteration is in the framebuffer.
void retire(swr_server* s, const WaveState& w, uint32_t seq_limit) {
  const uint32_t done = std::min(w.done_seq, seq_limit);
  std::vector<Request*> still;
  for (Request* r : s->active) {
    for (Iter& it : r->iters) {
      if (!it.done && it.seq < done) {
        it.done = true;
        std::lock_guard<std::mutex> lk(s->mu);
        s->free_slots.push_back(it.slot);  // its job has finished; later jobs may take the slot
      }
    }
    while (!r->iters.empty() && r->iters.front().done) {
      s->loop->filter(r->iters.front().p, true);
      {
        // Stream order: a later job's raygen writes the range only after this filter read it.
        std::lock_guard<std::mutex> lk(s->mu);
        s->sample_space.release(r->iters.front().sample_off, r->sample_n);
      }
      r->iters.pop_front();
      r->accumulated += 1;
    }
    if (r->accumulated < r->iterations * r->bands) {
      still.push_back(r);
      continue;
    }
    const size_t bytes = 12ull * r->pixel_n;
    r->pinned = take_pinned(s, bytes, &r->pinned_bytes);
    SWR_HIP(hipMemcpyAsync(r->pinned, s->loop->pixel_arena() + 3 * r->pixel_off, bytes, hipMemcpyDeviceToHost,
                           s->loop->stream()));
    r->ev = take_event(s);
    SWR_HIP(hipEventRecord(r->ev, s->loop->stream()));
    {
      // Later work on the stream runs after the copy, so the framebuffer range is free now.
      std::lock_guard<std::mutex> lk(s->mu);
      s->pixel_space.release(r->pixel_off, r->pixel_n);
    }
    s->completing.push_back(r);
  }
  s->active.swap(still);
  // Job completion still looks up the job of the last wave's first camera ray.
  s->loop->retire_ranges(w.raygen_base);
  std::lock_guard<std::mutex> lk(s->mu);
  s->last_cursor = w.cursor;
  s->stats.camera_rays = w.camera_rays;
  s->stats.closest_rays = w.closest_rays;
  s->stats.shadow_rays = w.shadow_rays + (w.counters >> 32);
  s->stats.waves = w.waves;
}

// Delivers every request whose copy has landed (all of them when `block`).
void deliver(swr_server* s, bool block) {
  std::vector<Request*> still;
  for (Request* r : s->completing) {
    const hipError_t q = block ? hipEventSynchronize(r->ev) : hipEventQuery(r->ev);
    if (q == hipErrorNotReady) {
      still.push_back(r);
      continue;
    }
    SWR_HIP(q);
    if (r->dst) {
      const float inv = 1.f / (float)r->iterations;
      const size_t n = 3 * r->pixel_n;
      for (size_t i = 0; i < n; ++i) r->dst[i] = r->pinned[i] * inv;
    }
    s->pinned_pool.emplace_back(r->pinned, r->pinned_bytes);
    s->event_pool.push_back(r->ev);
    r->pinned = nullptr;
    r->ev = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    r->t_done = Clock::now();
    r->finished = true;
    s->stats.completed += 1;
    s->done_cv.notify_all();
  }
  s->completing.swap(still);
}

// Server thread, on its way out: the device work of unfinished requests is abandoned.
void fail_all(swr_server* s, int code) {
  for (Request* r : s->completing) {
    if (r->pinned) s->pinned_pool.emplace_back(r->pinned, r->pinned_bytes);
    if (r->ev) s->event_pool.push_back(r->ev);
    r->pinned = nullptr;
    r->ev = nullptr;
  }
  s->completing.clear();
  s->active.clear();
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto& kv : s->reqs) {
    Request* r = kv.second.get();
    if (!r->finished) {
      r->status = code;
      r->finished = true;
      r->t_done = Clock::now();
    }
  }
  s->pending.clear();
  s->done_cv.notify_all();
}

void server_main(swr_server* s) {
  try {
    SWR_HIP(hipSetDevice(s->ctx->device));
    const int k = (int)(s->cfg.waves_per_check ? s->cfg.waves_per_check : 2);
    int ticket = -1;
    uint32_t ticket_limit = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(s->mu);
        s->work_cv.wait(lk, [&] {
          return s->stop || !s->pending.empty() || !s->active.empty() || !s->completing.empty() || ticket >= 0;
        });
        if (s->stop) break;
      }
      admit(s);
      const bool working = !s->active.empty();
      if (working && !s->busy) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->busy_since = Clock::now();
        s->busy = true;
      }
      if (working) s->loop->launch_graph_waves(k);
      int t = -1;
      uint32_t limit = 0;
      if (working || ticket >= 0) {
        t = s->loop->snapshot();
        limit = s->loop->admissions();
      }
      if (ticket >= 0) retire(s, s->loop->wait_snapshot(ticket), ticket_limit);
      ticket = t;
      ticket_limit = limit;
      deliver(s, s->active.empty() && ticket < 0);
      if (s->busy && s->active.empty() && s->completing.empty()) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.busy_s += std::chrono::duration<double>(Clock::now() - s->busy_since).count();
        s->busy = false;
      }
      if (!working && ticket >= 0) {
        // Nothing is tracing: settle the last snapshot now instead of spinning.
        retire(s, s->loop->wait_snapshot(ticket), ticket_limit);
        ticket = -1;
        deliver(s, true);
      }
    }
    if (s->busy) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stats.busy_s += std::chrono::duration<double>(Clock::now() - s->busy_since).count();
      s->busy = false;
    }
    SWR_HIP(hipStreamSynchronize(s->loop->stream()));
    fail_all(s, SWR_E_CANCELLED);
  } catch (const Error& e) {
    (void)hipStreamSynchronize(s->ctx->stream);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->error = e.code;
      s->error_msg = e.msg;
    }
    fail_all(s, e.code);
  } catch (const std::exception& e) {
    (void)hipStreamSynchronize(s->ctx->stream);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->error = SWR_E_INVALID;
      s->error_msg = e.what();
    }
    fail_all(s, SWR_E_INVALID);
  }
}

}  // namespace
}  // namespace swr

using swr::Error;

extern "C" {

int swr_server_create(swr_context* ctx, const swr_scene* const* scenes, int32_t num_scenes,
                      const swr_server_config* config, swr_server** out) {
  if (!ctx || !scenes || num_scenes <= 0 || !out) return SWR_E_INVALID;
  auto s = std::make_unique<swr_server>();
  s->ctx = ctx;
  if (config) s->cfg = *config;
  try {
    SWR_HIP(hipSetDevice(ctx->device));
    for (int i = 0; i < num_scenes; ++i) s->scenes.push_back(scenes[i]);
    const uint32_t max_req = s->cfg.max_requests ? s->cfg.max_requests : 256;
    if (max_req > 65536) swr::fail(SWR_E_INVALID, "server: at most 65536 requests on the device");
    uint32_t cap = s->cfg.wave_rays ? s->cfg.wave_rays : (1u << 23);
    cap = std::max<uint32_t>(64, std::min<uint32_t>(cap, 1u << 26));
    const uint64_t before = ctx->in_use;
    s->loop = std::make_unique<swr::WaveLoop>(ctx, s->bufs, s->scenes, max_req, cap);
    s->stats.queue_bytes = ctx->in_use - before;
    const uint64_t bytes = s->cfg.state_bytes ? s->cfg.state_bytes : (8ull << 30);
    const uint64_t pixels = bytes / 4 / 12;  // a quarter of the state for framebuffers
    const uint64_t samples = (bytes - 12 * pixels) / 24;
    if (!samples || !pixels) swr::fail(SWR_E_INVALID, "server: state budget too small");
    swr::ensure(ctx, s->bufs.radiance, 16 * samples);
    swr::ensure(ctx, s->bufs.sample_xy, 8 * samples);
    swr::ensure(ctx, s->bufs.pixels, 12 * pixels);
    s->stats.state_bytes = 24 * samples + 12 * pixels;
    s->loop->set_arenas(static_cast<float4*>(s->bufs.radiance.ptr), static_cast<float2*>(s->bufs.sample_xy.ptr),
                        static_cast<float*>(s->bufs.pixels.ptr));
    s->sample_space.reset(samples);
    s->pixel_space.reset(pixels);
    for (uint32_t i = 0; i < max_req; ++i) s->free_slots.push_back(max_req - 1 - i);
    s->loop->capture();
    swr_server* raw = s.get();
    s->thread = std::thread([raw] { swr::server_main(raw); });
  } catch (const Error& e) {
    s->loop.reset();
    s->bufs.release(ctx);
    ctx->err = e.msg;
    return e.code;
  }
  *out = s.release();
  return SWR_OK;
}

int swr_server_submit(swr_server* s, int32_t scene, const swr_render_desc* desc, int32_t iterations, float* pixels,
                      uint64_t* request) {
  if (!s || !desc || !request) return SWR_E_INVALID;
  try {
    if (scene < 0 || scene >= (int32_t)s->scenes.size()) swr::fail(SWR_E_INVALID, "server: scene index out of range");
    if (iterations < 1) swr::fail(SWR_E_INVALID, "server: iterations must be positive");
    const swr::FrameParams p = swr::make_params(*desc, &s->scenes[scene]->dev);
    if (p.max_depth > 65534) swr::fail(SWR_E_INVALID, "render: max depth above 65534");
    auto r = std::make_unique<swr::Request>();
    r->scene = scene;
    r->desc = *desc;
    r->iterations = iterations;
    r->dst = pixels;
    r->t_submit = swr::Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error) swr::fail(s->error, "server: stopped after an error: " + s->error_msg);
    s->stats.submitted += 1;
    // An iteration whose per-sample state exceeds the arena runs as column bands (out of core);
    // the framebuffer must fit whole.
    r->band_cols = swr::plan_band_cols(*desc, s->sample_space.capacity());
    if (!r->band_cols || swr::WaveLoop::job_pixels(*desc) > s->pixel_space.capacity()) {
      s->stats.refused += 1;
      swr::fail(SWR_E_BUDGET, "server: request larger than the state budget");
    }
    r->bands = (int)((desc->width + r->band_cols - 1) / r->band_cols);
    if ((int64_t)r->bands * iterations > 0x7FFFFFFF) swr::fail(SWR_E_INVALID, "server: too many band jobs");
    r->id = s->next_id++;
    *request = r->id;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->work_cv.notify_all();
    return SWR_OK;
  } catch (const Error& e) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->ctx->err = e.msg;
    return e.code;
  }
}

int swr_server_wait(swr_server* s, uint64_t request, swr_request_stats* st) {
  if (!s) return SWR_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) {
    s->ctx->err = "server: unknown request (or already waited for)";
    return SWR_E_INVALID;
  }
  swr::Request* r = it->second.get();
  s->waiters += 1;
  s->done_cv.wait(lk, [&] { return r->finished; });
  s->waiters -= 1;
  if (st) {
    st->status = r->status;
    st->iterations = (uint32_t)r->iterations;
    st->queue_ms = r->t_admit.time_since_epoch().count() ? swr::ms_between(r->t_submit, r->t_admit) : 0.0;
    st->latency_ms = swr::ms_between(r->t_submit, r->t_done);
  }
  const int status = r->status;
  s->reqs.erase(it);
  s->done_cv.notify_all();
  return status;
}

int swr_server_stats(swr_server* s, swr_server_totals* st) {
  if (!s || !st) return SWR_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  if (s->busy) st->busy_s += std::chrono::duration<double>(swr::Clock::now() - s->busy_since).count();
  return SWR_OK;
}

void swr_server_destroy(swr_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->work_cv.notify_all();
  }
  if (s->thread.joinable()) s->thread.join();
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->done_cv.wait(lk, [&] { return s->waiters == 0; });
  }
  (void)hipStreamSynchronize(s->ctx->stream);
  for (auto& kv : s->reqs) {
    if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
    if (kv.second->pinned) (void)hipHostFree(kv.second->pinned);
  }
  for (auto& p : s->pinned_pool) (void)hipHostFree(p.first);
  for (auto& e : s->event_pool) (void)hipEventDestroy(e);
  s->loop.reset();
  s->bufs.release(s->ctx);
  delete s;
}

}  // extern "C"

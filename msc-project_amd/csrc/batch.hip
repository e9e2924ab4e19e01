// batch.hip -- the batched driver of the wavefront loop: many render jobs (any resolution,
// spp and resident scene) traced together in one persistent loop replayed as HIP graphs.
//
// The reference renders one image per Pathtracer and drains that image's bins batch by batch
// (Pathtracer.cpp:451-475). Here the jobs' paths are concatenated into one global path index
// space; every wave tops up with camera rays from wherever the cursor is, so waves stay full
// across job boundaries and small jobs share waves instead of each paying a loop of its own.
// Job state lives in arenas laid out back to back (samples: 16 B radiance + 8 B film position;
// pixels: 12 B), all resident at once -- or, when a context budget cannot hold every job's state,
// in groups of consecutive jobs that fit, one persistent loop per group, each group's images
// kept in host memory (the out-of-core form of a batch).
#include <algorithm>
#include <memory>
#include <vector>

#include "wavefront.h"

struct swr_batch {
  swr_context* ctx = nullptr;
  struct Job {
    int scene;
    swr_render_desc desc;
    swr::FrameParams p;
  };
  std::vector<const swr_scene*> scenes;  // distinct, in first-use order
  std::vector<Job> jobs;
  swr::LoopBuffers bufs;
  std::unique_ptr<swr::WaveLoop> loop;   // the last run's loop (its arenas hold the results)
  bool done = false;
  bool grouped = false;                  // the last run went in groups: images in `images`
  std::vector<std::vector<float>> images;
};

namespace swr {
namespace {

uint32_t wave_capacity(uint32_t wave_rays, uint64_t total) {
  uint64_t cap = wave_rays ? wave_rays : (1u << 23);
  cap = std::min<uint64_t>(cap, 1u << 28);
  cap = std::min<uint64_t>(cap, std::max<uint64_t>(total, 64));
  return (uint32_t)std::max<uint64_t>(cap, 64);
}

// Builds the loop over `jobs` (indices into b->jobs), lays out the arenas, admits every job and
// runs the persistent loop; returns the device milliseconds from the first raygen to the last
// resolved framebuffer.
double run_jobs(swr_batch* b, const std::vector<swr_batch::Job>& jobs, uint32_t wave_rays, bool keep,
                swr_stats* stats, bool paced = false, const std::vector<size_t>* ids = nullptr) {
  swr_context* ctx = b->ctx;
  SWR_HIP(hipSetDevice(ctx->device));
  uint64_t samples = 0, pixels = 0;
  for (const auto& j : jobs) {
    samples += WaveLoop::job_samples(j.desc);
    pixels += WaveLoop::job_pixels(j.desc);
  }
  b->loop.reset();
  auto loop = std::make_unique<WaveLoop>(ctx, b->bufs, b->scenes, (uint32_t)std::max<size_t>(jobs.size(), 1),
                                         wave_capacity(wave_rays, samples));
  ensure(ctx, b->bufs.radiance, 16ull * samples);
  ensure(ctx, b->bufs.sample_xy, 8ull * samples);
  ensure(ctx, b->bufs.pixels, 12ull * pixels);
  loop->set_arenas(static_cast<float4*>(b->bufs.radiance.ptr), static_cast<float2*>(b->bufs.sample_xy.ptr),
                   static_cast<float*>(b->bufs.pixels.ptr));
  std::vector<FrameParams> params;
  uint64_t sb = 0, pb = 0;
  for (const auto& j : jobs) {
    params.push_back(loop->params((uint32_t)j.scene, j.desc, sb, pb));
    sb += WaveLoop::job_samples(j.desc);
    pb += WaveLoop::job_pixels(j.desc);
  }
  if (!paced) loop->capture();  // outside the timed region, like the plain render's allocations
  hipStream_t st = loop->stream();
  hipEvent_t ev[2];
  for (auto& e : ev) SWR_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      (void)hipEventDestroy(e[0]);
      (void)hipEventDestroy(e[1]);
    }
  } guard{ev};
  SWR_HIP(hipEventRecord(ev[0], st));
  for (size_t i = 0; i < jobs.size(); ++i) params[i].seq = loop->admit((uint32_t)i, params[i]);
  StageTimes times;
  if (paced)
    loop->run_paced(&times);
  else
    loop->run_graphs(4);
  for (const auto& p : params) loop->filter(p);
  SWR_HIP(hipEventRecord(ev[1], st));
  SWR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  SWR_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  if (stats) {
    const LoopTotals t = loop->totals();
    swr_stats s{};
    s.camera_rays = t.camera_rays;
    s.closest_rays = t.closest_rays;
    s.shadow_rays = t.shadow_rays;
    s.waves = (uint32_t)t.waves;
    s.max_wave = t.max_wave;
    s.ms_total = ms;
    s.ms_raygen = times.raygen;
    s.ms_sort_rays = times.sort_rays;
    s.ms_trace = times.trace;
    s.ms_sort_hits = times.sort_hits;
    s.ms_shade = times.shade;
    s.ms_shadow = times.shadow;
    *stats = s;
  }
  if (keep) {
    for (size_t i = 0; i < jobs.size(); ++i) b->jobs[ids ? (*ids)[i] : i].p = params[i];
    b->loop = std::move(loop);
  }
  return ms;
}

// Per-job state in bytes: 24 per sample, 12 per pixel.
uint64_t job_state_bytes(const swr_render_desc& d) {
  return 24ull * WaveLoop::job_samples(d) + 12ull * WaveLoop::job_pixels(d);
}

// Groups of consecutive jobs whose state fits `room` bytes (one group of every job when room is
// 0: no budget). A job that cannot fit alone is refused.
std::vector<std::vector<size_t>> plan_groups(const std::vector<swr_batch::Job>& jobs, uint64_t room) {
  std::vector<std::vector<size_t>> groups(1);
  uint64_t used = 0;
  for (size_t i = 0; i < jobs.size(); ++i) {
    // each arena allocation rounds up to 256 bytes
    const uint64_t need = job_state_bytes(jobs[i].desc);
    if (room && need + 768 > room)
      fail(SWR_E_BUDGET, "batch: job " + std::to_string(i) + " needs " + std::to_string(need) +
                             " B of state, more than the budget leaves (" + std::to_string(room) + " B)");
    if (room && used + need + 768 > room && !groups.back().empty()) {
      groups.emplace_back();
      used = 0;
    }
    groups.back().push_back(i);
    used += need;
  }
  return groups;
}

// Runs the batch: all jobs in one loop when their state fits the context budget (or there is
// none), else in groups that fit, each group's images copied to host memory as it finishes.
void run_batch(swr_batch* b, uint32_t wave_rays, swr_stats* stats, bool paced) {
  swr_context* ctx = b->ctx;
  b->images.clear();
  b->grouped = false;
  uint64_t total = 0, samples = 0;
  for (const auto& j : b->jobs) total += job_state_bytes(j.desc), samples += WaveLoop::job_samples(j.desc);
  // What the budget leaves for the arenas once the loop's own buffers (queues, tables, sized for
  // the whole batch so that no group grows them) are allocated.
  auto arena_room = [&]() -> uint64_t {
    const LoopBuffers& f = b->bufs;
    const uint64_t held = f.radiance.bytes + f.sample_xy.bytes + f.pixels.bytes + f.staging.bytes;
    const uint64_t others = ctx->in_use - std::min(ctx->in_use, held);
    return ctx->budget > others ? ctx->budget - others : 0;
  };
  if (!ctx->budget || total + 768 <= arena_room()) {
    run_jobs(b, b->jobs, wave_rays, true, stats, paced);
    return;
  }
  b->loop.reset();
  release(ctx, b->bufs.radiance);
  release(ctx, b->bufs.sample_xy);
  release(ctx, b->bufs.pixels);
  { WaveLoop sizing(ctx, b->bufs, b->scenes, (uint32_t)b->jobs.size(), wave_capacity(wave_rays, samples)); }
  const uint64_t room = arena_room();
  if (total + 768 <= room) {
    run_jobs(b, b->jobs, wave_rays, true, stats, paced);
    return;
  }
  // Out of core: groups of consecutive jobs whose arenas fit what is left.
  const std::vector<std::vector<size_t>> groups = plan_groups(b->jobs, room);
  b->images.assign(b->jobs.size(), {});
  swr_stats acc{};
  for (const auto& g : groups) {
    std::vector<swr_batch::Job> js;
    for (size_t i : g) js.push_back(b->jobs[i]);
    swr_stats st{};
    // fresh arenas per group: growing one while the others still hold the last group's sizes
    // could pass the budget
    b->loop.reset();
    release(ctx, b->bufs.radiance);
    release(ctx, b->bufs.sample_xy);
    release(ctx, b->bufs.pixels);
    run_jobs(b, js, wave_rays, true, &st, paced, &g);
    for (size_t i : g) {
      const FrameParams& p = b->jobs[i].p;
      b->images[i].resize(3ull * p.width * p.height);
      b->loop->copy_out(p, b->images[i].data(), nullptr);
    }
    acc.camera_rays += st.camera_rays;
    acc.closest_rays += st.closest_rays;
    acc.shadow_rays += st.shadow_rays;
    acc.waves += st.waves;
    acc.max_wave = std::max(acc.max_wave, st.max_wave);
    acc.ms_total += st.ms_total;
    acc.ms_raygen += st.ms_raygen;
    acc.ms_sort_rays += st.ms_sort_rays;
    acc.ms_trace += st.ms_trace;
    acc.ms_sort_hits += st.ms_sort_hits;
    acc.ms_shade += st.ms_shade;
    acc.ms_shadow += st.ms_shadow;
  }
  b->grouped = true;
  if (stats) *stats = acc;
}

template <class F>
int guarded_batch(swr_batch* b, F&& f) {
  if (!b || !b->ctx) return SWR_E_INVALID;  // no batch, or its context was destroyed
  try {
    f();
    return SWR_OK;
  } catch (const Error& e) {
    (void)hipStreamSynchronize(b->ctx->stream);
    b->ctx->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    (void)hipStreamSynchronize(b->ctx->stream);
    b->ctx->err = e.what();
    return SWR_E_INVALID;
  }
}

}  // namespace

void detach_batch(swr_batch* b) {
  if (!b->ctx) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  b->loop.reset();
  b->bufs.release(b->ctx);
  b->images.clear();
  b->done = false;
  b->ctx = nullptr;
}

}  // namespace swr

using swr::Error;

extern "C" {

int swr_batch_create(swr_context* ctx, swr_batch** out) {
  if (!ctx || !out) return SWR_E_INVALID;
  auto* b = new swr_batch();
  b->ctx = ctx;
  ctx->batches.insert(b);
  *out = b;
  return SWR_OK;
}

void swr_batch_destroy(swr_batch* b) {
  if (!b) return;
  if (b->ctx) {
    b->ctx->batches.erase(b);
    swr::detach_batch(b);
  }
  delete b;
}

int swr_batch_add(swr_batch* b, const swr_scene* scene, const swr_render_desc* desc, int32_t* job) {
  return swr::guarded_batch(b, [&] {
    if (!scene || !desc) swr::fail(SWR_E_INVALID, "batch: null scene or desc");
    if (scene->ctx != b->ctx) swr::fail(SWR_E_INVALID, "batch: scene belongs to another (or a destroyed) context");
    if (!scene->committed) swr::fail(SWR_E_INVALID, "batch: scene not committed");
    if (b->jobs.size() >= 65536) swr::fail(SWR_E_INVALID, "batch: at most 65536 jobs");
    auto it = std::find(b->scenes.begin(), b->scenes.end(), scene);
    if (it == b->scenes.end()) {
      if (b->scenes.size() >= 256) swr::fail(SWR_E_INVALID, "batch: at most 256 distinct scenes");
      b->scenes.push_back(scene);
      it = b->scenes.end() - 1;
    }
    // Validate now, so that a bad job is refused here and not at run time.
    swr::FrameParams p = swr::make_params(*desc, &scene->dev);
    b->jobs.push_back({(int)(it - b->scenes.begin()), *desc, p});
    b->done = false;
    if (job) *job = (int32_t)b->jobs.size() - 1;
  });
}

int swr_batch_run(swr_batch* b, uint32_t wave_rays, uint32_t flags, swr_stats* stats) {
  return swr::guarded_batch(b, [&] {
    if (b->jobs.empty()) swr::fail(SWR_E_INVALID, "batch: no jobs");
    if (flags & ~SWR_BATCH_PACED) swr::fail(SWR_E_INVALID, "batch: unknown flags");
    b->done = false;
    swr::run_batch(b, wave_rays, stats, (flags & SWR_BATCH_PACED) != 0);
    b->done = true;
  });
}

int swr_batch_result(swr_batch* b, int32_t job, float* pixels, float* samples) {
  return swr::guarded_batch(b, [&] {
    if (!b->done || !b->loop) swr::fail(SWR_E_INVALID, "batch: not run");
    if (job < 0 || (size_t)job >= b->jobs.size()) swr::fail(SWR_E_INVALID, "batch: job index out of range");
    if (b->grouped) {
      if (samples) swr::fail(SWR_E_INVALID, "batch: sample records of a batch run in groups are not kept");
      if (pixels) std::copy(b->images[job].begin(), b->images[job].end(), pixels);
      return;
    }
    b->loop->copy_out(b->jobs[job].p, pixels, samples);
  });
}

int swr_batch_autotune(swr_batch* b, uint32_t wave_rays, swr_tuning* out) {
  return swr::guarded_batch(b, [&] {
    if (b->jobs.empty()) swr::fail(SWR_E_INVALID, "batch: no jobs");
    swr_context* ctx = b->ctx;
    auto reduced = [](swr_render_desc d) {
      if (d.spp > 0 && d.sampler == SWR_SAMPLER_INDEPENDENT)
        d.spp = std::max(1, d.spp / 16);
      else
        d.base = std::max(1, d.base / 4), d.spp = 0;
      return d;
    };
    std::vector<swr_batch::Job> small;
    for (const auto& j : b->jobs) small.push_back({j.scene, reduced(j.desc), j.p});
    const swr::Tuning saved = ctx->tuning;
    swr::Tuning base = saved;
    base.trace_stack = 0;
    ctx->tuning = base;
    auto time_jobs = [&](const std::vector<swr_batch::Job>& js, int reps) {
      double best = 1e30;
      for (int r = 0; r < reps; ++r) best = std::min(best, swr::run_jobs(b, js, wave_rays, false, nullptr));
      return best;
    };
    try {
      // Traversal ring per scene, on that scene's jobs alone.
      for (size_t s = 0; s < b->scenes.size(); ++s) {
        std::vector<swr_batch::Job> js;
        for (const auto& j : small)
          if (j.scene == (int)s) js.push_back(j);
        swr_scene* sc = const_cast<swr_scene*>(b->scenes[s]);
        int best_ring = 0;
        double best_ms = 1e30;
        time_jobs(js, 1);  // first touch: allocations and graph capture outside the timings
        for (int ring : {8, 16, 32, 64}) {
          sc->tuned_ring = ring;
          const double ms = time_jobs(js, 2);
          if (ms < best_ms) best_ms = ms, best_ring = ring;
        }
        sc->tuned_ring = best_ring;
      }
      // Sort tile and key widths on the whole reduced batch.
      swr::Tuning best = ctx->tuning;
      time_jobs(small, 1);
      double best_ms = time_jobs(small, 2);
      auto trial = [&](swr::Tuning t) {
        ctx->tuning = t;
        const double ms = time_jobs(small, 2);
        if (ms < best_ms) best_ms = ms, best = t;
        ctx->tuning = best;
      };
      for (int alg : {SWR_SORT_ONESWEEP, SWR_SORT_REDUCE_SCAN})
        for (int kpt : {8, 12, 16, 20, 24}) {
          swr::Tuning t = best;
          t.sort_algorithm = alg;
          t.sort_kpt = kpt;
          trial(t);
        }
      for (int rb : {32, 24, 20, 16, 12, 8}) {
        swr::Tuning t = best;
        t.ray_key_bits = rb;
        trial(t);
      }
      for (int hb : {0, 24, 20, 16, 12, 8}) {
        swr::Tuning t = best;
        t.hit_key_bits = hb;
        trial(t);
      }
      ctx->tuning = best;
    } catch (...) {
      ctx->tuning = saved;
      throw;
    }
    b->loop.reset();
    b->done = false;
    if (out) swr_get_tuning(ctx, out);
  });
}

}  // extern "C"

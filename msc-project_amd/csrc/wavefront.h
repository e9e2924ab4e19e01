// wavefront.h -- the sorted wavefront loop (render.hip) as its three drivers use it: one plain
// render with a host-paced loop, a batch of renders in one persistent loop replayed as HIP
// graphs, and the serving loop that admits and retires jobs while the loop runs.
//
// A loop owns a fixed set of resident scenes, a table of job slots (one render job = one
// Pathtracer::process() iteration of one camera), the ray queues of one wave and the look-up
// tables the kernels read. Job state (per-sample radiance and film positions, pixels) lives in
// arenas whose offsets the driver chooses, so that jobs of any size can come and go.
#pragma once

#include <cstdint>
#include <vector>

#include "integrator.h"
#include "scene.h"

namespace swr {

// Device bookkeeping of a loop (one copy in device memory). The append counter that shading
// hits with atomics has a cache line of its own: reads of a line under atomics wait behind them,
// and every wave of every kernel reads the wave size.
struct alignas(128) WaveState {
  uint32_t n, k;          // rays in the wave in flight; camera rays among them (its last k slots)
  uint32_t cap;           // queue capacity
  uint32_t range_lo, num_ranges;  // live part of the job range table
  uint32_t done_seq;      // every job with a smaller admission number has finished
  uint32_t max_wave;
  uint32_t live_min;      // smallest admission number among the last wave's continuations
  uint32_t shadow_n;      // shadow-record positions of the last shaded wave (its size), which its
                          // shadow pass reads while the next wave may already have begun
  uint32_t live_min_prev; // live_min of the wave before the last
  uint64_t raygen_base;   // global path index of the wave's first camera ray
  uint64_t cursor;        // next global path index to generate
  uint64_t total;         // global path indices admitted so far
  uint64_t camera_rays, closest_rays, shadow_rays, waves;  // running totals
  alignas(128) uint64_t counters;  // appends of the wave in flight: low word continuations, high word shadow rays
};

// Global path indices [path_begin, next range's path_begin) belong to job slot `job`.
struct JobRange {
  uint64_t path_begin;
  uint32_t job;
  uint32_t seq;
};

struct LoopTotals {
  uint64_t camera_rays = 0, closest_rays = 0, shadow_rays = 0, waves = 0;
  uint32_t max_wave = 0;
};

// Per-stage device times of a host-paced wave (events on the loop's stream).
struct StageTimes {
  double raygen = 0, sort_rays = 0, trace = 0, sort_hits = 0, shade = 0, shadow = 0;
};

// Device buffers of a loop, owned by its driver so that they can outlive one loop: the plain
// render keeps one set per context and reuses it frame after frame.
struct LoopBuffers {
  DeviceBuffer queue[2], keys[2], hit_keys, perm, alt_k, alt_v, shadow, spill;
  DeviceBuffer spill_shadow;  // the shadow pass's stack spill when it overlaps the next traversal
  DeviceBuffer jobs, ranges, scene_tab, scene_ring, wave, block_min;
  DeviceBuffer radiance, sample_xy, pixels, staging;  // arenas, when the driver keeps them here
  WaveState* host_state = nullptr;                     // pinned snapshot slots
  WaveState* host_wave = nullptr;  // coherent pinned slots the run-ahead loop's wave end writes
  hipEvent_t snap_ev[8] = {};
  hipEvent_t stage_ev[3][8] = {};  // per-stage events of up to three waves in flight
  // The run-ahead loop's second stream (default priority): a wave's shadow pass runs there while
  // the next wave's ray generation, sorts and traversal run on the loop's stream.
  hipStream_t side = nullptr;
  // per queue parity: where the shadow pass may start (after the next wave's ray sort), its end
  hipEvent_t fork_ev[2] = {}, shadow_ev[2] = {};
  void release(swr_context* ctx);
};

class WaveLoop {
 public:
  // scenes: the resident scenes (at most 256); max_jobs: job slots (at most 65536);
  // capacity: rays per wave (64 .. 2^28).
  WaveLoop(swr_context* ctx, LoopBuffers& bufs, const std::vector<const swr_scene*>& scenes, uint32_t max_jobs,
           uint32_t capacity);
  ~WaveLoop();
  WaveLoop(const WaveLoop&) = delete;
  WaveLoop& operator=(const WaveLoop&) = delete;

  // Samples and pixels of a job (arena sizes the driver must provide).
  static uint64_t job_samples(const swr_render_desc& d);
  static uint64_t job_pixels(const swr_render_desc& d);

  // Validates desc and fills the frame constants of a job on scene `scene` (an index into the
  // constructor's list) whose state starts at the given arena offsets.
  FrameParams params(uint32_t scene, const swr_render_desc& desc, uint64_t sample_base, uint64_t pixel_base) const;

  // Arenas shared by the jobs: per-sample radiance (16 B) and film position (8 B), pixels (12 B).
  void set_arenas(float4* radiance, float2* sample_xy, float* pixels);

  // Admits a job into `slot` (stream-ordered, so it is safe while waves are queued): its paths
  // follow every path admitted before. Returns its admission number. The slot must be free: its
  // previous job (if any) finished (snapshot done_seq above that job's number).
  uint32_t admit(uint32_t slot, FrameParams p);
  // Drops range-table entries of jobs whose paths were all generated before `cursor` (a value
  // read from a snapshot), making room for more admissions.
  void retire_ranges(uint64_t cursor);

  // One wave: raygen refill, ray sort, trace, hit sort, shade, shadow. With `exact` the wave
  // size is known on the host and launches are sized to it; otherwise they are sized for the
  // capacity and read the size on the device (the form captured into graphs).
  // With `times`, per-stage events are recorded into event set `ev_set` (0..2).
  // With `overlap` (run-ahead loop only) the wave's shadow pass goes to the second stream: the
  // loop's stream goes on to the next wave and waits for it only before that wave's shading,
  // the first kernel after it that adds to per-sample radiance.
  // With `snap`, the wave's last kernel also copies the device state there (host memory).
  void enqueue_wave(int cur, bool exact, uint32_t n, uint32_t k, StageTimes* times, int ev_set = 0,
                    bool overlap = false, WaveState* snap = nullptr);
  // Adds the stage times of a finished wave recorded into event set `ev_set`.
  void add_stage_times(StageTimes* times, int ev_set);

  // Host-paced loop until every admitted path has finished; per-stage times if asked.
  void run_paced(StageTimes* times);
  // Run-ahead loop: device-sized waves enqueued without host round trips, each checked through
  // a snapshot one wave behind, so the device always has the next wave queued; per-stage times
  // (read once a wave is known finished) if asked. One empty wave runs after the last.
  void run_ahead(StageTimes* times);

  // Persistent loop: captures the wave for both queue parities once, then replays graphs until
  // every admitted path has finished (checked every `check_every` waves, two checks in flight).
  void capture();
  void run_graphs(int check_every = 4);
  // Replays `waves` graph waves without waiting (serving loop); a snapshot then reads the state.
  void launch_graph_waves(int waves);
  // Asynchronously snapshots the device state into host memory (done_seq recomputed against
  // every admission enqueued before it); returns a ticket.
  int snapshot();
  // Waits for the snapshot with this ticket and returns it.
  const WaveState& wait_snapshot(int ticket);

  // Filters job `p` into its pixels (in the pixel arena); with accumulate, adds the filtered
  // iteration to them (the reference's pixels += across iterations, TentFilter.cpp:57-60).
  void filter(const FrameParams& p, bool accumulate = false);
  // Zeroes job `p`'s pixels (stream-ordered).
  void clear_pixels(const FrameParams& p);
  // Copies the job's pixels and, if asked, its sample records (x, y, r, g, b in reference
  // sample-id order) to the host (synchronises).
  void copy_out(const FrameParams& p, float* pixels_host, float* samples_host);
  // Copies the sample records of the columns job `p` filters (out_x0..out_x1) to their place in
  // a whole frame's records (synchronises).
  void copy_samples(const FrameParams& p, float* samples_host);

  LoopTotals totals();  // running totals from a fresh snapshot (synchronises)
  hipStream_t stream() const { return stream_; }
  uint32_t capacity() const { return cap_; }
  int num_scenes() const { return (int)scenes_.size(); }
  const swr_scene* scene(int i) const { return scenes_[i]; }
  uint64_t paths_admitted() const { return total_; }
  uint32_t admissions() const { return seq_; }  // the next admission number
  float* pixel_arena() const { return pixels_; }

 private:
  swr_context* ctx_;
  LoopBuffers& b_;
  hipStream_t stream_;
  std::vector<const swr_scene*> scenes_;
  std::vector<int> rings_;        // ring size per scene
  std::vector<int> ring_set_;     // distinct ring sizes
  uint32_t max_jobs_, cap_;
  int scene_bits_ = 0, ray_bits_ = 32, hit_bits_ = 0, hit_shift_ = 0;
  uint32_t spill_stride_ = 0;
  bool spill_ = false;
  bool textured_ = false;         // some scene has textured shaders
  uint64_t total_ = 0;            // paths admitted
  uint32_t seq_ = 0, num_ranges_ = 0, range_lo_ = 0, range_cap_ = 0;
  float4* radiance_ = nullptr;
  float2* sxy_ = nullptr;
  float* pixels_ = nullptr;
  hipGraphExec_t exec_[2] = {nullptr, nullptr};
  hipGraph_t graph_[2] = {nullptr, nullptr};
  int cur_ = 0;
  int passes_per_wave_ = 0;
  bool shadow_pending_ = false;  // a shadow pass on the second stream not yet waited for
  int shadow_last_ = 0;          // its queue parity
  bool shadow_deferred_ = false; // a shaded wave whose shadow pass is not launched yet
  // Each wave's shadow pass runs with the next wave (run-ahead and graph loops): a job is then
  // done only once no ray of the last wave belongs to it (its last shadow pass has run).
  bool lagged_ = false;
  void ensure_side();            // the second stream, its events and the shadow pass's spill
  int deferred_cur_ = 0, deferred_set_ = 0;
  // Shadow pass of the last shaded wave on stream sst with stack spill sspill, over nb positions
  // (the wave's size when the host knows it, else the capacity).
  void launch_shadow(hipStream_t sst, uint32_t* sspill, uint32_t nb, StageTimes* times, int ev_set);
  void drain_side();             // the loop's stream waits for the shadow pass on the second stream
  void flush_shadow(StageTimes* times);  // launches a deferred shadow pass, then drain_side
  int snap_next_ = 0;
  std::vector<uint64_t> range_begin_;  // host copy of the range table (path_begin per range)
};

}  // namespace swr

// swr_internal.h -- host-side internals shared by the libswr translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "swr.h"

namespace swr {

// Raised inside the library and turned into a status code at the C boundary.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

// Launch shape of a kernel, checked on the host before the launch is issued or captured into a
// graph: a grid of at least one workgroup, 1..1024 lanes per workgroup and fewer than 2^32
// work-items in all (the dispatch packet's grid size is a 32-bit count of work-items). A shape
// outside these is refused with SWR_E_INVALID instead of reaching the hardware queue.
inline uint32_t checked_grid(uint64_t groups, uint32_t block, const char* kernel) {
  if (groups == 0 || block == 0 || block > 1024 || groups * block > 0xFFFFFFFFull)
    fail(SWR_E_INVALID, std::string("launch shape out of range for ") + kernel + ": " + std::to_string(groups) +
                            " workgroups of " + std::to_string(block) + " lanes");
  return (uint32_t)groups;
}

#define SWR_HIP(call)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      ::swr::fail(SWR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
  } while (0)

// Device buffer owned by a context (accounted against its budget).
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Workspace of the onesweep radix sort.
struct SortWorkspace {
  DeviceBuffer alt_keys, alt_vals;   // ping-pong partners of the caller's buffers
  DeviceBuffer hist;                 // 8 copies x 4 passes x 256 digits (summed by the pass kernel)
  DeviceBuffer counters;             // 4 tile tickets and the device epoch counter
  DeviceBuffer status;               // look-back words, tiles x 256 x u64
  DeviceBuffer tile_counts;          // per-tile digit counts of the reduce-then-scan passes, 256 x tile_stride
  uint32_t tile_stride = 0;
  uint64_t epoch = 0;                // host mirror of the device epoch counter (tags issued so far)
};

}  // namespace swr

struct swr_scene;
struct swr_batch;

namespace swr {
struct LoopBuffers;
void destroy_loop_buffers(swr_context* ctx);
// Launch shapes chosen per context (by swr_autotune or swr_set_tuning).
struct Tuning {
  int sort_kpt = 16;       // keys per thread of a onesweep tile (tile = 512 * kpt)
  int ray_key_bits = 32;   // top bits of the octant + Morton ray key that the ray sort orders by
  int hit_key_bits = 0;    // top bits of the hit key sorted (0 = all: full (geomID, primID) order)
  int trace_stack = 0;     // LDS ring entries of the traversal stack (0 = from the BVH depth)
  int sort_algorithm = SWR_SORT_ONESWEEP;  // pass structure of the radix sort
  int shadow_overlap = 1;  // the shadow pass beside the next wave's traversal (second stream)
};
}  // namespace swr

struct swr_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::atomic<bool> cancel{false};
  uint64_t budget = 0;               // 0 = unlimited
  uint64_t in_use = 0;
  std::unordered_map<void*, size_t> allocs;
  swr::SortWorkspace sort;
  swr::Tuning tuning;
  swr::LoopBuffers* loop = nullptr;          // wavefront buffers of the plain render (render.hip)
  swr::DeviceBuffer trace_spill;             // traversal stack spill of the stand-alone queries
  std::unordered_set<swr_scene*> scenes;     // live scenes; destroying the context detaches them
  std::unordered_set<swr_batch*> batches;    // live batches; likewise
  int num_cus = 0;
  uint32_t* sort_scratch = nullptr;          // device address of sort.hip's scratch array on `device`
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // brackets of the last timed entry point
  double last_ms = 0.0;                     // its device time
};

namespace swr {

// Device time of one entry point: events on the context stream around its launches.
struct Timer {
  swr_context* ctx;
  explicit Timer(swr_context* c) : ctx(c) { SWR_HIP(hipEventRecord(ctx->ev0, ctx->stream)); }
  void stop() { SWR_HIP(hipEventRecord(ctx->ev1, ctx->stream)); }
  void finish() {
    float ms = 0.f;
    SWR_HIP(hipEventSynchronize(ctx->ev1));
    SWR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
  }
};

// Called by swr_destroy for every live scene: the scene forgets its (now freed) device memory
// and context; its handle stays valid for swr_scene_destroy.
void detach_scene(swr_scene* s);
// Called by swr_destroy for every live batch: the batch frees its device buffers and forgets its
// context; its handle stays valid for swr_batch_destroy (every other call then fails).
void detach_batch(swr_batch* b);

// Allocation accounted against the context budget.
void* dev_alloc(swr_context* ctx, size_t bytes);
void dev_free(swr_context* ctx, void* p);
// Grow a context-owned buffer to at least bytes (contents not preserved).
void ensure(swr_context* ctx, DeviceBuffer& b, size_t bytes);
void release(swr_context* ctx, DeviceBuffer& b);

// ---- sort.hip ----
// Stable sort of n (key, value) pairs on key bits [0, key_bits). keys/vals hold the input; the
// sorted output lands in whichever of (keys, vals) or (alt_keys, alt_vals) the last pass wrote,
// returned through out_keys / out_vals. iota: values are 0..n-1 (vals is not read).
// With n_dev, the element count is read on the device at run time and n is only its upper
// bound (launch sizes); such a sort may be captured into a graph once sort_reserve(n) has run.
void sort_pairs(swr_context* ctx, uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals, size_t n,
                int key_bits, bool iota, uint32_t** out_keys, uint32_t** out_vals, hipStream_t stream,
                const uint32_t* n_dev = nullptr);
// The look-back tags come from a 32-bit device counter; before it could wrap (checked against
// a host mirror), the status words are cleared. Graph replays call the guard between launches.
void sort_epoch_guard(swr_context* ctx, uint64_t passes_ahead, hipStream_t stream);
void sort_reset_epochs(swr_context* ctx, hipStream_t stream);
// Scratch sizes the sort needs for n elements.
size_t sort_status_bytes(size_t n);
void sort_reserve(swr_context* ctx, size_t n);
#ifndef SWR_SORT_BLOCK
#define SWR_SORT_BLOCK 512
#endif
constexpr int kSortBlock = SWR_SORT_BLOCK;  // lanes per workgroup of the sort passes
constexpr int kSortMinKeysPerThread = 8;
// The sort without allocation (swr_sort_pairs_u32_noalloc): at most this many pairs, with all
// of its scratch in a file-scope device array of sort.hip (sort_scratch_bytes(), about 8.3 MiB).
constexpr size_t kNoAllocMaxPairs = size_t(1) << 20;
size_t sort_scratch_bytes();
// Looks up that array's device address on the current device (once per context, at creation).
void sort_scratch_init(swr_context* ctx);
// Reduce-then-scan passes with the scratch array as workspace; the sorted pairs end in keys/vals.
// Stream-ordered; calls on one device share the array, so calls that overlap on a device race.
void sort_pairs_noalloc(swr_context* ctx, uint32_t* keys, uint32_t* vals, size_t n, int key_bits, bool iota,
                        hipStream_t stream);

// ---- keys.hip ----
void launch_ray_keys(const float4* rays, size_t n, float3 lo, float3 scale, uint32_t* keys, hipStream_t s);
void launch_hit_keys(const uint4* hits, size_t n, uint32_t* keys, hipStream_t s);
void launch_mesh_face_keys(const uint32_t* mesh, const uint32_t* face, size_t n, uint32_t* keys, hipStream_t s);
void launch_gather_u32(const uint32_t* in, const uint32_t* perm, size_t n, int words, uint32_t* out, hipStream_t s);

}  // namespace swr

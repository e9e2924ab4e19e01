// trace.hip -- stand-alone closest-hit and occlusion queries (the rtcIntersect / rtcOccluded
// entry points of the C ABI). The traversal itself lives in traverse.h and is shared with the
// render loop's kernels.
#include "traverse.h"

namespace swr {
namespace {

template <int kShort, bool kAny>
__global__ __launch_bounds__(kTraceBlock) void trace_kernel(const Node4* __restrict__ nodes,
                                                            const float4* __restrict__ leaf_tris,
                                                            const float4* __restrict__ rays, uint32_t n,
                                                            uint32_t mask, uint4* __restrict__ hits,
                                                            uint32_t* __restrict__ keys, uint32_t* __restrict__ flags,
                                                            uint32_t* __restrict__ spill, uint32_t spill_stride) {
  __shared__ uint32_t s_ring[kShort * kTraceBlock];
  const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  const RayState r = make_ray(rays[2 * (size_t)i], rays[2 * (size_t)i + 1]);
  const HitResult h = traverse<kTraceBlock, kShort, kAny>(nodes, leaf_tris, r, mask, s_ring + threadIdx.x,
                                                          spill + (size_t)i * spill_stride);
  if (kAny) {
    flags[i] = h.id != kMiss ? 1u : 0u;
  } else {
    hits[i] = make_uint4(__float_as_uint(h.t), __float_as_uint(h.u), __float_as_uint(h.v), h.id);
    if (keys) keys[i] = h.id + 1u;
  }
}

template <bool kAny>
void launch_trace(swr_context* ctx, const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask, uint4* hits,
                  uint32_t* keys, uint32_t* flags, hipStream_t stream) {
  if (!n) return;
  const int ring = choose_ring(ctx->tuning.trace_stack, s.max_depth);
  const uint32_t need = stack_need(s.max_depth);
  uint32_t* spill = nullptr;
  if ((uint32_t)ring < need) {
    ensure(ctx, ctx->trace_spill, (size_t)n * need * sizeof(uint32_t));
    spill = static_cast<uint32_t*>(ctx->trace_spill.ptr);
  }
  with_ring(ring, [&](auto kShort) {
    trace_kernel<kShort.value, kAny><<<(n + kTraceBlock - 1) / kTraceBlock, kTraceBlock, 0, stream>>>(
        s.nodes, s.leaf_tris, rays, n, mask, hits, keys, flags, spill, need);
  });
  SWR_HIP(hipGetLastError());
}

}  // namespace

uint32_t max_supported_depth() { return 1u << 20; }

void launch_intersect(swr_context* ctx, const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask,
                      uint4* hits, uint32_t* keys, hipStream_t stream) {
  launch_trace<false>(ctx, s, rays, n, mask, hits, keys, nullptr, stream);
}

void launch_occluded(swr_context* ctx, const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask,
                     uint32_t* flags, hipStream_t stream) {
  launch_trace<true>(ctx, s, rays, n, mask, nullptr, nullptr, flags, stream);
}

}  // namespace swr

// trace.hip -- stand-alone closest-hit and occlusion queries (the rtcIntersect / rtcOccluded
// entry points of the C ABI). The traversal itself lives in traverse.h and is shared with the
// render loop's fused kernels.
#include "traverse.h"

namespace swr {
namespace {

template <int kStack, int kBlock, bool kAny>
__global__ __launch_bounds__(kBlock) void trace_kernel(const Node4* __restrict__ nodes,
                                                       const float4* __restrict__ leaf_tris,
                                                       const float4* __restrict__ rays, uint32_t n, uint32_t mask,
                                                       uint4* __restrict__ hits, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ flags) {
  __shared__ uint32_t s_stack[kStack * kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const RayState r = make_ray(rays[2 * (size_t)i], rays[2 * (size_t)i + 1]);
  const HitResult h = traverse<kBlock, kAny>(nodes, leaf_tris, r, mask, s_stack + threadIdx.x);
  if (kAny) {
    flags[i] = h.id != kMiss ? 1u : 0u;
  } else {
    hits[i] = make_uint4(__float_as_uint(h.t), __float_as_uint(h.u), __float_as_uint(h.v), h.id);
    if (keys) keys[i] = h.id + 1u;
  }
}

template <bool kAny>
void launch_trace(const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask, uint4* hits, uint32_t* keys,
                  uint32_t* flags, hipStream_t stream) {
  if (!n) return;
  const uint32_t need = stack_need(s.max_depth);
  if (need <= 32) {
    constexpr int B = 256;
    trace_kernel<32, B, kAny><<<(n + B - 1) / B, B, 0, stream>>>(s.nodes, s.leaf_tris, rays, n, mask, hits, keys, flags);
  } else if (need <= 64) {
    constexpr int B = 128;
    trace_kernel<64, B, kAny><<<(n + B - 1) / B, B, 0, stream>>>(s.nodes, s.leaf_tris, rays, n, mask, hits, keys, flags);
  } else if (need <= 128) {
    constexpr int B = 64;
    trace_kernel<128, B, kAny><<<(n + B - 1) / B, B, 0, stream>>>(s.nodes, s.leaf_tris, rays, n, mask, hits, keys,
                                                                 flags);
  } else {
    fail(SWR_E_INVALID, "traversal: BVH too deep for the LDS stack (depth " + std::to_string(s.max_depth) + ")");
  }
  SWR_HIP(hipGetLastError());
}

}  // namespace

uint32_t max_supported_depth() { return (128 - 4) / 3; }

void launch_intersect(const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask, uint4* hits, uint32_t* keys,
                      hipStream_t stream) {
  launch_trace<false>(s, rays, n, mask, hits, keys, nullptr, stream);
}

void launch_occluded(const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask, uint32_t* flags,
                     hipStream_t stream) {
  launch_trace<true>(s, rays, n, mask, nullptr, nullptr, flags, stream);
}

}  // namespace swr

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
  const uint32_t waves = gridDim.x * (kTraceBlock / 64);
  const uint32_t chunk = chunk_rays(n, waves);
  const uint32_t wave = blockIdx.x * (kTraceBlock / 64) + threadIdx.x / 64;
  const uint32_t begin = wave * chunk;
  if (begin >= n) return;
  const uint32_t end = min(n, begin + chunk);
  trace_chunk<kTraceBlock, kShort, kAny>(
      begin, end, mask, s_ring + threadIdx.x, spill + ((size_t)blockIdx.x * kTraceBlock + threadIdx.x) * spill_stride,
      [&](uint32_t i, auto& wk) {
        wk.start(make_ray(rays[2 * (size_t)i], rays[2 * (size_t)i + 1]), nodes, leaf_tris);
        return true;
      },
      [&](uint32_t i, auto& wk) {
        if (kAny) {
          flags[i] = wk.best.id != kMiss ? 1u : 0u;
        } else {
          hits[i] = make_uint4(__float_as_uint(wk.best.t), __float_as_uint(wk.best.u), __float_as_uint(wk.best.v),
                               wk.best.id);
          if (keys) keys[i] = wk.best.id + 1u;
        }
      });
}

template <bool kAny>
void launch_trace(swr_context* ctx, const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask, uint4* hits,
                  uint32_t* keys, uint32_t* flags, hipStream_t stream) {
  if (!n) return;
  const int ring = choose_ring(ctx->tuning.trace_stack, s.max_depth);
  const uint32_t need = stack_need(s.max_depth);
  const unsigned tb = trace_blocks(refill_waves(n));
  uint32_t* spill = nullptr;
  if ((uint32_t)ring < need) {
    // one spill column per traversal lane
    ensure(ctx, ctx->trace_spill, (size_t)tb * kTraceBlock * need * sizeof(uint32_t));
    spill = static_cast<uint32_t*>(ctx->trace_spill.ptr);
  }
  with_ring(ring, [&](auto kShort) {
    trace_kernel<kShort.value, kAny><<<tb, kTraceBlock, 0, stream>>>(s.nodes, s.leaf_tris, rays, n, mask, hits, keys,
                                                                      flags, spill, need);
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

// keys.hip -- sort keys and gathers of the sorted wavefront loop (gfx950).
//
// Ray key: the reference bins rays into six cardinal direction bins (Camera.cpp:42-46,
// Integrator.cpp:240-244) and then sorts each batch recursively by origin and direction
// (RaySort.cpp:7-118). Here one 32-bit key carries both: the direction octant in bits 31..29
// and a 29-bit Morton code of the origin quantised to (10, 10, 9) bits over the scene box.
// Hit key: CompareHit orders by (geomID, primID) as signed ints, so misses (-1) come first
// (RayUncompressed.h:54-60). Global triangle ids are geomID-major, so id + 1 (miss -> 0)
// reproduces that order.
#include "device_util.h"
#include "swr_internal.h"

namespace swr {
namespace {

__global__ void ray_keys_kernel(const float4* __restrict__ rays, uint32_t n, float3 lo, float3 scale,
                                uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 o = rays[2 * (size_t)i];
  const float4 d = rays[2 * (size_t)i + 1];
  keys[i] = ray_key(o.x, o.y, o.z, d.x, d.y, d.z, lo, scale);
}

__global__ void hit_keys_kernel(const uint4* __restrict__ hits, uint32_t n, uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = hits[i].w + 1u;
}

__global__ void mesh_face_keys_kernel(const uint32_t* __restrict__ mesh, const uint32_t* __restrict__ face, uint32_t n,
                                      uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = ((mesh[i] + 1u) << 21) | (face[i] & 0x1FFFFFu);
}

template <int W>
__global__ void gather_kernel(const uint32_t* __restrict__ in, const uint32_t* __restrict__ perm, uint32_t n,
                              uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t src = (size_t)perm[i] * W, dst = (size_t)i * W;
  if constexpr (W % 4 == 0) {
#pragma unroll
    for (int j = 0; j < W; j += 4)
      *reinterpret_cast<uint4*>(out + dst + j) = *reinterpret_cast<const uint4*>(in + src + j);
  } else if constexpr (W % 2 == 0) {
#pragma unroll
    for (int j = 0; j < W; j += 2)
      *reinterpret_cast<uint2*>(out + dst + j) = *reinterpret_cast<const uint2*>(in + src + j);
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) out[dst + j] = in[src + j];
  }
}

inline unsigned grid_of(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

void launch_ray_keys(const float4* rays, size_t n, float3 lo, float3 scale, uint32_t* keys, hipStream_t s) {
  if (!n) return;
  ray_keys_kernel<<<grid_of(n, 256), 256, 0, s>>>(rays, (uint32_t)n, lo, scale, keys);
}

void launch_hit_keys(const uint4* hits, size_t n, uint32_t* keys, hipStream_t s) {
  if (!n) return;
  hit_keys_kernel<<<grid_of(n, 256), 256, 0, s>>>(hits, (uint32_t)n, keys);
}

void launch_mesh_face_keys(const uint32_t* mesh, const uint32_t* face, size_t n, uint32_t* keys, hipStream_t s) {
  if (!n) return;
  mesh_face_keys_kernel<<<grid_of(n, 256), 256, 0, s>>>(mesh, face, (uint32_t)n, keys);
}

void launch_gather_u32(const uint32_t* in, const uint32_t* perm, size_t n, int words, uint32_t* out, hipStream_t s) {
  if (!n) return;
  const unsigned g = grid_of(n, 256);
  switch (words) {
#define SWR_GATHER_CASE(W) \
  case W:                  \
    gather_kernel<W><<<g, 256, 0, s>>>(in, perm, (uint32_t)n, out); \
    break;
    SWR_GATHER_CASE(1) SWR_GATHER_CASE(2) SWR_GATHER_CASE(3) SWR_GATHER_CASE(4) SWR_GATHER_CASE(5)
    SWR_GATHER_CASE(6) SWR_GATHER_CASE(7) SWR_GATHER_CASE(8) SWR_GATHER_CASE(9) SWR_GATHER_CASE(10)
    SWR_GATHER_CASE(11) SWR_GATHER_CASE(12) SWR_GATHER_CASE(13) SWR_GATHER_CASE(14) SWR_GATHER_CASE(15)
    SWR_GATHER_CASE(16)
#undef SWR_GATHER_CASE
    default:
      fail(SWR_E_INVALID, "gather: rec_words must be in 1..16");
  }
}

}  // namespace swr

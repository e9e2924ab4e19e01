// device_util.h -- small device helpers shared by the kernels: the ray binning key, the
// counter-based RNG and wave-aggregated queue appends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace swr {

// ---- ray binning key --------------------------------------------------------------------
// Spread the low 10 bits of v to every third bit.
__device__ inline uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Bits from the top: x9 y9 z8 x8 y8 z7 ... x1 y1 z0 x0 y0 (the 30-bit code of (x, y, 2z) >> 1).
__device__ inline uint32_t morton29(uint32_t qx, uint32_t qy, uint32_t qz) {
  return ((spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz << 1)) >> 1;
}

__device__ inline uint32_t quantise(float o, float lo, float scale, float qmax) {
  float f = floorf((o - lo) * scale);
  f = fminf(fmaxf(f, 0.f), qmax);
  return (uint32_t)f;
}

// Direction octant in bits 31..29 over a 29-bit Morton code of the quantised origin. The
// reference's cardinal bins (Camera.cpp:42-46, Integrator.cpp:240-244) group rays by dominant
// axis and sign only; the octant plus origin code also orders rays spatially inside a bin.
__device__ inline uint32_t ray_key(float ox, float oy, float oz, float dx, float dy, float dz, float3 lo, float3 scale) {
  const uint32_t oct = (dx < 0.f ? 1u : 0u) | (dy < 0.f ? 2u : 0u) | (dz < 0.f ? 4u : 0u);
  return (oct << 29) | morton29(quantise(ox, lo.x, scale.x, 1023.f), quantise(oy, lo.y, scale.y, 1023.f),
                                quantise(oz, lo.z, scale.z, 511.f));
}

// ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC 2011): 10 rounds of two 32x32 -> 64-bit products with Weyl-sequence key bumps. -------
struct Uniform4 {
  float r[4];
};

__device__ inline Uniform4 philox_uniform4(uint32_t key0, uint32_t key1, uint32_t c0, uint32_t c1, uint32_t c2,
                                           uint32_t c3) {
  uint32_t k0 = key0, k1 = key1;
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    if (round) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  // top 24 bits -> [0, 1)
  Uniform4 u;
  u.r[0] = (float)(c0 >> 8) * (1.0f / 16777216.0f);
  u.r[1] = (float)(c1 >> 8) * (1.0f / 16777216.0f);
  u.r[2] = (float)(c2 >> 8) * (1.0f / 16777216.0f);
  u.r[3] = (float)(c3 >> 8) * (1.0f / 16777216.0f);
  return u;
}

// Counter word 2 of a shading draw: ((depth + 1) << 8) | slot; the camera uses word 2 = 0.
__device__ inline uint32_t shade_word(uint32_t depth, uint32_t slot) { return ((depth + 1u) << 8) | slot; }

// ---- wave-aggregated append --------------------------------------------------------------
// Every active lane calls this in converged control flow; lanes with pred get consecutive
// slots from one atomic per wave. Returns the slot (meaningful only where pred holds).
__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return 0;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  return base + (uint32_t)__popcll(mask & lt);
}

// ---- block-aggregated append -------------------------------------------------------------
// Every thread of a kBlock-thread workgroup calls this once, in converged control flow; the
// workgroup reserves its slots with a single atomic. Same-address device atomics serialise at
// about 11 ns each on MI355X (measured), so a per-wave atomic on a 5M-ray wave costs ~1 ms
// where one per 1024-thread block costs ~60 us. s_scratch: 2 * (kBlock / 64) + 1 words of LDS.
// counter[0] (low word) hands out the slots of lanes with pred; counter[1] (high word) only counts
// the lanes with tally, in the same 64-bit atomic. The low word must stay below 2^32.
template <int kBlock>
__device__ inline uint32_t block_append(uint64_t* counter, bool pred, bool tally, uint32_t* s_scratch) {
  constexpr int kWaves = kBlock / 64;
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  const uint64_t mask = __ballot(pred);
  const uint64_t tmask = __ballot(tally);
  if (lane == 0) {
    s_scratch[w] = (uint32_t)__popcll(mask);
    s_scratch[kWaves + 1 + w] = (uint32_t)__popcll(tmask);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0, tallied = 0;
    for (int i = 0; i < kWaves; ++i) {
      const uint32_t c = s_scratch[i];
      s_scratch[i] = total;
      total += c;
      tallied += s_scratch[kWaves + 1 + i];
    }
    const uint64_t add = (uint64_t)total | ((uint64_t)tallied << 32);
    s_scratch[kWaves] = add ? (uint32_t)atomicAdd((unsigned long long*)counter, (unsigned long long)add) : 0u;
  }
  __syncthreads();
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint32_t slot = s_scratch[kWaves] + s_scratch[w] + (uint32_t)__popcll(mask & lt);
  __syncthreads();
  return slot;
}

}  // namespace swr

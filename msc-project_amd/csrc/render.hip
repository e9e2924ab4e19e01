// render.hip -- the sorted wavefront loop on gfx950 (Pathtracer::process,
// src/core/Pathtracer.cpp:425-493), shared by the plain render, the batched loop and the
// serving loop.
//
// The reference generates every camera ray up front, bins rays into six direction files,
// then repeatedly pops a batch, sorts it, traces it with Embree, sorts the hits and shades
// them, writing continuation rays back to the bins (Pathtracer.cpp:451-475). Here one device
// queue of rays replaces the bins and batches, and one wave is:
//   wave begin  continuations of the last wave + room for camera rays -> wave size (device);
//   raygen      tops the queue up with camera rays for the next global path indices, which map
//               to (job, sample id) through the job range table (Camera.cpp:6-59);
//   ray sort    scene + octant + Morton keys (written by raygen and shade) -> stable radix sort;
//   trace       closest hit in ray-key order -> (t, leaf) into the ray's record + hit key;
//   hit sort    stable radix sort by (scene, geomID, primID) (CompareHit, RayUncompressed.h:54-60);
//   shade       MIS shading in hit order (Integrator.cpp:12-255): emission and environment
//               radiance, one shadow ray per NEE sample at the same hit-order position, one
//               continuation per survivor appended with one atomic per 1024-lane workgroup;
//   shadow      any-hit traversal; unoccluded rays add their contribution;
//   wave end    the oldest job that still has a live path (job completion for the drivers).
// A path has one live ray, so within a launch no two lanes touch the same sample: radiance is
// added to the sample in place, in bounce order, as the reference's samples[sampleID] +=
// (Integrator.cpp:66-71, 171-173) -- no float atomics, and every job's image is a pure function
// of its inputs whatever else shares its waves. The tent/box resolve gathers each pixel's 3x3
// neighbourhood (TentFilter.cpp:17-61) from LDS tiles, sample index outermost.
//
// Every kernel reads the wave size from device memory, so one wave can be launched with sizes
// known on the host (the plain, host-paced loop) or captured once into a HIP graph sized for the
// queue capacity and replayed (the persistent loop of the batched and serving drivers).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "device_util.h"
#include "traverse.h"
#include "wavefront.h"

namespace swr {

namespace {

// ---- kernels --------------------------------------------------------------------------------
__global__ void wave_begin_kernel(WaveState* w) {
  const uint64_t c = w->counters;
  const uint32_t n_cont = (uint32_t)c;
  w->shadow_rays += c >> 32;
  const uint64_t left = w->total - w->cursor;
  const uint32_t k = (uint32_t)min((uint64_t)(w->cap - n_cont), left);
  const uint32_t n = n_cont + k;
  w->raygen_base = w->cursor;
  w->cursor += k;
  w->n = n;
  w->k = k;
  w->counters = 0;
  w->camera_rays += k;
  w->closest_rays += n;
  if (n) w->waves += 1;
  w->max_wave = max(w->max_wave, n);
}

// Index (in [range_lo, num_ranges)) of the range holding global path g.
__device__ inline uint32_t find_range(const JobRange* __restrict__ ranges, uint32_t mask, uint32_t lo, uint32_t hi_excl,
                                      uint64_t g) {
  uint32_t hi = hi_excl - 1;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (ranges[mid & mask].path_begin <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo & mask;
}

// Lanes of a wave take turns by `value` (a scene or a job): each turn serves the value of the
// lowest pending lane, read into a wave-uniform register so that what it indexes is read with
// scalar loads. Queues are sorted by scene first and camera rays come in job order, so a wave
// almost always finishes in one turn.
#define SWR_BY_UNIFORM(pending, value, u, ...)                                                    \
  for (;;) {                                                                                     \
    const uint64_t m_ = __ballot(pending);                                                       \
    if (m_ == 0) break;                                                                          \
    const uint32_t u = __builtin_amdgcn_readfirstlane(                                           \
        __shfl((int)(value), __ffsll((unsigned long long)m_) - 1, 64));                          \
    if ((pending) && (value) == u) {                                                             \
      pending = false;                                                                           \
      __VA_ARGS__                                                                                \
    }                                                                                            \
  }

// Camera rays for the wave's k new paths into its last k queue slots; resets each sample's
// radiance and records its film position (Camera.cpp:23-33).
// Paths are generated in sample-id order (a pixel's samples together: coherent camera rays), but
// per-sample state is stored sample-major (slot = (x * spp + k) * H + y), so a lane's own slot lies
// a column height away from its neighbour's. A workgroup whose paths are one run of sample ids of
// one job (nearly all of them) stages the film positions in LDS and writes the state in slot
// order instead: lane e takes the e-th sample of the workgroup's pixels in (sample, pixel) order,
// so that consecutive lanes write consecutive slots (measured on the terrain frame: raygen 3.85 ->
// 3.26 ms; 256- and 1024-lane workgroups 3.33 and 3.35 ms).
constexpr uint32_t kRaygenBlock = 512;
template <bool kOne>
__global__ __launch_bounds__(kRaygenBlock) void raygen_kernel(const FrameParams* __restrict__ jobs,
                                                              const JobRange* __restrict__ ranges, uint32_t range_mask,
                                                              const WaveState* __restrict__ w, RayRec* __restrict__ q,
                                                              uint32_t* __restrict__ keys,
                                                              float2* __restrict__ sample_xy,
                                                              float4* __restrict__ radiance) {
  __shared__ float2 s_xy[kRaygenBlock];
  __shared__ uint32_t s_first[3];
  const uint32_t k = w->k;
  if (blockIdx.x * blockDim.x >= k) return;
  // Workgroups go round robin to the 8 XCDs (one L2 each); a sample's radiance and film position
  // are part of a slot row that the neighbouring pixels' workgroups complete. Each XCD therefore
  // takes one contiguous run of the wave's active workgroups, so that most writes to a line meet
  // in one L2. Only which workgroup computes which rays changes, not what is written where.
  const uint32_t active = (k + blockDim.x - 1) / blockDim.x, per = active / 8u, rem = active % 8u;
  const uint32_t xcd = blockIdx.x % 8u;
  const uint32_t logical = xcd * per + min(xcd, rem) + blockIdx.x / 8u;
  const uint32_t t = logical * blockDim.x + threadIdx.x;
  bool pending = t < k;
  uint32_t job = 0, sid_local = 0, range = 0;
  if (pending) {
    const uint64_t g = w->raygen_base + t;
    range = find_range(ranges, range_mask, w->range_lo, w->num_ranges, g);
    const JobRange rg = ranges[range];
    job = rg.job;
    sid_local = (uint32_t)(g - rg.path_begin);  // index within the job; its sample id adds sid_base
  }
  const uint32_t slot = w->n - k + t;
  auto body = [&](uint32_t j0, bool state) {
    const FrameParams& p = jobs[j0];
    const uint32_t sid = p.sid_base + sid_local;
    V3 o, d;
    float sx, sy;
    camera_ray(p, sid, &o, &d, &sx, &sy);
    RayRec r;
    r.o = make_float4(o.x, o.y, o.z, 0.001f);    // RayDecompress.cpp:15
    r.d = make_float4(d.x, d.y, d.z, 100000.f);  // RayDecompress.cpp:16
    r.w = make_float4(1.f, 1.f, 1.f, 1.f);       // weight 1, lastPdf 1 (Camera.cpp:23-26)
    r.id = make_uint4(sid, j0 << 16, p.scene, kMiss);
    q[slot] = r;
    keys[slot] = queue_key(p, o, d);
    if (state) {
      const uint64_t s = p.sample_base + slot_of_sid(p, sid);
      sample_xy[s] = make_float2(sx, sy);
      radiance[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      s_xy[threadIdx.x] = make_float2(sx, sy);
    }
  };
  // The workgroup's pending lanes are a prefix of it; they hold one run of sample ids of one job
  // when they all lie in the first lane's range.
  const uint32_t cnt = min(kRaygenBlock, k - (t - threadIdx.x));
  if (threadIdx.x == 0) {
    s_first[0] = range;
    s_first[1] = sid_local;
    s_first[2] = job;
  }
  __syncthreads();
  if (__syncthreads_and(!pending || range == s_first[0])) {
    const uint32_t j0 = kOne ? 0u : __builtin_amdgcn_readfirstlane(s_first[2]);
    const FrameParams& p = jobs[j0];
    const uint32_t spp = p.spp, sid0 = p.sid_base + s_first[1];
    // Pixels (sample id / spp, column-major) the run touches, the first and last partly; lanes
    // take the samples of those pixels in (sample, pixel) order and skip those outside the run.
    // Only for a few pixels' worth of samples (else each lane writes its own).
    const uint32_t pix0 = sid0 / spp;
    const uint64_t npix = ((uint64_t)sid0 + cnt - 1u) / spp - pix0 + 1u;
    const bool staged = npix * spp <= 4u * kRaygenBlock;
    if (pending) body(j0, !staged);
    if (!staged) return;  // the whole workgroup
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < (uint32_t)(npix * spp); e += kRaygenBlock) {
      const uint32_t sid = (pix0 + e % (uint32_t)npix) * spp + e / (uint32_t)npix;
      if (sid < sid0 || sid - sid0 >= cnt) continue;
      const uint64_t s = p.sample_base + slot_of_sid(p, sid);
      sample_xy[s] = s_xy[sid - sid0];
      radiance[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  } else if (kOne) {
    if (pending) body(0u, true);
  } else {
    SWR_BY_UNIFORM(pending, job, j0, { body(j0, true); })
  }
}

// Stand-alone camera rays (swr_raygen): ray t as 8 floats, film position t.
__global__ void raygen_rays_kernel(FrameParams p, uint32_t sid0, uint32_t k, float4* __restrict__ rays,
                                   float2* __restrict__ sample_xy) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k) return;
  V3 o, d;
  float sx, sy;
  camera_ray(p, sid0 + t, &o, &d, &sx, &sy);
  rays[2 * (size_t)t] = make_float4(o.x, o.y, o.z, 0.001f);
  rays[2 * (size_t)t + 1] = make_float4(d.x, d.y, d.z, 100000.f);
  if (sample_xy) sample_xy[t] = make_float2(sx, sy);
}


// Closest hit of every queued ray, visited in ray-key order (perm). The hit (t, leaf) goes into
// the ray's record, the hit key (the scene's hit_base + the triangle's position in the BVH's
// leaf array + 1; a miss hit_base) into hit_keys. The leaf array is in Morton order of the
// triangle centroids, so a key truncated to its top bits still groups hits that lie close
// together (shading reads their records together, and their shadow rays, traced in hit order,
// start close together). With
// kRingFilter, only rays whose scene uses this launch's ring size are traced.
// The scene of ray j is the top scene_bits bits of its sorted key (sorted_keys[j] >> scene_shift):
// the ray record's scene word is overwritten with the hit t by whichever launch traces it.
//
// Tile modes (several scenes in the loop, whose keys carry the scene on top, so the rays of a
// scene are contiguous in sorted order and almost every tile of kTraceBlock rays holds one scene):
//   kTileAll:   every tile, each lane with its ray's scene (scene-uniform groups of lanes in turn);
//   kTileOne:   only tiles whose rays all belong to one scene, traversed with that scene's
//               constants in scalar registers and the one-scene traversal (as kOne);
//   kTileMixed: only the tiles that hold rays of several scenes, per lane as kTileAll; workgroup b
//               takes the tile holding the first ray of scene b + 1.
// kTileOne plus kTileMixed trace every ray exactly once.
constexpr int kTileAll = 0, kTileOne = 1, kTileMixed = 2;

// First position of [0, n) whose sorted key is >= v (n when none); every lane computes the same.
__device__ inline uint32_t lower_bound_u32(const uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <bool kOne, int kTile>
constexpr bool kUniformTrace = kOne || kTile == kTileOne;

// Held to 8 waves per SIMD (64 VGPRs): the one-scene traversal fits without spilling (63), and
// more waves hide more of the node-fetch latency; the per-lane multi-scene form, left for tiles at
// scene boundaries, spills 24 bytes.
template <int kShort, bool kRingFilter, bool kOne, int kTile = kTileAll>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(8))) void trace_queue_kernel(const DeviceScene* __restrict__ scenes,
                                                                  const uint8_t* __restrict__ scene_ring, int ring,
                                                                  const uint32_t* __restrict__ perm,
                                                                  const uint32_t* __restrict__ sorted_keys,
                                                                  uint32_t scene_shift,
                                                                  const WaveState* __restrict__ w,
                                                                  RayRec* __restrict__ q,
                                                                  uint32_t* __restrict__ hit_keys,
                                                                  uint32_t hit_shift,
                                                                  uint32_t* __restrict__ spill, uint32_t spill_stride) {
  __shared__ uint32_t s_ring[kShort * kTraceBlock];
  __shared__ uint32_t s_child[kUniformTrace<kOne, kTile> ? 4 * kTraceBlock : 1];
  const uint32_t n = w->n;
  uint32_t tile = blockIdx.x;
  if (kTile == kTileMixed) {
    const uint32_t b = blockIdx.x + 1u;  // scene whose first ray opens the boundary
    const uint32_t start = lower_bound_u32(sorted_keys, n, b << scene_shift);
    if (start >= n || start % kTraceBlock == 0u) return;  // no boundary inside a tile here
    const uint32_t prev = lower_bound_u32(sorted_keys, n, (b - 1u) << scene_shift);
    // The tile is taken by its lowest boundary: skip when scene b - 1 also starts inside it.
    if (b > 1u && prev / kTraceBlock == start / kTraceBlock && prev % kTraceBlock != 0u) return;
    tile = start / kTraceBlock;
  }
  if (tile * kTraceBlock >= n) return;
  uint32_t tile_scene = 0;
  if (kTile == kTileOne) {
    const uint32_t last = min(tile * kTraceBlock + kTraceBlock, n) - 1u;
    const uint32_t a = __builtin_amdgcn_readfirstlane(sorted_keys[tile * kTraceBlock] >> scene_shift);
    const uint32_t z = __builtin_amdgcn_readfirstlane(sorted_keys[last] >> scene_shift);
    if (a != z) return;  // a mixed tile
    if (kRingFilter && scene_ring[a] != (uint32_t)ring) return;
    tile_scene = a;
  }
  constexpr bool kUniform = kOne || kTile == kTileOne;
  const uint32_t j = tile * kTraceBlock + threadIdx.x;
  bool pending = j < n;
  uint32_t i = 0, scene = 0;
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f), d = o;
  if (pending) {
    i = perm[j];
    if (!kUniform) scene = sorted_keys[j] >> scene_shift;
    if (!kUniform && kRingFilter && scene_ring[scene] != (uint32_t)ring) pending = false;
  }
  if (pending) {
    o = q[i].o;
    d = q[i].d;
  }
  auto body = [&](uint32_t s0) {
    const RayState r = make_ray(o, d);
    constexpr bool kQ = kUniform && kQuantNodes;
    const void* nodes = kQ ? (const void*)scenes[s0].qnodes : (const void*)scenes[s0].nodes;
    auto run = [&](auto oct) {
      return traverse<kTraceBlock, kShort, false, kUniform, kUniform, kQ, kUniform && kLdsChildren, decltype(oct)::value>(
          nodes, scenes[s0].leaf_tris, r, SWR_MASK_ALL, s_ring + threadIdx.x, spill + (size_t)j * spill_stride,
          s_child + (kUniform && kLdsChildren ? threadIdx.x : 0u));
    };
    HitResult h;
    // Rays are sorted octant first (and camera-ray waves are one pixel's samples), so nearly
    // every wave has one direction octant: its node tests then take their near and far planes
    // without per-lane selects (terrain frame: closest-hit traversal 26.8-27.0 -> 26.4-26.5 ms;
    // the same for shadow rays changed nothing and is not used).
    const uint32_t oct = (r.ix < 0.f ? 1u : 0u) | (r.iy < 0.f ? 2u : 0u) | (r.iz < 0.f ? 4u : 0u);
    const uint32_t oct0 = __builtin_amdgcn_readfirstlane(oct);
    if constexpr (!kQ) {
      h = run(std::integral_constant<int, -1>{});
    } else if (__ballot(oct != oct0) == 0ull) {
      switch (oct0) {
        case 0: h = run(std::integral_constant<int, 0>{}); break;
        case 1: h = run(std::integral_constant<int, 1>{}); break;
        case 2: h = run(std::integral_constant<int, 2>{}); break;
        case 3: h = run(std::integral_constant<int, 3>{}); break;
        case 4: h = run(std::integral_constant<int, 4>{}); break;
        case 5: h = run(std::integral_constant<int, 5>{}); break;
        case 6: h = run(std::integral_constant<int, 6>{}); break;
        default: h = run(std::integral_constant<int, 7>{}); break;
      }
    } else {
      h = run(std::integral_constant<int, -1>{});
    }
    reinterpret_cast<uint2*>(&q[i].id)[1] = make_uint2(__float_as_uint(h.t), h.leaf);
    hit_keys[i] = (scenes[s0].hit_base + h.leaf + 1u) >> hit_shift;
  };
  if (kUniform) {
    if (pending) body(tile_scene);
  } else {
    SWR_BY_UNIFORM(pending, scene, s0, { body(s0); })
  }
}

struct ShadeOut {
  RayRec* next;          // continuation queue (compacted)
  uint32_t* next_keys;
  float4* shadow;        // 3 float4 per hit-order position m: origin + tag, direction + tmax, rgb + slot
  float4* radiance;      // per-sample radiance arena, accumulated in place
  uint32_t* block_min;   // per workgroup: smallest admission number among its continuations
};

// Workgroup sizes of the shading kernel: one untextured job is held to 64 VGPRs, so that two
// 1024-lane workgroups (8 waves per SIMD) share a CU (at 66 VGPRs and 7 waves shading took 10.1
// instead of 9.1 ms per terrain frame).
// The multi-job variant holds its lanes' job constants in vector registers; it is held to 80 VGPRs
// (6 waves per SIMD), so that three 512-lane workgroups share a CU.
constexpr int kShadeBlock = 1024;
constexpr int kShadeBlockMulti = 512;

// Integrator::operator() for every queued ray, in hit-key order (Integrator.cpp:12-255).
// Radiance is added in place: a path has one live ray, so no two lanes of a launch touch the
// same sample and each sample receives its contributions in bounce order, exactly as the
// reference's samples[sampleID] += (Integrator.cpp:66-71, 171-173), without atomics.
// Shadow record at position m: origin (tmin is the reference's 0.001, Integrator.cpp:154; w
// carries scene << 24 | sample slot bits 55..32), direction + tmax (< 0: no shadow ray), rgb
// contribution (already times throughput) + sample slot bits 31..0.
//
// kOne: the loop holds one job on one scene, whose constants are then wave-uniform (scalar
// loads). Otherwise each lane reads its job's constants and scene pointers.
// kTex: some scene of the loop has textured shaders (a separate instance, so that untextured
// loops do not carry the texture code's registers and stack).
template <bool kOne, int kBlock, bool kTex>
__global__ __launch_bounds__(kBlock, kOne ? (kTex ? 1 : 8) : 6) void shade_kernel(const DeviceScene* __restrict__ scenes,
                                                            const FrameParams* __restrict__ jobs,
                                                            const uint32_t* __restrict__ perm, WaveState* w,
                                                            const RayRec* __restrict__ q, ShadeOut out) {
  __shared__ uint32_t s_append[2 * (kBlock / 64) + 1];
  __shared__ uint32_t s_min[kBlock / 64];
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = m < w->n;
  if (m == 0u) w->shadow_n = w->n;  // for this wave's shadow pass (one record per position)
  bool has_c = false, has_s = false, has_n = false;
  V3 contrib = v3(0.f, 0.f, 0.f), srgb = v3(0.f, 0.f, 0.f);
  V3 P = v3(0.f, 0.f, 0.f), sdir = v3(0.f, 0.f, 0.f), ndir = v3(0.f, 0.f, 0.f), nw = v3(0.f, 0.f, 0.f);
  float sdist = 0.f, npdf = 0.f;
  uint32_t sid = 0, depth = 0, job = 0, scene = 0, seq = 0xFFFFFFFFu;
  uint64_t sample = 0;
  uint32_t nkey = 0;
  RayRec rec;
  if (valid) {
    rec = q[perm[m]];
    sid = rec.id.x;
    job = kOne ? 0u : rec.id.y >> 16;
    depth = rec.id.y & 0xFFFFu;
    if (!kOne) scene = jobs[job].scene;
  }
  auto body = [&](const FrameParams& p, const DeviceScene& s) {
    const uint32_t leaf = rec.id.w;
    seq = p.seq;
    sample = p.sample_base + slot_of_sid(p, sid);
    const V3 w = v3(rec.w.x, rec.w.y, rec.w.z);
    const float last_pdf = rec.w.w;
    if (leaf == kMiss) {
      // Misses add nothing in the reference (Integrator.cpp:26-27); a constant environment
      // adds throughput * radiance.
      if (p.env_nonzero) {
        has_c = true;
        contrib = v3(w.x * p.env.x, w.y * p.env.y, w.z * p.env.z);
      }
    } else {
      const float t = __uint_as_float(rec.id.z);
      const SWR_GLOBAL gf4* lt = global_f4(s.leaf_tris) + 3 * (size_t)leaf;
      const float4 e1 = gload(lt + 1), e2 = gload(lt + 2);
      const uint32_t material = __float_as_uint(e2.w);  // the geometry's light or shader
      const V3 dn = normalize(v3(rec.d.x, rec.d.y, rec.d.z));
      if (material & kLeafLight) {  // light hit (Integrator.cpp:37-75)
        const LightRec L = gload_rec(s.lights + (material & ~kLeafLight));
        const float c = dot(v3(L.n[0], L.n[1], L.n[2]), dn * -1.f);  // QuadLight::radiance
        const V3 le = c < 0.f ? v3(0.f, 0.f, 0.f) : v3(L.intensity, L.intensity, L.intensity);
        const float pdfa = 1.f / (L.sx * L.sy);
        float mis = 1.f;
        if (depth > 0) {
          const float pdfw = pdfa * (t * t) / fabsf(c);  // areaToAngleProbability (Common.h:35-38)
          mis = last_pdf / (last_pdf + pdfw * p.pick);   // misBalance (Common.h:45-48)
        }
        if (maxc(le) > kEps) {
          has_c = true;
          contrib = v3(le.x * mis * w.x, le.y * mis * w.y, le.z * mis * w.z);
        }
      } else {
        ShaderRec sh = gload_rec(s.shaders + material);
        if (kTex && sh.tex_len > 0) {
          // The texture coordinate of the hit (PolygonObject::texture, PolygonObject.cpp:45-63)
          // from the triangle's barycentrics, recomputed as traversal computed them.
          const float4 v0 = gload(lt);
          float bu, bv;
          tri_uv(rec.o, rec.d, v0, e1, e2, &bu, &bv);
          float tu = 0.f, tv = 0.f;
          if (s.uv) {
            const uint4 tr = s.tri[__float_as_uint(v0.w)];
            const float2 a = s.uv[tr.x], b = s.uv[tr.y], c = s.uv[tr.z];
            const float w0 = 1.f - bu - bv;
            tu = w0 * a.x + bu * b.x + bv * c.x;
            tv = w0 * a.y + bu * b.y + bv * c.y;
          }
          const V3 col = texture_colour(s, sh, tu, tv);
          sh.colour[0] = col.x;
          sh.colour[1] = col.y;
          sh.colour[2] = col.z;
        }
        // Ng = (v0 - v1) x (v2 - v0), the leaf's e1 x e2; N = -normalize(Ng) (Integrator.cpp:121-125)
        const V3 nrm = normalize(cross(v3(e1.x, e1.y, e1.z), v3(e2.x, e2.y, e2.z))) * -1.f;
        P = v3(rec.o.x, rec.o.y, rec.o.z) + dn * t;  // Integrator.cpp:126
        const V3 out_dir = dn * -1.f;
        const Uniform4 r = philox_uniform4(p.seed_lo, p.seed_hi, sid, p.iteration, shade_word(depth, 1u), 0u);
        // Next event estimation (Integrator.cpp:102-178)
        if (p.num_lights) {
          const uint32_t l = min(p.num_lights - 1u, (uint32_t)((float)p.num_lights * r.r[0]));
          V3 in, le;
          float dist, pdfw;
          const LightRec Ll = gload_rec(s.lights + l);
          light_illuminate(Ll, r.r[1], r.r[2], P, &in, &dist, &le, &pdfw);
          if (maxc(le) > kEps && isfinite(pdfw)) {
            const Eval e = bsdf_eval(sh, in, out_dir, nrm);
            if (maxc(e.f) > kEps) {
              const float mis = (pdfw * p.pick) / (pdfw * p.pick + e.pdf);
              const float k = mis * e.cos / (pdfw * p.pick);
              const V3 cn = v3(k * (le.x * e.f.x), k * (le.y * e.f.y), k * (le.z * e.f.z));
              has_s = true;
              sdir = in;
              sdist = dist;
              srgb = v3(cn.x * w.x, cn.y * w.y, cn.z * w.z);
            }
          }
        }
        // Russian roulette and continuation (Integrator.cpp:181-246)
        if ((int32_t)depth < p.max_depth) {
          const float cont = sh.type == SWR_SHADER_NULL ? 1.f : sh.reflectance;
          float qk = fminf(1.f, cont / p.threshold);
          if ((int32_t)depth < p.min_depth) qk = 1.f;
          if (!(r.r[3] > qk)) {
            const Uniform4 r2 = philox_uniform4(p.seed_lo, p.seed_hi, sid, p.iteration, shade_word(depth, 2u), 0u);
            V3 in;
            Eval e;
            if (bsdf_sample(sh, r2.r[0], r2.r[1], out_dir, nrm, &in, &e)) {
              const float k = e.cos / e.pdf;
              nw = v3(w.x * e.f.x * k / qk, w.y * e.f.y * k / qk, w.z * e.f.z * k / qk);
              if (maxc(nw) > 0.f) {
                has_n = true;
                ndir = in;
                npdf = e.pdf;
                nkey = queue_key(p, P, ndir);
              }
            }
          }
        }
      }
    }
  };
  if (valid) body(jobs[kOne ? 0u : job], scenes[kOne ? 0u : scene]);
  // Oldest job with a live continuation in this workgroup (job completion).
  uint32_t live = has_n ? seq : 0xFFFFFFFFu;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) live = min(live, (uint32_t)__shfl_xor((int)live, off, 64));
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = live;
  const uint32_t cslot = block_append<kBlock>(&w->counters, has_n, has_s, s_append);
  if (threadIdx.x == 0) {
    uint32_t b = 0xFFFFFFFFu;
    for (int i = 0; i < kBlock / 64; ++i) b = min(b, s_min[i]);
    out.block_min[blockIdx.x] = b;
  }
  if (has_n) {
    RayRec r;
    r.o = make_float4(P.x, P.y, P.z, 0.001f);
    r.d = make_float4(ndir.x, ndir.y, ndir.z, 100000.f);
    r.w = make_float4(nw.x, nw.y, nw.z, npdf);
    r.id = make_uint4(sid, rec.id.y + 1u, scene, kMiss);  // same job, depth + 1
    out.next[cslot] = r;
    out.next_keys[cslot] = nkey;
  }
  if (!valid) return;
  // The shadow ray stays at hit-order position m (no compaction); tmax < 0 marks none.
  float4* sr = out.shadow + 3 * (size_t)m;
  if (has_s) {
    sr[0] = make_float4(P.x, P.y, P.z, __uint_as_float((scene << 24) | (uint32_t)(sample >> 32)));
    sr[1] = make_float4(sdir.x, sdir.y, sdir.z, sdist);
    sr[2] = make_float4(srgb.x, srgb.y, srgb.z, __uint_as_float((uint32_t)sample));
  } else {
    sr[1] = make_float4(0.f, 0.f, 0.f, -1.f);
  }
  if (has_c) {
    float4 acc = out.radiance[sample];
    acc.x = acc.x + contrib.x;
    acc.y = acc.y + contrib.y;
    acc.z = acc.z + contrib.z;
    out.radiance[sample] = acc;
  }
}

// Shadow rays (mask 0x0FFFFFFF, Integrator.cpp:153-174), one per hit-order position; positions
// without a shadow ray carry tmax < 0. An unoccluded ray adds its contribution to its sample in
// place (one live path per sample, so no two lanes share a sample).
template <int kShort, bool kRingFilter, bool kOne>
__global__ __launch_bounds__(kTraceBlock) void shadow_kernel(const DeviceScene* __restrict__ scenes,
                                                             const uint8_t* __restrict__ scene_ring, int ring,
                                                             const WaveState* __restrict__ w,
                                                             const float4* __restrict__ shadow,
                                                             float4* __restrict__ radiance,
                                                             uint32_t* __restrict__ spill, uint32_t spill_stride) {
  __shared__ uint32_t s_ring[kShort * kTraceBlock];
  __shared__ uint32_t s_scenes[2];
  const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
  const uint32_t n = w->shadow_n;
  if (blockIdx.x * kTraceBlock >= n) return;
  const float4* sr = shadow + 3 * (size_t)i;
  float4 d = make_float4(0.f, 0.f, 0.f, -1.f);
  if (i < n) d = sr[1];
  bool pending = d.w >= 0.f;
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t tag = 0, scene = 0;
  if (pending) {
    o = sr[0];
    tag = __float_as_uint(o.w);
    scene = tag >> 24;
  }
  // Several scenes: shadow rays come in hit order, whose keys carry the scene on top, so most
  // workgroups hold rays of one scene; those trace them with that scene's constants in scalar
  // registers and the one-scene traversal, the others lane by lane.
  uint32_t lo = 0, hi = 0;
  if (!kOne && kSceneTiles) {
    if (threadIdx.x == 0) {
      s_scenes[0] = 0xFFFFFFFFu;
      s_scenes[1] = 0u;
    }
    __syncthreads();
    if (pending) {
      atomicMin(&s_scenes[0], scene);
      atomicMax(&s_scenes[1], scene);
    }
    __syncthreads();
    lo = __builtin_amdgcn_readfirstlane(s_scenes[0]);
    hi = __builtin_amdgcn_readfirstlane(s_scenes[1]);
  }
  auto finish = [&](const HitResult& h) {
    if (h.id == kMiss) {
      const float4 c = sr[2];
      const uint64_t s = ((uint64_t)(tag & 0xFFFFFFu) << 32) | __float_as_uint(c.w);
      float4 acc = radiance[s];
      acc.x = acc.x + c.x;
      acc.y = acc.y + c.y;
      acc.z = acc.z + c.z;
      radiance[s] = acc;
    }
  };
  if (kOne || (kSceneTiles && lo >= hi)) {  // one scene (lo > hi: no shadow ray in the workgroup)
    const uint32_t s0 = kOne ? 0u : lo;
    if (pending && (!kRingFilter || scene_ring[s0] == (uint32_t)ring)) {
      const RayState r = make_ray(make_float4(o.x, o.y, o.z, 0.001f), d);
      const void* nodes = kQuantNodes ? (const void*)scenes[s0].qnodes : (const void*)scenes[s0].nodes;
      if constexpr (kPostponeShadow)
        finish(traverse_postponed<kTraceBlock, kShort, true, false, true, kQuantNodes>(
            nodes, scenes[s0].leaf_tris, r, SWR_MASK_SHADOW, s_ring + threadIdx.x, spill + (size_t)i * spill_stride));
      else
        finish(traverse<kTraceBlock, kShort, true, false, true, kQuantNodes>(
            nodes, scenes[s0].leaf_tris, r, SWR_MASK_SHADOW, s_ring + threadIdx.x, spill + (size_t)i * spill_stride));
    }
  } else if (pending && (!kRingFilter || scene_ring[scene] == (uint32_t)ring)) {
    const RayState r = make_ray(make_float4(o.x, o.y, o.z, 0.001f), d);
    finish(traverse<kTraceBlock, kShort, true, false, false>(
        scenes[scene].nodes, scenes[scene].leaf_tris, r, SWR_MASK_SHADOW, s_ring + threadIdx.x, spill + (size_t)i * spill_stride));
  }
}

// After a wave: every job admitted before the oldest job that still has a live path (a
// continuation in the next queue, or paths not yet generated) has finished -- its last shadow
// rays were traced in this wave.
// The smallest admission number of a job that may still receive radiance: one with paths not
// generated yet, or with rays in the last wave (lagged: its shadow pass runs with the next wave),
// or else with continuations after the last wave.
__device__ inline uint32_t unfinished_min(const JobRange* __restrict__ ranges, uint32_t range_mask,
                                          const WaveState* w, bool lagged) {
  uint32_t m = lagged ? w->live_min_prev : w->live_min;
  if (lagged && w->k > 0u) m = min(m, ranges[find_range(ranges, range_mask, w->range_lo, w->num_ranges, w->raygen_base)].seq);
  if (w->cursor < w->total) m = min(m, ranges[find_range(ranges, range_mask, w->range_lo, w->num_ranges, w->cursor)].seq);
  return m;
}

__global__ __launch_bounds__(1024) void wave_end_kernel(const uint32_t* __restrict__ block_min, uint32_t nblocks,
                                                        const JobRange* __restrict__ ranges, uint32_t range_mask,
                                                        WaveState* w, bool lagged, WaveState* snap) {
  __shared__ uint32_t s_min[1024 / 64];
  uint32_t m = 0xFFFFFFFFu;
  for (uint32_t i = threadIdx.x; i < nblocks; i += 1024) m = min(m, block_min[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off, 64));
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < 1024 / 64; ++i) m = min(m, s_min[i]);
    w->live_min_prev = w->live_min;
    w->live_min = m;
    w->done_seq = unfinished_min(ranges, range_mask, w, lagged);
    if (snap) *snap = *w;  // the run-ahead loop's snapshot, straight into host memory
  }
}

// Stream-ordered admission of one job: its frame constants, its range and the new path total.
__global__ void admit_kernel(FrameParams* jobs, FrameParams p, JobRange* ranges, uint32_t range_index, JobRange r,
                             WaveState* w, uint32_t num_ranges, uint64_t total) {
  jobs[p.job] = p;
  ranges[range_index] = r;
  w->num_ranges = num_ranges;
  w->total = total;
}

__global__ void set_range_lo_kernel(WaveState* w, uint32_t range_lo) { w->range_lo = range_lo; }

// Job completion as of now: admissions since the last wave moved the cursor's job.
__global__ void done_kernel(const JobRange* __restrict__ ranges, uint32_t range_mask, WaveState* w, bool lagged) {
  w->done_seq = unfinished_min(ranges, range_mask, w, lagged);
}

__global__ void clear_kernel(float* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Tent / box resolve (TentFilter.cpp:17-61, BoxFilter.cpp:17-36). A workgroup owns a tile of
// kFilterCols columns x 64 rows of pixels, one lane per pixel (a wave per column). For each sample
// index k it stages sample k of the tile and its one-pixel halo, (kFilterCols + 2) x 66 slots, in
// LDS with row-contiguous loads (slots of one column and sample index are consecutive in y), and
// every lane adds its 3 x 3 neighbourhood from LDS. Each sample is read from HBM about
// (kFilterCols + 2) * 66 / (kFilterCols * 64) times instead of nine.
//
// Summation order per pixel: sample index k outer, then mask x, mask y (the reference sums mask
// x, mask y, then k). The terms and weights are the reference's; only the order of the float
// additions differs. Out-of-image neighbours are staged with weight 0, which adds exactly zero.
constexpr int kFilterCols = 4;
constexpr int kFilterRows = 64;
constexpr int kFilterBlock = kFilterCols * kFilterRows;
constexpr int kHaloRows = kFilterRows + 2, kHaloCols = kFilterCols + 2;
constexpr int kHaloSlots = kHaloRows * kHaloCols;
constexpr int kHaloPerLane = (kHaloSlots + kFilterBlock - 1) / kFilterBlock;

__global__ __launch_bounds__(kFilterBlock) void filter_kernel(FrameParams p, const float4* __restrict__ radiance,
                                                              const float2* __restrict__ sxy,
                                                              float* __restrict__ pixels, int accumulate) {
  __shared__ float4 s_rad[2][kHaloSlots];
  __shared__ float2 s_xy[2][kHaloSlots];
  const int W = (int)p.width, H = (int)p.height, spp = (int)p.spp;
  // Columns out_x0.. of the job's band; its samples are those of columns band_x0.. band_x1 - 1.
  const int bx0 = (int)p.band_x0, bx1 = (int)(p.band_x0 + p.band_cols);
  const int x0 = (int)p.out_x0 + (int)blockIdx.y * kFilterCols, y0 = (int)blockIdx.x * kFilterRows;
  const int lx = (int)threadIdx.x / kFilterRows, ly = (int)threadIdx.x % kFilterRows;
  const int px = x0 + lx, py = y0 + ly;
  const bool tent = p.filter != SWR_FILTER_BOX;

  // The halo slots this lane stages: slot e = c * kHaloRows + r is pixel (x0 - 1 + c, y0 - 1 + r).
  int64_t base[kHaloPerLane];  // slot of sample 0, or -1 outside the image / past the halo
#pragma unroll
  for (int j = 0; j < kHaloPerLane; ++j) {
    const int e = (int)threadIdx.x + j * kFilterBlock;
    const int c = e / kHaloRows, r = e - c * kHaloRows;
    const int qx = x0 - 1 + c, qy = y0 - 1 + r;
    const bool in = e < kHaloSlots && qx >= bx0 && qx < bx1 && qy >= 0 && qy < H &&
                    (tent || (c >= 1 && c <= kFilterCols && r >= 1 && r <= kFilterRows));
    base[j] = in ? (int64_t)p.sample_base + (int64_t)(qx - bx0) * spp * H + qy : -1;
  }
  float4 pr[kHaloPerLane];
  float2 pq[kHaloPerLane];
  auto fetch = [&](int k) {
#pragma unroll
    for (int j = 0; j < kHaloPerLane; ++j) {
      if (base[j] >= 0) {
        const size_t sl = (size_t)base[j] + (size_t)k * H;
        pr[j] = radiance[sl];
        pq[j] = sxy[sl];
      } else {
        pr[j] = make_float4(0.f, 0.f, 0.f, 0.f);
        pq[j] = make_float2(-1e30f, -1e30f);  // weight 0 for every pixel of the tile
      }
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int j = 0; j < kHaloPerLane; ++j) {
      const int e = (int)threadIdx.x + j * kFilterBlock;
      if (e < kHaloSlots) {
        s_rad[buf][e] = pr[j];
        s_xy[buf][e] = pq[j];
      }
    }
  };

  float s0 = 0.f, s1 = 0.f, s2 = 0.f, den = 0.f;
  const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
  if (spp > 0) {
    fetch(0);
    stage(0);
  }
  for (int k = 0; k < spp; ++k) {
    const int buf = k & 1;
    if (k + 1 < spp) fetch(k + 1);  // in flight while this sample index is summed
    __syncthreads();
    if (tent) {
#pragma unroll
      for (int mx = 0; mx < 3; ++mx)
#pragma unroll
        for (int my = 0; my < 3; ++my) {
          const int e = (lx + mx) * kHaloRows + ly + my;
          const float2 q = s_xy[buf][e];
          const float dx = fmaxf(0.f, 1.f - fabsf(q.x - cx));
          const float dy = fmaxf(0.f, 1.f - fabsf(q.y - cy));
          const float w = dx * dy;
          den = den + w;
          const float4 L = s_rad[buf][e];
          s0 += L.x * w;
          s1 += L.y * w;
          s2 += L.z * w;
        }
    } else {
      const float4 L = s_rad[buf][(lx + 1) * kHaloRows + ly + 1];
      s0 += L.x;
      s1 += L.y;
      s2 += L.z;
    }
    if (k + 1 < spp) stage(buf ^ 1);
  }
  if (px >= (int)p.out_x1 || py >= H) return;
  float* o = pixels + 3 * (p.pixel_base + (size_t)py * W + px);
  float v0, v1, v2;
  if (tent) {
    v0 = s0 / den;
    v1 = s1 / den;
    v2 = s2 / den;
  } else {
    v0 = s0 / (float)spp;
    v1 = s1 / (float)spp;
    v2 = s2 / (float)spp;
  }
  if (accumulate) {
    v0 = o[0] + v0;
    v1 = o[1] + v1;
    v2 = o[2] + v2;
  }
  o[0] = v0;
  o[1] = v1;
  o[2] = v2;
}

// Sample records (x, y, r, g, b) of one job in reference sample-id order (Image.h:10-14).
__global__ void pack_samples_kernel(FrameParams p, uint32_t sid0, size_t n, const float2* __restrict__ sxy,
                                    const float4* __restrict__ radiance, float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = p.sample_base + slot_of_sid(p, sid0 + (uint32_t)i);
  const float2 q = sxy[s];
  const float4 L = radiance[s];
  float* o = out + 5 * i;
  o[0] = q.x;
  o[1] = q.y;
  o[2] = L.x;
  o[3] = L.y;
  o[4] = L.z;
}

inline unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---- host side ------------------------------------------------------------------------------
static void axis_rotation(int axis, float deg, float m[9]) {
  const float a = deg * kPi180, c = std::cos(a), s = std::sin(a);
  std::memset(m, 0, 9 * sizeof(float));
  const int i = (axis + 1) % 3, j = (axis + 2) % 3;
  m[axis * 3 + axis] = 1.f;
  m[i * 3 + i] = c;
  m[i * 3 + j] = -s;
  m[j * 3 + i] = s;
  m[j * 3 + j] = c;
}

static void mat_mul(const float a[9], const float b[9], float out[9]) {
  float r[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
  std::memcpy(out, r, sizeof r);
}

// PinHoleCamera::construct / ThinLensCamera::construct (PinHoleCamera.cpp:5-14).
CameraRec make_camera(const swr_camera_desc& d) {
  CameraRec c{};
  float ry[9], rx[9], rz[9], t[9];
  axis_rotation(1, d.rotation[1], ry);
  axis_rotation(0, d.rotation[0], rx);
  axis_rotation(2, d.rotation[2], rz);
  mat_mul(ry, rx, t);
  mat_mul(t, rz, c.rot);
  std::memcpy(c.t, d.translation, sizeof c.t);
  const float n0 = c.rot[0], n1 = c.rot[3], n2 = c.rot[6];
  const float len = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  const float inv = len > 0.f ? 1.f / len : 1.f;
  c.n[0] = n0 * inv;
  c.n[1] = n1 * inv;
  c.n[2] = n2 * inv;
  c.type = d.type;
  c.focal_length = d.focal_length;
  c.focal_distance = d.focal_distance;
  c.aperture = d.aperture;
  return c;
}

uint64_t spp_of(const swr_render_desc& d) {
  return d.spp > 0 ? (uint64_t)d.spp : d.base > 0 ? (uint64_t)d.base * (uint64_t)d.base : 0u;
}

FrameParams make_params(const swr_render_desc& d, const DeviceScene* s) {
  if (d.width <= 0 || d.height <= 0) fail(SWR_E_INVALID, "render: width and height must be positive");
  if (d.base <= 0 && d.spp <= 0) fail(SWR_E_INVALID, "render: sample base or spp must be positive");
  if (d.base > 65535) fail(SWR_E_INVALID, "render: sample base above 65535 (spp = base^2 must fit 32 bits)");
  if ((uint64_t)d.width * (uint64_t)d.height * spp_of(d) > 0xFFFFFFFFull)
    fail(SWR_E_INVALID, "render: more than 2^32 - 1 samples in one iteration");
  if (d.sampler < 0 || d.sampler > 2 || d.filter < 0 || d.filter > 1 || d.camera.type < 0 || d.camera.type > 1)
    fail(SWR_E_INVALID, "render: unknown sampler, filter or camera type");
  if (!(d.threshold > 0.f)) fail(SWR_E_INVALID, "render: threshold must be positive");
  if (d.max_depth < 0 || d.max_depth > 1000000) fail(SWR_E_INVALID, "render: max depth out of range");
  if (d.camera.type == SWR_CAMERA_THINLENS && !(d.camera.aperture > 0.f && d.camera.focal_length > 0.f))
    fail(SWR_E_INVALID, "render: thin lens needs positive focal length and aperture");
  FrameParams p{};
  p.width = (uint32_t)d.width;
  p.height = (uint32_t)d.height;
  p.spp = (uint32_t)spp_of(d);
  p.base = d.base > 0 ? (uint32_t)d.base : 1u;
  if (d.sampler != SWR_SAMPLER_INDEPENDENT && p.base * p.base != p.spp)
    fail(SWR_E_INVALID, "render: stratified and grid samplers need spp = base^2 (Image.h:65)");
  p.sampler = d.sampler;
  p.filter = d.filter;
  p.min_depth = d.min_depth;
  p.max_depth = d.max_depth;
  p.threshold = d.threshold;
  p.iteration = d.iteration;
  p.seed_lo = (uint32_t)d.seed;
  p.seed_hi = (uint32_t)(d.seed >> 32);
  p.cam = make_camera(d.camera);
  if (s) {
    p.num_lights = s->num_lights;
    p.pick = s->num_lights ? 1.f / (float)s->num_lights : 0.f;
    p.env = v3(s->env[0], s->env[1], s->env[2]);
    p.env_nonzero = (s->env[0] != 0.f || s->env[1] != 0.f || s->env[2] != 0.f) ? 1 : 0;
    float lo[3], sc[3];
    const float bits[3] = {1024.f, 1024.f, 512.f};
    for (int k = 0; k < 3; ++k) {
      lo[k] = s->lo[k];
      const float ext = s->hi[k] - s->lo[k];
      sc[k] = ext > 0.f ? bits[k] / ext : 0.f;
    }
    p.key_lo = make_float3(lo[0], lo[1], lo[2]);
    p.key_scale = make_float3(sc[0], sc[1], sc[2]);
  }
  set_band(&p, 0, p.width);
  return p;
}

void set_band(FrameParams* p, uint32_t x0, uint32_t x1) {
  if (x0 >= x1 || x1 > p->width) fail(SWR_E_INVALID, "render: column band out of range");
  const BandLayout b = band_layout(p->width, p->height, p->spp, p->filter == SWR_FILTER_BOX, x0, x1);
  p->out_x0 = x0;
  p->out_x1 = x1;
  p->band_x0 = b.band_x0;
  p->band_cols = b.band_cols;
  p->sid_base = b.sid_base;
}

int bits_for(uint64_t max_value) {
  int b = 1;
  while (b < 32 && (max_value >> b) != 0) ++b;
  return b;
}


// ---- WaveLoop --------------------------------------------------------------------------------
void LoopBuffers::release(swr_context* ctx) {
  if (side) (void)hipStreamSynchronize(side);
  for (DeviceBuffer* b : {&queue[0], &queue[1], &keys[0], &keys[1], &hit_keys, &perm, &alt_k, &alt_v, &shadow, &spill,
                          &spill_shadow, &jobs, &ranges, &scene_tab, &scene_ring, &wave, &block_min, &radiance,
                          &sample_xy, &pixels, &staging})
    swr::release(ctx, *b);
  for (hipEvent_t* e : {&fork_ev[0], &fork_ev[1], &shadow_ev[0], &shadow_ev[1]})
    if (*e) (void)hipEventDestroy(*e), *e = nullptr;
  if (side) (void)hipStreamDestroy(side);
  side = nullptr;
  if (host_state) (void)hipHostFree(host_state);
  host_state = nullptr;
  if (host_wave) (void)hipHostFree(host_wave);
  host_wave = nullptr;
  for (auto& e : snap_ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  for (auto& row : stage_ev)
    for (auto& e : row)
      if (e) (void)hipEventDestroy(e), e = nullptr;
}

uint64_t WaveLoop::job_samples(const swr_render_desc& d) {
  return (uint64_t)std::max(d.width, 0) * (uint64_t)std::max(d.height, 0) * spp_of(d);
}
uint64_t WaveLoop::job_pixels(const swr_render_desc& d) {
  return (uint64_t)std::max(d.width, 0) * (uint64_t)std::max(d.height, 0);
}

WaveLoop::WaveLoop(swr_context* ctx, LoopBuffers& bufs, const std::vector<const swr_scene*>& scenes,
                   uint32_t max_jobs, uint32_t capacity)
    : ctx_(ctx), b_(bufs), stream_(ctx->stream), scenes_(scenes), max_jobs_(max_jobs), cap_(capacity) {
  if (scenes.empty() || scenes.size() > 256) fail(SWR_E_INVALID, "wavefront: 1 to 256 scenes");
  if (max_jobs == 0 || max_jobs > 65536) fail(SWR_E_INVALID, "wavefront: 1 to 65536 job slots");
  if (capacity < 64 || capacity > (1u << 28)) fail(SWR_E_INVALID, "wavefront: wave capacity must be 64 .. 2^28");
  uint32_t max_depth = 0;
  for (const swr_scene* s : scenes) {
    if (!s || s->ctx != ctx) fail(SWR_E_INVALID, "wavefront: scene belongs to another (or a destroyed) context");
    if (!s->committed) fail(SWR_E_INVALID, "wavefront: scene not committed");
    max_depth = std::max(max_depth, s->dev.max_depth);
  }
  scene_bits_ = scenes.size() > 1 ? bits_for(scenes.size() - 1) : 0;
  // Hit keys: scene s takes keys hit_base[s] .. hit_base[s] + num_tris (0 = miss), in scene order.
  std::vector<uint32_t> hit_base(scenes.size());
  uint64_t key_end = 0;
  for (size_t i = 0; i < scenes.size(); ++i) {
    hit_base[i] = (uint32_t)key_end;
    key_end += (uint64_t)scenes[i]->dev.num_tris + 1u;
    if (key_end > (1ull << 32)) fail(SWR_E_INVALID, "wavefront: the scenes' triangles exceed a 32-bit hit key");
  }
  const int hit_full = bits_for(key_end - 1u);
  const Tuning& tu = ctx->tuning;
  ray_bits_ = std::max(scene_bits_ + 3, std::min(tu.ray_key_bits, 32));
  hit_bits_ = tu.hit_key_bits > 0 ? std::min(tu.hit_key_bits, hit_full) : hit_full;
  hit_shift_ = hit_full - hit_bits_;
  passes_per_wave_ = (ray_bits_ + 7) / 8 + (hit_bits_ + 7) / 8;
  spill_stride_ = stack_need(max_depth);
  std::vector<uint8_t> ring_tab(scenes.size());
  for (size_t i = 0; i < scenes.size(); ++i) {
    const int ring = tu.trace_stack > 0 ? tu.trace_stack
                                        : choose_ring(scenes[i]->tuned_ring, scenes[i]->dev.max_depth);
    rings_.push_back(ring);
    ring_tab[i] = (uint8_t)ring;
    textured_ |= scenes[i]->textured;
    if (std::find(ring_set_.begin(), ring_set_.end(), ring) == ring_set_.end()) ring_set_.push_back(ring);
    if ((uint32_t)ring < stack_need(scenes[i]->dev.max_depth)) spill_ = true;
  }
  // Live ranges: jobs not fully generated, plus those a lagging retire_ranges still holds.
  range_cap_ = 1;
  while (range_cap_ < 4 * max_jobs) range_cap_ <<= 1;

  ensure(ctx, b_.queue[0], sizeof(RayRec) * (size_t)cap_);
  ensure(ctx, b_.queue[1], sizeof(RayRec) * (size_t)cap_);
  ensure(ctx, b_.keys[0], 4ull * cap_);
  ensure(ctx, b_.keys[1], 4ull * cap_);
  ensure(ctx, b_.hit_keys, 4ull * cap_);
  ensure(ctx, b_.perm, 4ull * cap_);
  ensure(ctx, b_.alt_k, 4ull * cap_);
  ensure(ctx, b_.alt_v, 4ull * cap_);
  ensure(ctx, b_.shadow, 48ull * cap_);
  if (spill_) ensure(ctx, b_.spill, 4ull * cap_ * spill_stride_);
  ensure(ctx, b_.jobs, sizeof(FrameParams) * (size_t)max_jobs_);
  ensure(ctx, b_.ranges, sizeof(JobRange) * (size_t)range_cap_);
  ensure(ctx, b_.scene_tab, sizeof(DeviceScene) * scenes.size());
  ensure(ctx, b_.scene_ring, scenes.size());
  ensure(ctx, b_.wave, sizeof(WaveState));
  ensure(ctx, b_.block_min, 4ull * blocks(cap_, std::min(kShadeBlock, kShadeBlockMulti)));
  sort_reserve(ctx, cap_);
  if (!b_.host_state) {
    SWR_HIP(hipHostMalloc((void**)&b_.host_state, 8 * sizeof(WaveState), hipHostMallocDefault));
    SWR_HIP(hipHostMalloc((void**)&b_.host_wave, 8 * sizeof(WaveState), hipHostMallocCoherent));
    for (auto& e : b_.snap_ev) SWR_HIP(hipEventCreate(&e));
    for (auto& row : b_.stage_ev)
      for (auto& e : row) SWR_HIP(hipEventCreate(&e));
  }
  std::vector<DeviceScene> tab;
  for (size_t i = 0; i < scenes.size(); ++i) {
    tab.push_back(scenes[i]->dev);
    tab.back().hit_base = hit_base[i];
  }
  SWR_HIP(hipMemcpyAsync(b_.scene_tab.ptr, tab.data(), sizeof(DeviceScene) * tab.size(), hipMemcpyHostToDevice,
                         stream_));
  SWR_HIP(hipMemcpyAsync(b_.scene_ring.ptr, ring_tab.data(), ring_tab.size(), hipMemcpyHostToDevice, stream_));
  WaveState w{};
  w.cap = cap_;
  w.done_seq = 0;
  w.live_min = w.live_min_prev = 0xFFFFFFFFu;
  SWR_HIP(hipMemcpyAsync(b_.wave.ptr, &w, sizeof w, hipMemcpyHostToDevice, stream_));
  SWR_HIP(hipStreamSynchronize(stream_));
}

WaveLoop::~WaveLoop() {
  (void)hipStreamSynchronize(stream_);  // replays may still be in flight
  if (b_.side) (void)hipStreamSynchronize(b_.side);  // and a shadow pass of the run-ahead loop
  for (int c = 0; c < 2; ++c) {
    if (exec_[c]) (void)hipGraphExecDestroy(exec_[c]);
    if (graph_[c]) (void)hipGraphDestroy(graph_[c]);
  }
}

FrameParams WaveLoop::params(uint32_t scene, const swr_render_desc& desc, uint64_t sample_base,
                             uint64_t pixel_base) const {
  if (scene >= scenes_.size()) fail(SWR_E_INVALID, "wavefront: scene index out of range");
  FrameParams p = make_params(desc, &scenes_[scene]->dev);
  if (p.max_depth > 65534) fail(SWR_E_INVALID, "render: max depth above 65534");
  p.ray_key_shift = (uint32_t)(32 - ray_bits_);
  p.key_scene_bits = (uint32_t)scene_bits_;
  p.key_prefix = scene_bits_ ? scene << (32 - scene_bits_) : 0u;
  p.scene = scene;
  p.sample_base = sample_base;
  p.pixel_base = pixel_base;
  return p;
}

void WaveLoop::set_arenas(float4* radiance, float2* sample_xy, float* pixels) {
  radiance_ = radiance;
  sxy_ = sample_xy;
  pixels_ = pixels;
}

uint32_t WaveLoop::admit(uint32_t slot, FrameParams p) {
  if (slot >= max_jobs_) fail(SWR_E_INVALID, "wavefront: job slot out of range");
  if (num_ranges_ - range_lo_ >= range_cap_) fail(SWR_E_BUSY, "wavefront: job range table full");
  p.job = slot;
  p.seq = seq_++;
  const uint64_t paths = band_samples(p);
  JobRange r{total_, slot, p.seq};
  total_ += paths;
  range_begin_.push_back(r.path_begin);
  ++num_ranges_;
  admit_kernel<<<1, 1, 0, stream_>>>(static_cast<FrameParams*>(b_.jobs.ptr), p, static_cast<JobRange*>(b_.ranges.ptr),
                                     (num_ranges_ - 1) & (range_cap_ - 1), r, static_cast<WaveState*>(b_.wave.ptr),
                                     num_ranges_, total_);
  SWR_HIP(hipGetLastError());
  return p.seq;
}

void WaveLoop::retire_ranges(uint64_t cursor) {
  // Keep the range holding `cursor` (and everything after it).
  uint32_t lo = range_lo_;
  while (lo + 1 < num_ranges_ && range_begin_[lo + 1] <= cursor) ++lo;
  if (lo != range_lo_) {
    range_lo_ = lo;
    set_range_lo_kernel<<<1, 1, 0, stream_>>>(static_cast<WaveState*>(b_.wave.ptr), range_lo_);
    SWR_HIP(hipGetLastError());
  }
}

void WaveLoop::enqueue_wave(int cur, bool exact, uint32_t n, uint32_t k, StageTimes* times, int ev_set,
                            bool overlap, WaveState* snap) {
  hipStream_t st = stream_;
  const uint32_t nb = exact ? n : cap_, kb = exact ? k : cap_;
  WaveState* w = static_cast<WaveState*>(b_.wave.ptr);
  const uint32_t* n_dev = exact ? nullptr : &w->n;
  RayRec* queue[2] = {static_cast<RayRec*>(b_.queue[0].ptr), static_cast<RayRec*>(b_.queue[1].ptr)};
  uint32_t* keys[2] = {static_cast<uint32_t*>(b_.keys[0].ptr), static_cast<uint32_t*>(b_.keys[1].ptr)};
  uint32_t* perm = static_cast<uint32_t*>(b_.perm.ptr);
  uint32_t* alt_k = static_cast<uint32_t*>(b_.alt_k.ptr);
  uint32_t* alt_v = static_cast<uint32_t*>(b_.alt_v.ptr);
  uint32_t* hit_keys = static_cast<uint32_t*>(b_.hit_keys.ptr);
  const FrameParams* jobs = static_cast<const FrameParams*>(b_.jobs.ptr);
  const JobRange* ranges = static_cast<const JobRange*>(b_.ranges.ptr);
  const DeviceScene* stab = static_cast<const DeviceScene*>(b_.scene_tab.ptr);
  const uint8_t* sring = static_cast<const uint8_t*>(b_.scene_ring.ptr);
  uint32_t* spill = static_cast<uint32_t*>(b_.spill.ptr);
  const uint32_t mask = range_cap_ - 1;
  const bool filter = ring_set_.size() > 1;
  auto mark = [&](int i) {
    if (times) SWR_HIP(hipEventRecord(b_.stage_ev[ev_set][i], st));
  };

  wave_begin_kernel<<<1, 1, 0, st>>>(w);
  mark(0);
  // Several scenes always carry scene bits in the ray key (the multi-scene trace reads them).
  const bool one_job = max_jobs_ == 1 && scenes_.size() == 1, one_scene = scenes_.size() == 1;
  if (kb) {
    if (one_job)
      raygen_kernel<true><<<checked_grid(blocks(kb, kRaygenBlock), kRaygenBlock, "raygen"), kRaygenBlock, 0, st>>>(jobs, ranges, mask, w, queue[cur], keys[cur], sxy_, radiance_);
    else
      raygen_kernel<false><<<checked_grid(blocks(kb, kRaygenBlock), kRaygenBlock, "raygen"), kRaygenBlock, 0, st>>>(jobs, ranges, mask, w, queue[cur], keys[cur], sxy_, radiance_);
  }
  mark(1);
  const uint32_t shade_blocks = nb ? checked_grid(blocks(nb, one_job ? kShadeBlock : kShadeBlockMulti),
                                                 one_job ? kShadeBlock : kShadeBlockMulti, "shade")
                                    : 0u;
  if (nb) {
    const uint32_t trace_grid = checked_grid(blocks(nb, kTraceBlock), kTraceBlock, "trace");
    uint32_t *rk = nullptr, *rperm = nullptr;
    sort_pairs(ctx_, keys[cur], perm, alt_k, alt_v, nb, ray_bits_, true, &rk, &rperm, st, n_dev);
    mark(2);
    if (overlap && shadow_deferred_) {
      // The previous wave's shadow pass touches only the shadow records, per-sample radiance and
      // a stack spill: on the second stream it uses a spill of its own, and this wave's kernels up
      // to its shading write none of the rest (its camera rays start samples of their own). It
      // starts with this wave's traversal, so its workgroups fill the traversal's tail.
      SWR_HIP(hipEventRecord(b_.fork_ev[deferred_cur_], st));
      SWR_HIP(hipStreamWaitEvent(b_.side, b_.fork_ev[deferred_cur_], 0));
      launch_shadow(b_.side, static_cast<uint32_t*>(b_.spill_shadow.ptr), cap_, times, deferred_set_);
      SWR_HIP(hipEventRecord(b_.shadow_ev[deferred_cur_], b_.side));
      shadow_deferred_ = false;
      shadow_pending_ = true;
      shadow_last_ = deferred_cur_;
    }
    const uint32_t scene_shift = (uint32_t)(ray_bits_ - scene_bits_);  // used only with scene_bits_ > 0
    for (int ring : ring_set_) {
      with_ring(ring, [&](auto kShort) {
        if (one_scene) {
          trace_queue_kernel<kShort.value, false, true><<<trace_grid, kTraceBlock, 0, st>>>(
              stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_, spill,
              spill_stride_);
        } else if (kSceneTiles) {
          // Scene-uniform tiles, then the tiles at scene boundaries (at most one per scene).
          const uint32_t nscenes = (uint32_t)scenes_.size();
          if (filter) {
            trace_queue_kernel<kShort.value, true, false, kTileOne><<<trace_grid, kTraceBlock, 0, st>>>(
                stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_,
                spill, spill_stride_);
            trace_queue_kernel<kShort.value, true, false, kTileMixed><<<checked_grid(nscenes - 1, kTraceBlock, "trace (scene boundaries)"), kTraceBlock, 0, st>>>(
                stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_,
                spill, spill_stride_);
          } else {
            trace_queue_kernel<kShort.value, false, false, kTileOne><<<trace_grid, kTraceBlock, 0, st>>>(
                stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_,
                spill, spill_stride_);
            trace_queue_kernel<kShort.value, false, false, kTileMixed><<<checked_grid(nscenes - 1, kTraceBlock, "trace (scene boundaries)"), kTraceBlock, 0, st>>>(
                stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_,
                spill, spill_stride_);
          }
        } else if (filter) {
          trace_queue_kernel<kShort.value, true, false><<<trace_grid, kTraceBlock, 0, st>>>(
              stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_, spill,
              spill_stride_);
        } else {
          trace_queue_kernel<kShort.value, false, false><<<trace_grid, kTraceBlock, 0, st>>>(
              stab, sring, ring, rperm, rk, scene_shift, w, queue[cur], hit_keys, hit_shift_, spill,
              spill_stride_);
        }
      });
    }
    mark(3);
    // The ray permutation is consumed; the hit sort reuses its buffers.
    uint32_t *hk = nullptr, *hperm = nullptr;
    uint32_t* spare = rperm == perm ? alt_v : perm;
    sort_pairs(ctx_, hit_keys, spare, keys[cur], rperm, nb, hit_bits_, true, &hk, &hperm, st, n_dev);
    mark(4);
    ShadeOut so;
    so.next = queue[cur ^ 1];
    so.next_keys = keys[cur ^ 1];
    so.shadow = static_cast<float4*>(b_.shadow.ptr);
    so.radiance = radiance_;
    so.block_min = static_cast<uint32_t*>(b_.block_min.ptr);
    // The previous wave's shadow pass adds to the same samples: it goes first.
    if (overlap && shadow_pending_) SWR_HIP(hipStreamWaitEvent(st, b_.shadow_ev[shadow_last_], 0));
    if (one_job && textured_)
      shade_kernel<true, kShadeBlock, true><<<shade_blocks, kShadeBlock, 0, st>>>(stab, jobs, hperm, w, queue[cur], so);
    else if (one_job)
      shade_kernel<true, kShadeBlock, false><<<shade_blocks, kShadeBlock, 0, st>>>(stab, jobs, hperm, w, queue[cur], so);
    else if (textured_)
      shade_kernel<false, kShadeBlockMulti, true><<<shade_blocks, kShadeBlockMulti, 0, st>>>(stab, jobs, hperm, w,
                                                                                             queue[cur], so);
    else
      shade_kernel<false, kShadeBlockMulti, false><<<shade_blocks, kShadeBlockMulti, 0, st>>>(stab, jobs, hperm, w,
                                                                                              queue[cur], so);
    mark(5);
    if (overlap) {
      // Launched by the next wave once its rays are sorted, beside its traversal, or by
      // flush_shadow after the last wave.
      shadow_deferred_ = true;
      deferred_cur_ = cur;
      deferred_set_ = ev_set;
    } else {
      launch_shadow(st, spill, nb, times, ev_set);
    }
  }
  wave_end_kernel<<<1, 1024, 0, st>>>(static_cast<const uint32_t*>(b_.block_min.ptr), nb ? shade_blocks : 0u, ranges,
                                      mask, w, overlap, snap);
  SWR_HIP(hipGetLastError());
}

void WaveLoop::run_paced(StageTimes* times) {
  lagged_ = false;  // the shadow pass runs in line
  WaveState* hs = b_.host_state;
  SWR_HIP(hipMemcpyAsync(hs, b_.wave.ptr, sizeof(WaveState), hipMemcpyDeviceToHost, stream_));
  SWR_HIP(hipStreamSynchronize(stream_));
  uint32_t n_cont = (uint32_t)hs->counters;
  uint64_t cursor = hs->cursor;
  while (true) {
    if (ctx_->cancel.load()) fail(SWR_E_CANCELLED, "render cancelled");
    const uint32_t k = (uint32_t)std::min<uint64_t>(cap_ - n_cont, total_ - cursor);
    const uint32_t n = n_cont + k;
    if (n == 0) break;
    enqueue_wave(cur_, true, n, k, times);
    SWR_HIP(hipMemcpyAsync(hs, b_.wave.ptr, sizeof(WaveState), hipMemcpyDeviceToHost, stream_));
    SWR_HIP(hipStreamSynchronize(stream_));
    add_stage_times(times, 0);
    n_cont = (uint32_t)hs->counters;
    cursor += k;
    cur_ ^= 1;
  }
}

void WaveLoop::launch_shadow(hipStream_t sst, uint32_t* sspill, uint32_t nb, StageTimes* times, int ev_set) {
  WaveState* w = static_cast<WaveState*>(b_.wave.ptr);
  const DeviceScene* stab = static_cast<const DeviceScene*>(b_.scene_tab.ptr);
  const uint8_t* sring = static_cast<const uint8_t*>(b_.scene_ring.ptr);
  const float4* shadow = static_cast<const float4*>(b_.shadow.ptr);
  const bool filter = ring_set_.size() > 1, one_scene = scenes_.size() == 1;
  // The pass reads its wave's size (shadow_n) on the device; the grid covers nb positions.
  const uint32_t grid = checked_grid(blocks(nb, kTraceBlock), kTraceBlock, "shadow");
  if (times) SWR_HIP(hipEventRecord(b_.stage_ev[ev_set][7], sst));
  for (int ring : ring_set_) {
    with_ring(ring, [&](auto kShort) {
      if (filter)
        shadow_kernel<kShort.value, true, false><<<grid, kTraceBlock, 0, sst>>>(stab, sring, ring, w, shadow,
                                                                                radiance_, sspill, spill_stride_);
      else if (one_scene)
        shadow_kernel<kShort.value, false, true><<<grid, kTraceBlock, 0, sst>>>(stab, sring, ring, w, shadow,
                                                                                radiance_, sspill, spill_stride_);
      else
        shadow_kernel<kShort.value, false, false><<<grid, kTraceBlock, 0, sst>>>(stab, sring, ring, w, shadow,
                                                                                 radiance_, sspill, spill_stride_);
    });
  }
  if (times) SWR_HIP(hipEventRecord(b_.stage_ev[ev_set][6], sst));
  SWR_HIP(hipGetLastError());
}

void WaveLoop::add_stage_times(StageTimes* times, int ev_set) {
  if (!times) return;
  SWR_HIP(hipEventSynchronize(b_.stage_ev[ev_set][6]));  // the shadow pass may run on the second stream
  auto ms = [&](int a, int b) {
    float v = 0.f;
    SWR_HIP(hipEventElapsedTime(&v, b_.stage_ev[ev_set][a], b_.stage_ev[ev_set][b]));
    return (double)v;
  };
  times->raygen += ms(0, 1);
  times->sort_rays += ms(1, 2);
  times->trace += ms(2, 3);
  times->sort_hits += ms(3, 4);
  times->shade += ms(4, 5);
  times->shadow += ms(7, 6);  // the shadow pass's own launches (it may run on the second stream)
}

void WaveLoop::drain_side() {
  if (shadow_pending_) SWR_HIP(hipStreamWaitEvent(stream_, b_.shadow_ev[shadow_last_], 0));
  shadow_pending_ = false;
}

void WaveLoop::flush_shadow(StageTimes* times) {
  if (shadow_deferred_)
    launch_shadow(stream_, static_cast<uint32_t*>(b_.spill.ptr), cap_, times, deferred_set_);
  shadow_deferred_ = false;
  drain_side();
}

void WaveLoop::ensure_side() {
  if (!b_.side) {
    SWR_HIP(hipStreamCreateWithFlags(&b_.side, hipStreamNonBlocking));
    for (hipEvent_t* e : {&b_.fork_ev[0], &b_.fork_ev[1], &b_.shadow_ev[0], &b_.shadow_ev[1]})
      SWR_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  if (spill_) ensure(ctx_, b_.spill_shadow, 4ull * cap_ * spill_stride_);
}

void WaveLoop::run_ahead(StageTimes* times) {
  const bool overlap = ctx_->tuning.shadow_overlap != 0;
  if (overlap) ensure_side();
  lagged_ = overlap;
  shadow_pending_ = shadow_deferred_ = false;
  int pending = -1, pending_set = 0;
  for (uint64_t wave = 0;; ++wave) {
    if (ctx_->cancel.load()) {
      shadow_deferred_ = false;
      drain_side();
      SWR_HIP(hipStreamSynchronize(stream_));
      fail(SWR_E_CANCELLED, "render cancelled");
    }
    // Sorts sized on the device skip their own epoch check.
    sort_epoch_guard(ctx_, 2ull * (uint64_t)passes_per_wave_, stream_);
    const int set = (int)(wave % 3);
    // The wave's last kernel writes its state into a host slot; an event marks it written.
    const int t = snap_next_;
    snap_next_ = (snap_next_ + 1) & 7;
    enqueue_wave(cur_, false, cap_, cap_, times, set, overlap, b_.host_wave + t);
    SWR_HIP(hipEventRecord(b_.snap_ev[t], stream_));
    cur_ ^= 1;
    if (pending >= 0) {
      SWR_HIP(hipEventSynchronize(b_.snap_ev[pending]));  // the previous wave has finished
      const WaveState& s = b_.host_wave[pending];
      add_stage_times(times, pending_set);
      if (s.cursor == s.total && s.total == total_ && (uint32_t)s.counters == 0) {
        pending = t;
        pending_set = set;
        break;
      }
    }
    pending = t;
    pending_set = set;
  }
  flush_shadow(times);
  SWR_HIP(hipStreamSynchronize(stream_));
  add_stage_times(times, pending_set);  // the empty wave after the last
}

void WaveLoop::capture() {
  if (exec_[0]) return;
  const bool overlap = ctx_->tuning.shadow_overlap != 0;
  if (overlap) ensure_side();  // no allocation inside a capture
  lagged_ = overlap;
  const uint64_t epoch_mirror = ctx_->sort.epoch;
  for (int c = 0; c < 2; ++c) {
    // Each captured wave starts the previous wave's shadow pass on a branch of its own beside
    // its traversal (the first replay's is empty: shadow_n starts at 0) and defers its own.
    shadow_deferred_ = overlap;
    deferred_cur_ = c ^ 1;
    shadow_pending_ = false;
    SWR_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_wave(c, false, cap_, cap_, nullptr, 0, overlap);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      ctx_->sort.epoch = epoch_mirror;
      shadow_deferred_ = shadow_pending_ = false;
      throw;
    }
    shadow_deferred_ = shadow_pending_ = false;
    SWR_HIP(hipStreamEndCapture(stream_, &graph_[c]));
    SWR_HIP(hipGraphInstantiate(&exec_[c], graph_[c], nullptr, nullptr, 0));
  }
  ctx_->sort.epoch = epoch_mirror;  // capturing issued no tags; replays do
}

void WaveLoop::launch_graph_waves(int waves) {
  capture();
  sort_epoch_guard(ctx_, (uint64_t)passes_per_wave_ * (uint64_t)(waves + 16), stream_);
  for (int i = 0; i < waves; ++i) {
    SWR_HIP(hipGraphLaunch(exec_[cur_], stream_));
    ctx_->sort.epoch += (uint64_t)passes_per_wave_;
    cur_ ^= 1;
  }
}

int WaveLoop::snapshot() {
  const int t = snap_next_;
  snap_next_ = (snap_next_ + 1) & 7;
  done_kernel<<<1, 1, 0, stream_>>>(static_cast<const JobRange*>(b_.ranges.ptr), range_cap_ - 1,
                                    static_cast<WaveState*>(b_.wave.ptr), lagged_);
  SWR_HIP(hipMemcpyAsync(b_.host_state + t, b_.wave.ptr, sizeof(WaveState), hipMemcpyDeviceToHost, stream_));
  SWR_HIP(hipEventRecord(b_.snap_ev[t], stream_));
  return t;
}

const WaveState& WaveLoop::wait_snapshot(int ticket) {
  SWR_HIP(hipEventSynchronize(b_.snap_ev[ticket]));
  return b_.host_state[ticket];
}

void WaveLoop::run_graphs(int check_every) {
  check_every = std::max(1, check_every);
  capture();
  int pending = -1;
  while (true) {
    if (ctx_->cancel.load()) {
      SWR_HIP(hipStreamSynchronize(stream_));
      fail(SWR_E_CANCELLED, "render cancelled");
    }
    launch_graph_waves(check_every);
    const int t = snapshot();
    if (pending >= 0) {
      const WaveState& s = wait_snapshot(pending);
      if (s.cursor == s.total && s.total == total_ && (uint32_t)s.counters == 0) break;
    }
    pending = t;
  }
  SWR_HIP(hipStreamSynchronize(stream_));
}

void WaveLoop::filter(const FrameParams& p, bool accumulate) {
  dim3 fgrid(blocks(p.height, kFilterRows), blocks(p.out_x1 - p.out_x0, kFilterCols));
  filter_kernel<<<fgrid, kFilterBlock, 0, stream_>>>(p, radiance_, sxy_, pixels_, accumulate ? 1 : 0);
  SWR_HIP(hipGetLastError());
}

void WaveLoop::clear_pixels(const FrameParams& p) {
  const size_t n = 3ull * p.width * p.height;
  clear_kernel<<<blocks(n), 256, 0, stream_>>>(pixels_ + 3 * p.pixel_base, n);
  SWR_HIP(hipGetLastError());
}

void WaveLoop::copy_out(const FrameParams& p, float* pixels_host, float* samples_host) {
  hipStream_t st = stream_;
  if (pixels_host)
    SWR_HIP(hipMemcpyAsync(pixels_host, pixels_ + 3 * p.pixel_base, 12ull * p.width * p.height,
                           hipMemcpyDeviceToHost, st));
  if (samples_host) copy_samples(p, samples_host);
  SWR_HIP(hipStreamSynchronize(st));
}

void WaveLoop::copy_samples(const FrameParams& p, float* samples_host) {
  hipStream_t st = stream_;
  const uint32_t per_col = p.height * p.spp, sid0 = p.out_x0 * per_col;
  const uint64_t n = (uint64_t)(p.out_x1 - p.out_x0) * per_col;
  ensure(ctx_, b_.staging, 20ull * n);
  pack_samples_kernel<<<blocks(n), 256, 0, st>>>(p, sid0, n, sxy_, radiance_, static_cast<float*>(b_.staging.ptr));
  SWR_HIP(hipGetLastError());
  SWR_HIP(hipMemcpyAsync(samples_host + 5ull * sid0, b_.staging.ptr, 20ull * n, hipMemcpyDeviceToHost, st));
  SWR_HIP(hipStreamSynchronize(st));
  release(ctx_, b_.staging);
}

LoopTotals WaveLoop::totals() {
  const WaveState& s = wait_snapshot(snapshot());
  LoopTotals t;
  t.camera_rays = s.camera_rays;
  t.closest_rays = s.closest_rays;
  t.shadow_rays = s.shadow_rays + (s.counters >> 32);
  t.waves = s.waves;
  t.max_wave = s.max_wave;
  return t;
}

void destroy_loop_buffers(swr_context* ctx) {
  if (!ctx->loop) return;
  ctx->loop->release(ctx);
  delete ctx->loop;
  ctx->loop = nullptr;
}

// ---- plain render ------------------------------------------------------------------------------
namespace {

uint32_t wave_capacity(const swr_render_desc& desc, uint64_t total) {
  const int exp = std::max(10, std::min(desc.bin_exponent > 0 ? desc.bin_exponent : 20, 26));
  return (uint32_t)std::min<uint64_t>(uint64_t(1) << exp, std::max<uint64_t>(total, 64));
}

// Image columns per band when a frame's per-sample state does not fit the context budget: the
// arenas may take what the budget leaves after everything else the context holds and the
// framebuffer; a band holds its columns plus the tent's neighbour columns. The whole frame
// (width columns) when there is no budget or it fits.
uint32_t band_columns(swr_context* ctx, const LoopBuffers& b, const FrameParams& p, bool samples_out) {
  const uint64_t per_sample = 24 + (samples_out ? 20 : 0);  // radiance + film position (+ staging)
  const uint64_t per_col = (uint64_t)p.height * p.spp;
  if (!ctx->budget) return p.width;
  const uint64_t held = b.radiance.bytes + b.sample_xy.bytes + b.pixels.bytes + b.staging.bytes;
  const uint64_t others = ctx->in_use - std::min(ctx->in_use, held);
  const uint64_t fb = 12ull * p.width * p.height + 256;
  if (ctx->budget <= others + fb) fail(SWR_E_BUDGET, "render: the framebuffer does not fit the memory budget");
  // room in samples (each allocation is rounded up to 256 bytes: 512 spare bytes per column)
  const uint64_t room_cols = (ctx->budget - others - fb) / (per_sample * per_col + 512);
  const uint64_t cols = band_width_for(p.width, per_col, p.filter == SWR_FILTER_BOX, room_cols * per_col);
  if (!cols) fail(SWR_E_BUDGET, "render: one column of samples does not fit the memory budget");
  return (uint32_t)cols;
}

void render_frame(swr_context* ctx, const swr_scene* scene, const swr_render_desc& desc, float* pixels_host,
                  float* samples_host, swr_stats* stats) {
  SWR_HIP(hipSetDevice(ctx->device));
  if (!ctx->loop) ctx->loop = new LoopBuffers();
  LoopBuffers& b = *ctx->loop;
  const uint64_t total = WaveLoop::job_samples(desc);
  WaveLoop loop(ctx, b, {scene}, 1, wave_capacity(desc, total));
  const FrameParams full = loop.params(0, desc, 0, 0);
  // Out of core: a frame whose per-sample state exceeds the budget is traced in column bands,
  // one after another through the same arenas (the reference spills its ray bins to disk
  // instead, DirectionalBins.cpp:37-115). Paths are pure functions of their sample ids and each
  // pixel is filtered from the same samples in the same order, so the image is the same.
  const uint32_t cols = band_columns(ctx, b, full, samples_host != nullptr);
  uint64_t band_max = total;
  if (cols < full.width) {
    band_max = band_max_samples(full.width, (uint64_t)full.height * full.spp, full.filter == SWR_FILTER_BOX, cols);
    release(ctx, b.radiance);  // make room for the smaller arenas
    release(ctx, b.sample_xy);
  }
  ensure(ctx, b.radiance, 16ull * band_max);
  ensure(ctx, b.sample_xy, 8ull * band_max);
  ensure(ctx, b.pixels, 12ull * full.width * full.height);
  loop.set_arenas(static_cast<float4*>(b.radiance.ptr), static_cast<float2*>(b.sample_xy.ptr),
                  static_cast<float*>(b.pixels.ptr));
  hipStream_t st = ctx->stream;
  hipEvent_t ev[4];
  for (auto& e : ev) SWR_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  StageTimes times;
  float filter_ms = 0.f;
  SWR_HIP(hipEventRecord(ev[0], st));
  for (uint32_t x0 = 0; x0 < full.width; x0 += cols) {
    FrameParams p = full;
    set_band(&p, x0, std::min(full.width, x0 + cols));
    loop.admit(0, p);
    loop.run_ahead(&times);
    loop.retire_ranges(loop.paths_admitted());
    SWR_HIP(hipEventRecord(ev[1], st));
    loop.filter(p);
    SWR_HIP(hipEventRecord(ev[2], st));
    if (samples_host) loop.copy_samples(p, samples_host);
    if (cols < full.width) {
      float f = 0.f;
      SWR_HIP(hipEventSynchronize(ev[2]));
      SWR_HIP(hipEventElapsedTime(&f, ev[1], ev[2]));
      filter_ms += f;
    }
  }
  loop.copy_out(full, pixels_host, nullptr);
  SWR_HIP(hipStreamSynchronize(st));
  if (cols >= full.width) SWR_HIP(hipEventElapsedTime(&filter_ms, ev[1], ev[2]));
  const LoopTotals t = loop.totals();
  if (stats) {
    swr_stats acc{};
    acc.camera_rays = t.camera_rays;
    acc.closest_rays = t.closest_rays;
    acc.shadow_rays = t.shadow_rays;
    acc.waves = (uint32_t)t.waves;
    acc.max_wave = t.max_wave;
    acc.ms_raygen = times.raygen;
    acc.ms_sort_rays = times.sort_rays;
    acc.ms_trace = times.trace;
    acc.ms_sort_hits = times.sort_hits;
    acc.ms_shade = times.shade;
    acc.ms_shadow = times.shadow;
    float tot = 0.f;
    SWR_HIP(hipEventElapsedTime(&tot, ev[0], ev[2]));
    acc.ms_filter = filter_ms;
    acc.ms_total = tot;
    *stats = acc;
  }
}

}  // namespace
}  // namespace swr

using swr::Error;

extern "C" {

int swr_raygen(swr_context* ctx, const swr_render_desc* desc, uint32_t sid0, size_t n, float* rays_dev,
               float* xy_dev) {
  if (!ctx || !desc) return SWR_E_INVALID;
  try {
    const swr::FrameParams p = swr::make_params(*desc, nullptr);
    const uint64_t total = (uint64_t)p.width * p.height * p.spp;
    if (n && (!rays_dev || (reinterpret_cast<uintptr_t>(rays_dev) & 15u)))
      swr::fail(SWR_E_INVALID, "raygen: rays must be a 16-byte aligned device buffer");
    if ((uint64_t)sid0 + n > total) swr::fail(SWR_E_INVALID, "raygen: sample ids beyond width*height*spp");
    if (n) {
      swr::Timer timer(ctx);
      swr::raygen_rays_kernel<<<swr::blocks(n), 256, 0, ctx->stream>>>(p, sid0, (uint32_t)n,
                                                                       reinterpret_cast<float4*>(rays_dev),
                                                                       reinterpret_cast<float2*>(xy_dev));
      SWR_HIP(hipGetLastError());
      timer.stop();
      SWR_HIP(hipStreamSynchronize(ctx->stream));
      timer.finish();
    }
    return SWR_OK;
  } catch (const Error& e) {
    ctx->err = e.msg;
    return e.code;
  }
}

// Coordinate descent over the launch shapes on a reduced render (about 1/16 of the samples of
// desc: base/4 for square samplers, spp/16 otherwise), best of two device timings each.
int swr_autotune(swr_context* ctx, const swr_scene* scene, const swr_render_desc* desc, swr_tuning* out) {
  if (!ctx || !scene || !desc || scene->ctx != ctx || !scene->committed) return SWR_E_INVALID;
  try {
    swr_render_desc d = *desc;
    if (d.spp > 0 && d.sampler == SWR_SAMPLER_INDEPENDENT)
      d.spp = std::max(1, d.spp / 16);
    else
      d.base = std::max(1, d.base / 4), d.spp = 0;
    auto time_it = [&](const swr::Tuning& t, int reps = 3) {
      ctx->tuning = t;
      double best = 1e30;
      for (int rep = 0; rep < reps; ++rep) {
        swr_stats st{};
        swr::render_frame(ctx, scene, d, nullptr, nullptr, &st);
        best = std::min(best, st.ms_total);
      }
      return best;
    };
    swr::Tuning best = ctx->tuning;
    swr::render_frame(ctx, scene, d, nullptr, nullptr, nullptr);  // allocations outside the timings
    double best_ms = time_it(best);
    // A candidate replaces the best so far only when it is faster by more than run-to-run noise
    // on the reduced frame (1%); key widths are tried from the fewest sort passes up, so that
    // ties keep the cheaper sort.
    auto consider = [&](const swr::Tuning& t) {
      const double ms = time_it(t);
      if (ms < best_ms * 0.99) best_ms = ms, best = t;
    };
    for (int alg : {SWR_SORT_ONESWEEP, SWR_SORT_REDUCE_SCAN})
      for (int kpt : {8, 12, 16, 20, 24}) {
        swr::Tuning t = best;
        t.sort_algorithm = alg;
        t.sort_kpt = kpt;
        consider(t);
      }
    for (int rb : {8, 12, 16, 20, 24, 32}) {
      swr::Tuning t = best;
      t.ray_key_bits = rb;
      consider(t);
    }
    const int full = swr::bits_for((uint64_t)scene->dev.num_tris);
    for (int hb : {8, full - 12, full - 6, 0}) {
      if (hb < 0 || (hb > 0 && hb < 4) || (hb == 8 && full - 12 <= 8)) continue;
      swr::Tuning t = best;
      t.hit_key_bits = hb;
      consider(t);
    }
    // No stack ring is cheaper than another by construction, so the fastest one is kept (best of
    // five runs each), without the noise margin: with the margin the first ring that beat the
    // default stayed even where a later one was faster (terrain: ring 8 kept, 16 faster by 1.5%
    // of closest-hit traversal on the full frame).
    {
      swr::Tuning pick = best;
      double pick_ms = time_it(best, 5);
      for (int ring : {8, 16, 32, 64}) {
        swr::Tuning t = best;
        t.trace_stack = ring;
        const double ms = time_it(t, 5);
        if (ms < pick_ms) pick_ms = ms, pick = t;
      }
      best = pick;
    }
    ctx->tuning = best;
    if (out) swr_get_tuning(ctx, out);
    return SWR_OK;
  } catch (const Error& e) {
    (void)hipStreamSynchronize(ctx->stream);
    ctx->err = e.msg;
    return e.code;
  }
}

int swr_render(swr_context* ctx, const swr_scene* scene, const swr_render_desc* desc, float* pixels, float* samples,
               swr_stats* stats) {
  if (!ctx || !scene || !desc) return SWR_E_INVALID;
  try {
    if (scene->ctx != ctx) swr::fail(SWR_E_INVALID, "render: scene belongs to another (or a destroyed) context");
    if (!scene->committed) swr::fail(SWR_E_INVALID, "render: scene not committed");
    swr::render_frame(ctx, scene, *desc, pixels, samples, stats);
    return SWR_OK;
  } catch (const Error& e) {
    (void)hipStreamSynchronize(ctx->stream);
    ctx->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    (void)hipStreamSynchronize(ctx->stream);
    ctx->err = e.what();
    return SWR_E_INVALID;
  }
}

}  // extern "C"


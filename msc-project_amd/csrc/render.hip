// render.hip -- the sorted wavefront loop of one render on gfx950 (Pathtracer::process,
// src/core/Pathtracer.cpp:425-493).
//
// The reference generates every camera ray up front, bins rays into six direction files,
// then repeatedly pops a batch, sorts it, traces it with Embree, sorts the hits and shades
// them, writing continuation rays back to the bins (Pathtracer.cpp:451-475). Here one device
// queue of 2^bin_exponent rays replaces the bins and batches:
//   raygen      tops the queue up with camera rays for the next sample ids (Camera.cpp:6-59);
//   ray sort    octant + Morton keys (written by raygen and shade) -> stable radix sort;
//   trace       closest hit in ray-key order -> (t, leaf) into the ray's record + hit key;
//   hit sort    stable radix sort by (geomID, primID) (CompareHit, RayUncompressed.h:54-60);
//   shade       MIS shading in hit order (Integrator.cpp:12-255): emission and environment
//               radiance, one shadow ray per NEE sample at the same hit-order position, one
//               continuation per survivor appended with one atomic per 1024-lane workgroup;
//   shadow      any-hit traversal; unoccluded rays add their contribution.
// A path has one live ray, so within a launch no two lanes touch the same sample: radiance is
// added to the sample in place, in bounce order, as the reference's samples[sampleID] +=
// (Integrator.cpp:66-71, 171-173) -- no float atomics, and the image is a pure function of the
// inputs whatever the wave composition. The tent/box resolve gathers each pixel's 3x3
// neighbourhood (TentFilter.cpp:17-61) from LDS tiles, sample index outermost.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_util.h"
#include "traverse.h"

namespace swr {

void commit_scene(swr_scene* s);

namespace {

const float kPi = 3.14159265358979323846f;  // Constants.h:10-15
const float kInvPi = 0.31830988618379067154f;
const float kPi180 = 0.01745329251994329577f;
const float kEps = 1.19209290e-07f;         // FLT_EPSILON

// ---- small vector helpers (no contraction: -ffp-contract=off) ----------------------------
struct V3 {
  float x, y, z;
};
__host__ __device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ inline float norm(V3 a) { return sqrtf(dot(a, a)); }
__device__ inline V3 normalize(V3 a) {
  const float n = norm(a);
  return n > 0.f ? a * (1.f / n) : a;
}
__device__ inline float maxc(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ inline V3 apply(const float* m, V3 v) {
  return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// ---- frame constants -----------------------------------------------------------------------
struct CameraRec {
  float rot[9];  // Ry * Rx * Rz (PinHoleCamera.cpp:7-11)
  float t[3];
  float n[3];    // R * x-hat, normalised (PinHoleCamera.cpp:13)
  int32_t type;
  float focal_length, focal_distance, aperture;
};

struct FrameParams {
  uint32_t width, height, spp, base;
  int32_t sampler, filter, min_depth, max_depth;
  float threshold;
  uint32_t iteration;
  uint32_t seed_lo, seed_hi;
  uint32_t num_lights;
  float pick;           // 1 / number of lights
  float3 key_lo, key_scale;
  uint32_t ray_key_shift;  // the ray sort orders by the top (32 - shift) key bits
  V3 env;
  int32_t env_nonzero;
  CameraRec cam;
};

// ---- BSDFs (LambertShader.cpp, NullShader.cpp; Glossy = GGX reflection) -----------------
struct Eval {
  V3 f;
  float cos;
  float pdf;
};

// Local frame of the samplers (LambertShader.cpp:65-67).
__device__ inline void shading_frame(V3 n, V3* s, V3* t) {
  const V3 temp = (n.x > 0.99f) ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
  *s = normalize(cross(temp, n));
  *t = normalize(cross(n, *s));
}

__device__ inline float ggx_d(float a2, float ch) {
  const float k = ch * ch * (a2 - 1.f) + 1.f;
  return a2 / (kPi * k * k);
}
__device__ inline float ggx_g1(float a2, float c) { return 2.f * c / (c + sqrtf(a2 + (1.f - a2) * c * c)); }

__device__ inline Eval bsdf_eval(const ShaderRec& sh, V3 in, V3 out, V3 n) {
  Eval e;
  if (sh.type == SWR_SHADER_NULL) {  // NullShader::evaluate (NullShader.cpp:25-38)
    e.f = v3(0.f, 0.f, 0.f);
    e.cos = 1.f;
    e.pdf = 1.f * kInvPi * 0.5f;
    return e;
  }
  const float ci = dot(n, in), co = dot(n, out);
  const V3 col = v3(sh.colour[0], sh.colour[1], sh.colour[2]);
  if (sh.type == SWR_SHADER_LAMBERT) {  // LambertShader::evaluate (LambertShader.cpp:27-47)
    e.f = (ci < 0.f || co < 0.f) ? v3(0.f, 0.f, 0.f) : col * sh.reflectance * kInvPi;
    e.cos = ci < 0.f ? 0.f : ci;
    e.pdf = e.cos * kInvPi;
    return e;
  }
  e.cos = ci < 0.f ? 0.f : ci;
  if (ci <= 0.f || co <= 0.f) {
    e.f = v3(0.f, 0.f, 0.f);
    e.pdf = 0.f;
    return e;
  }
  const float a = sh.roughness, a2 = a * a;
  const V3 h = normalize(in + out);
  const float ch = dot(n, h), oh = dot(out, h);
  const float D = ggx_d(a2, ch);
  const float G = ggx_g1(a2, ci) * ggx_g1(a2, co);
  e.f = col * (sh.reflectance * D * G / (4.f * ci * co));
  e.pdf = oh > 0.f ? D * ch / (4.f * oh) : 0.f;
  return e;
}

// Samples an input direction; false when the sample carries no energy (the path ends).
__device__ inline bool bsdf_sample(const ShaderRec& sh, float u1, float u2, V3 out, V3 n, V3* in, Eval* e) {
  if (sh.type == SWR_SHADER_NULL) return false;  // NullShader::sample: zero weight
  V3 s, t;
  shading_frame(n, &s, &t);
  const V3 col = v3(sh.colour[0], sh.colour[1], sh.colour[2]);
  if (sh.type == SWR_SHADER_LAMBERT) {  // LambertShader::sample (LambertShader.cpp:49-73)
    const float azimuth = 2.f * kPi * u1;
    const float zenith = acosf(sqrtf(u2));
    // sphericalToCartesianCoord (Common.h:28-33)
    const float x = cosf(azimuth) * sinf(zenith);
    const float y = sinf(azimuth) * sinf(zenith);
    const float z = cosf(zenith);
    *in = t * x + s * y + n * z;
    e->f = col * sh.reflectance * kInvPi;
    e->cos = dot(n, *in);
    e->pdf = e->cos * kInvPi;
    return e->pdf > 0.f;
  }
  const float a = sh.roughness, a2 = a * a;
  const float phi = 2.f * kPi * u1;
  const float tan2 = a2 * u2 / (1.f - u2);
  const float ch = 1.f / sqrtf(1.f + tan2);
  const float sh_ = sqrtf(fmaxf(0.f, 1.f - ch * ch));
  const V3 h = t * (cosf(phi) * sh_) + s * (sinf(phi) * sh_) + n * ch;
  const float oh = dot(out, h);
  *in = h * (2.f * oh) - out;
  *e = bsdf_eval(sh, *in, out, n);
  return e->pdf > 0.f && maxc(e->f) > 0.f;
}

// QuadLight::illuminate (QuadLight.cpp:41-67).
__device__ inline void light_illuminate(const LightRec& L, float ux, float uy, V3 p, V3* dir, float* dist, V3* rad,
                                        float* pdfw) {
  const V3 local = v3(0.f, L.sx * (ux - 0.5f), L.sy * (uy - 0.5f));
  const V3 smp = v3(L.t[0], L.t[1], L.t[2]) + apply(L.rot, local);
  const V3 ln = v3(L.n[0], L.n[1], L.n[2]);
  float c = dot(normalize(p - smp), ln);
  *rad = c < 0.f ? v3(0.f, 0.f, 0.f) : v3(L.intensity, L.intensity, L.intensity);
  c = c < 0.f ? 0.f : c;
  const V3 dv = smp - p;
  *dir = normalize(dv);
  *dist = norm(dv);
  *pdfw = (dot(dv, dv) / c) / (L.sx * L.sy);
}

// ---- camera (Camera.cpp:15-38, PinHoleCamera.cpp:16-32, ThinLensCamera.cpp:16-56) -------
__device__ inline void camera_ray(const FrameParams& p, uint32_t sid, V3* org, V3* dir, float* sx, float* sy) {
  const uint32_t W = p.width, H = p.height, spp = p.spp;
  const uint32_t x = sid / (H * spp);
  const uint32_t rem = sid - x * H * spp;
  const uint32_t y = rem / spp;
  const uint32_t k = rem - y * spp;
  const Uniform4 r = philox_uniform4(p.seed_lo, p.seed_hi, sid, p.iteration, 0u, 0u);
  const uint32_t base = p.base > 0 ? p.base : 1u;
  const float cell = 1.f / (float)base;
  const uint32_t ix = k / base, iy = k % base;
  float jx, jy;
  if (p.sampler == SWR_SAMPLER_INDEPENDENT) {  // IndependentSampler.cpp:11-12
    jx = r.r[0];
    jy = r.r[1];
  } else if (p.sampler == SWR_SAMPLER_GRID) {  // GridSampler.cpp:13-14
    jx = (float)ix * cell + 0.5f * cell;
    jy = (float)iy * cell + 0.5f * cell;
  } else {  // StratifiedSampler.cpp:13-14
    jx = (float)ix * cell + r.r[0] * cell;
    jy = (float)iy * cell + r.r[1] * cell;
  }
  *sx = (float)x + jx;
  *sy = (float)y + jy;
  const float fx = (((float)x + jx) * 2.f - (float)W) / (float)W * 36.f;  // Camera.cpp:34
  const float fy = (((float)y + jy) * 2.f - (float)H) / (float)W * 36.f;  // Camera.cpp:35
  const CameraRec& c = p.cam;
  const V3 ct = v3(c.t[0], c.t[1], c.t[2]), cn = v3(c.n[0], c.n[1], c.n[2]);
  const V3 nodal = ct + cn * c.focal_length;
  const V3 film = ct + apply(c.rot, v3(0.f, fy, fx));
  if (c.type == SWR_CAMERA_PINHOLE) {
    *org = nodal;
    *dir = normalize(nodal - film);
    return;
  }
  // Thin lens: rejection-sampled aperture disk, rays through the focal-plane point.
  const float ratio = c.focal_distance / c.focal_length;
  const V3 focal = film + (nodal - film) * ratio;
  float ax = 0.f, ay = 0.f;
  float cx0 = r.r[2] - 0.5f, cy0 = r.r[3] - 0.5f;
  for (uint32_t attempt = 0; attempt < 64; ++attempt) {
    if (attempt) {
      const Uniform4 q = philox_uniform4(p.seed_lo, p.seed_hi, sid, p.iteration, 0u, attempt);
      cx0 = q.r[0] - 0.5f;
      cy0 = q.r[1] - 0.5f;
    }
    if (sqrtf(cx0 * cx0 + cy0 * cy0) < 0.5f) {
      ax = cx0;
      ay = cy0;
      break;
    }
  }
  const float scale = c.focal_length / c.aperture;
  const V3 ap = ct + apply(c.rot, v3(0.f, ax * scale, ay * scale));
  *org = ap + cn * c.focal_length;
  *dir = normalize(focal - *org);
}

// ---- queue records ----------------------------------------------------------------------------
// One queued ray is one 64-byte line: origin + tmin, direction + tmax, throughput + lastPdf, and
// (sample id, depth, hit t, hit leaf). Traversal fills the last two words, so shading gathers
// each ray with a single line read.
struct alignas(16) RayRec {
  float4 o;
  float4 d;
  float4 w;
  uint4 id;
};
static_assert(sizeof(RayRec) == 64, "ray record must be one 64-byte line");

// Per-sample state is stored column-major by sample: slot = (x * spp + k) * H + y, so that the
// 64 lanes of a resolve wave (consecutive y) read consecutive slots.
__host__ __device__ inline uint32_t sample_slot(uint32_t x, uint32_t y, uint32_t k, uint32_t H, uint32_t spp) {
  return (x * spp + k) * H + y;
}
__device__ inline uint32_t slot_of_sid(const FrameParams& p, uint32_t sid) {
  const uint32_t x = sid / (p.height * p.spp);
  const uint32_t rem = sid - x * p.height * p.spp;
  const uint32_t y = rem / p.spp;
  return sample_slot(x, y, rem - y * p.spp, p.height, p.spp);
}

// ---- kernels --------------------------------------------------------------------------------
// Camera rays for sample ids [sid0, sid0 + k) into queue slots [slot0, slot0 + k); resets each
// sample's radiance and records its film position (Camera.cpp:23-33).
__global__ void raygen_kernel(FrameParams p, uint32_t sid0, uint32_t k, uint32_t slot0, RayRec* __restrict__ q,
                              uint32_t* __restrict__ keys, float2* __restrict__ sample_xy,
                              float4* __restrict__ radiance) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k) return;
  const uint32_t sid = sid0 + t, slot = slot0 + t;
  V3 o, d;
  float sx, sy;
  camera_ray(p, sid, &o, &d, &sx, &sy);
  RayRec r;
  r.o = make_float4(o.x, o.y, o.z, 0.001f);    // RayDecompress.cpp:15
  r.d = make_float4(d.x, d.y, d.z, 100000.f);  // RayDecompress.cpp:16
  r.w = make_float4(1.f, 1.f, 1.f, 1.f);       // weight 1, lastPdf 1 (Camera.cpp:23-26)
  r.id = make_uint4(sid, 0u, 0u, kMiss);
  q[slot] = r;
  keys[slot] = ray_key(o.x, o.y, o.z, d.x, d.y, d.z, p.key_lo, p.key_scale) >> p.ray_key_shift;
  const uint32_t s = slot_of_sid(p, sid);
  sample_xy[s] = make_float2(sx, sy);
  radiance[s] = make_float4(0.f, 0.f, 0.f, 0.f);
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

// Closest hit of every queue slot, visited in ray-key order (perm). The hit (t, leaf) goes into
// the slot's record, the hit key (global triangle id + 1, misses 0) into keys.
template <int kShort>
__global__ __launch_bounds__(kTraceBlock) void trace_queue_kernel(const Node4* __restrict__ nodes,
                                                                  const float4* __restrict__ leaf_tris,
                                                                  const uint32_t* __restrict__ perm, uint32_t n,
                                                                  RayRec* __restrict__ q,
                                                                  uint32_t* __restrict__ hit_keys, uint32_t hit_shift,
                                                                  uint32_t* __restrict__ spill, uint32_t spill_stride) {
  __shared__ uint32_t s_ring[kShort * kTraceBlock];
  const uint32_t j = blockIdx.x * kTraceBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  const RayState r = make_ray(q[i].o, q[i].d);
  const HitResult h = traverse<kTraceBlock, kShort, false>(nodes, leaf_tris, r, SWR_MASK_ALL, s_ring + threadIdx.x,
                                                           spill + (size_t)j * spill_stride);
  reinterpret_cast<uint2*>(&q[i].id)[1] = make_uint2(__float_as_uint(h.t), h.leaf);
  hit_keys[i] = (h.id + 1u) >> hit_shift;
}

struct ShadeOut {
  RayRec* next;          // continuation queue (compacted)
  uint32_t* next_keys;
  float4* shadow_rays;   // 2 float4 per hit-order position m; tmax < 0 marks "no shadow ray"
  float4* shadow_col;    // rgb contribution (already times throughput), sample slot bits in w
  float4* radiance;      // per-sample radiance, accumulated in place
  uint64_t* counters;    // low word: continuations, high word: shadow rays
};

constexpr int kShadeBlock = 1024;

// Integrator::operator() for every slot, in hit-key order (Integrator.cpp:12-255).
// Radiance is added in place: a path has one live ray, so no two lanes of a launch touch the
// same sample and each sample receives its contributions in bounce order, exactly as the
// reference's samples[sampleID] += (Integrator.cpp:66-71, 171-173), without atomics.
__global__ __launch_bounds__(kShadeBlock) void shade_kernel(DeviceScene s, FrameParams p,
                                                            const uint32_t* __restrict__ perm, uint32_t n,
                                                            const RayRec* __restrict__ q, ShadeOut out) {
  __shared__ uint32_t s_append[2 * (kShadeBlock / 64) + 1];
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = m < n;
  bool has_c = false, has_s = false, has_n = false;
  V3 contrib = v3(0.f, 0.f, 0.f), srgb = v3(0.f, 0.f, 0.f);
  V3 P = v3(0.f, 0.f, 0.f), sdir = v3(0.f, 0.f, 0.f), ndir = v3(0.f, 0.f, 0.f), nw = v3(0.f, 0.f, 0.f);
  float sdist = 0.f, npdf = 0.f;
  uint32_t sid = 0, depth = 0;
  if (valid) {
    const RayRec rec = q[perm[m]];
    sid = rec.id.x;
    depth = rec.id.y;
    const uint32_t leaf = rec.id.w;
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
      const float4 e1 = s.leaf_tris[3 * (size_t)leaf + 1], e2 = s.leaf_tris[3 * (size_t)leaf + 2];
      const GeomRec g = s.geoms[__float_as_uint(e2.w)];
      const V3 dn = normalize(v3(rec.d.x, rec.d.y, rec.d.z));
      if (g.light >= 0) {  // light hit (Integrator.cpp:37-75)
        const LightRec& L = s.lights[g.light];
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
        const ShaderRec sh = s.shaders[g.shader];
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
          light_illuminate(s.lights[l], r.r[1], r.r[2], P, &in, &dist, &le, &pdfw);
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
              }
            }
          }
        }
      }
    }
  }
  const uint32_t cslot = block_append<kShadeBlock>(out.counters, has_n, has_s, s_append);
  if (has_n) {
    RayRec r;
    r.o = make_float4(P.x, P.y, P.z, 0.001f);
    r.d = make_float4(ndir.x, ndir.y, ndir.z, 100000.f);
    r.w = make_float4(nw.x, nw.y, nw.z, npdf);
    r.id = make_uint4(sid, depth + 1u, 0u, kMiss);
    out.next[cslot] = r;
    out.next_keys[cslot] = ray_key(P.x, P.y, P.z, ndir.x, ndir.y, ndir.z, p.key_lo, p.key_scale) >> p.ray_key_shift;
  }
  if (!valid) return;
  const uint32_t sample = slot_of_sid(p, sid);
  // The shadow ray stays at hit-order position m (no compaction); tmax < 0 marks none.
  out.shadow_rays[2 * (size_t)m] = make_float4(P.x, P.y, P.z, 0.001f);
  out.shadow_rays[2 * (size_t)m + 1] = make_float4(sdir.x, sdir.y, sdir.z, has_s ? sdist : -1.f);
  out.shadow_col[m] = make_float4(srgb.x, srgb.y, srgb.z, __uint_as_float(sample));
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
template <int kShort>
__global__ __launch_bounds__(kTraceBlock) void shadow_kernel(const Node4* __restrict__ nodes,
                                                             const float4* __restrict__ leaf_tris, uint32_t n,
                                                             const float4* __restrict__ rays,
                                                             const float4* __restrict__ col,
                                                             float4* __restrict__ radiance,
                                                             uint32_t* __restrict__ spill, uint32_t spill_stride) {
  __shared__ uint32_t s_ring[kShort * kTraceBlock];
  const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  const float4 d = rays[2 * (size_t)i + 1];
  if (!(d.w >= 0.f)) return;
  const RayState r = make_ray(rays[2 * (size_t)i], d);
  const HitResult h = traverse<kTraceBlock, kShort, true>(nodes, leaf_tris, r, SWR_MASK_SHADOW, s_ring + threadIdx.x,
                                                          spill + (size_t)i * spill_stride);
  if (h.id != kMiss) return;
  const float4 c = col[i];
  const uint32_t s = __float_as_uint(c.w);
  float4 acc = radiance[s];
  acc.x = acc.x + c.x;
  acc.y = acc.y + c.y;
  acc.z = acc.z + c.z;
  radiance[s] = acc;
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
                                                              float* __restrict__ pixels) {
  __shared__ float4 s_rad[2][kHaloSlots];
  __shared__ float2 s_xy[2][kHaloSlots];
  const int W = (int)p.width, H = (int)p.height, spp = (int)p.spp;
  const int x0 = (int)blockIdx.y * kFilterCols, y0 = (int)blockIdx.x * kFilterRows;
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
    const bool in = e < kHaloSlots && qx >= 0 && qx < W && qy >= 0 && qy < H && (tent || (c >= 1 && c <= kFilterCols && r >= 1 && r <= kFilterRows));
    base[j] = in ? (int64_t)qx * spp * H + qy : -1;
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
  if (px >= W || py >= H) return;
  float* o = pixels + 3 * ((size_t)py * W + px);
  if (tent) {
    o[0] = s0 / den;
    o[1] = s1 / den;
    o[2] = s2 / den;
  } else {
    o[0] = s0 / (float)spp;
    o[1] = s1 / (float)spp;
    o[2] = s2 / (float)spp;
  }
}

// Sample records (x, y, r, g, b) in reference sample-id order (Image.h:10-14).
__global__ void pack_samples_kernel(FrameParams p, size_t n, const float2* __restrict__ sxy,
                                    const float4* __restrict__ radiance, float* __restrict__ out) {
  const size_t sid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sid >= n) return;
  const uint32_t s = slot_of_sid(p, (uint32_t)sid);
  const float2 q = sxy[s];
  const float4 L = radiance[s];
  float* o = out + 5 * sid;
  o[0] = q.x;
  o[1] = q.y;
  o[2] = L.x;
  o[3] = L.y;
  o[4] = L.z;
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

inline unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// ---- host side ------------------------------------------------------------------------------
void axis_rotation(int axis, float deg, float m[9]) {
  const float a = deg * kPi180, c = std::cos(a), s = std::sin(a);
  std::memset(m, 0, 9 * sizeof(float));
  const int i = (axis + 1) % 3, j = (axis + 2) % 3;
  m[axis * 3 + axis] = 1.f;
  m[i * 3 + i] = c;
  m[i * 3 + j] = -s;
  m[j * 3 + i] = s;
  m[j * 3 + j] = c;
}

void mat_mul(const float a[9], const float b[9], float out[9]) {
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

uint32_t spp_of(const swr_render_desc& d) { return d.spp > 0 ? (uint32_t)d.spp : (uint32_t)(d.base * d.base); }

FrameParams make_params(const swr_render_desc& d, const DeviceScene* s) {
  if (d.width <= 0 || d.height <= 0) fail(SWR_E_INVALID, "render: width and height must be positive");
  if (d.base <= 0 && d.spp <= 0) fail(SWR_E_INVALID, "render: sample base or spp must be positive");
  if (d.sampler < 0 || d.sampler > 2 || d.filter < 0 || d.filter > 1 || d.camera.type < 0 || d.camera.type > 1)
    fail(SWR_E_INVALID, "render: unknown sampler, filter or camera type");
  if (!(d.threshold > 0.f)) fail(SWR_E_INVALID, "render: threshold must be positive");
  if (d.max_depth < 0 || d.max_depth > 1000000) fail(SWR_E_INVALID, "render: max depth out of range");
  if (d.camera.type == SWR_CAMERA_THINLENS && !(d.camera.aperture > 0.f && d.camera.focal_length > 0.f))
    fail(SWR_E_INVALID, "render: thin lens needs positive focal length and aperture");
  FrameParams p{};
  p.width = (uint32_t)d.width;
  p.height = (uint32_t)d.height;
  p.spp = spp_of(d);
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
  return p;
}

int bits_for(uint64_t max_value) {
  int b = 1;
  while (b < 32 && (max_value >> b) != 0) ++b;
  return b;
}

enum Ws {
  kQueueA, kQueueB, kKeysA, kKeysB, kHitKeys, kPerm, kSortAltK, kSortAltV,
  kShadowRays, kShadowCol, kSpill,
  kRadiance, kSampleXY, kCounters, kPixels
};

template <class T>
T* ws(swr_context* ctx, Ws w) {
  return static_cast<T*>(ctx->render_ws[w].ptr);
}

struct StageClock {
  hipEvent_t ev[8];
  StageClock() {
    for (auto& e : ev) SWR_HIP(hipEventCreate(&e));
  }
  ~StageClock() {
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  float ms(int a, int b) {
    float v = 0.f;
    SWR_HIP(hipEventElapsedTime(&v, ev[a], ev[b]));
    return v;
  }
};

void render_frame(swr_context* ctx, const swr_scene* scene, const swr_render_desc& desc, float* pixels_host,
                  float* samples_host, swr_stats* stats) {
  SWR_HIP(hipSetDevice(ctx->device));
  const DeviceScene& s = scene->dev;
  FrameParams p = make_params(desc, &s);
  const int ray_bits = std::max(3, std::min(ctx->tuning.ray_key_bits, 32));
  p.ray_key_shift = (uint32_t)(32 - ray_bits);
  const uint64_t total = (uint64_t)p.width * p.height * p.spp;
  if (total > 0xFFFFFFFFull) fail(SWR_E_INVALID, "render: more than 2^32 - 1 samples in one iteration");
  const int exp = std::max(10, std::min(desc.bin_exponent > 0 ? desc.bin_exponent : 20, 24));
  const uint32_t cap = (uint32_t)std::min<uint64_t>(uint64_t(1) << exp, std::max<uint64_t>(total, 64));
  hipStream_t st = ctx->stream;

  // Queues (two for the current and the next wave) and per-sample state.
  auto need = [&](Ws w, size_t bytes) { ensure(ctx, ctx->render_ws[w], bytes); };
  need(kQueueA, sizeof(RayRec) * (size_t)cap);
  need(kQueueB, sizeof(RayRec) * (size_t)cap);
  need(kKeysA, 4ull * cap);
  need(kKeysB, 4ull * cap);
  need(kHitKeys, 4ull * cap);
  need(kPerm, 4ull * cap);
  need(kSortAltK, 4ull * cap);
  need(kSortAltV, 4ull * cap);
  need(kShadowRays, 32ull * cap);
  need(kShadowCol, 16ull * cap);
  need(kRadiance, 16ull * total);
  need(kSampleXY, 8ull * total);
  need(kCounters, 64);
  need(kPixels, 12ull * p.width * p.height);
  sort_reserve(ctx, cap);

  uint32_t* counters = ws<uint32_t>(ctx, kCounters);
  RayRec* queue[2] = {ws<RayRec>(ctx, kQueueA), ws<RayRec>(ctx, kQueueB)};
  uint32_t* keys[2] = {ws<uint32_t>(ctx, kKeysA), ws<uint32_t>(ctx, kKeysB)};
  float4* radiance = ws<float4>(ctx, kRadiance);
  float2* sxy = ws<float2>(ctx, kSampleXY);
  const int hit_bits_full = bits_for((uint64_t)s.num_tris);  // keys are prim + 1 <= num_tris
  const int hit_bits = ctx->tuning.hit_key_bits > 0 ? std::min(ctx->tuning.hit_key_bits, hit_bits_full) : hit_bits_full;
  const uint32_t hit_shift = (uint32_t)(hit_bits_full - hit_bits);
  // Traversal stack: an LDS ring per lane, spilling to HBM only past the ring.
  const int ring = choose_ring(ctx->tuning.trace_stack, s.max_depth);
  const uint32_t spill_stride = stack_need(s.max_depth);
  uint32_t* spill = nullptr;
  if ((uint32_t)ring < spill_stride) {
    need(kSpill, 4ull * cap * spill_stride);
    spill = ws<uint32_t>(ctx, kSpill);
  }

  StageClock clk;
  swr_stats acc{};
  hipEvent_t ev_start, ev_end;
  SWR_HIP(hipEventCreate(&ev_start));
  SWR_HIP(hipEventCreate(&ev_end));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{ev_start, ev_end};

  SWR_HIP(hipEventRecord(ev_start, st));
  uint32_t cur = 0, n_active = 0;
  uint64_t next_sid = 0;
  while (true) {
    if (ctx->cancel.load()) fail(SWR_E_CANCELLED, "render cancelled");
    const uint32_t k = (uint32_t)std::min<uint64_t>(cap - n_active, total - next_sid);
    const uint32_t n = n_active + k;
    if (n == 0) break;
    SWR_HIP(hipEventRecord(clk.ev[0], st));
    if (k)
      raygen_kernel<<<blocks(k), 256, 0, st>>>(p, (uint32_t)next_sid, k, n_active, queue[cur], keys[cur], sxy,
                                               radiance);
    SWR_HIP(hipEventRecord(clk.ev[1], st));
    uint32_t *rk = nullptr, *perm = nullptr;
    sort_pairs(ctx, keys[cur], ws<uint32_t>(ctx, kPerm), ws<uint32_t>(ctx, kSortAltK), ws<uint32_t>(ctx, kSortAltV), n,
               ray_bits, true, &rk, &perm, st);
    SWR_HIP(hipEventRecord(clk.ev[2], st));
    with_ring(ring, [&](auto kShort) {
      trace_queue_kernel<kShort.value><<<blocks(n, kTraceBlock), kTraceBlock, 0, st>>>(
          s.nodes, s.leaf_tris, perm, n, queue[cur], ws<uint32_t>(ctx, kHitKeys), hit_shift, spill, spill_stride);
    });
    SWR_HIP(hipEventRecord(clk.ev[3], st));
    // The ray permutation is consumed; the hit sort reuses its buffers.
    uint32_t *hk = nullptr, *hperm = nullptr;
    uint32_t* spare = perm == ws<uint32_t>(ctx, kPerm) ? ws<uint32_t>(ctx, kSortAltV) : ws<uint32_t>(ctx, kPerm);
    sort_pairs(ctx, ws<uint32_t>(ctx, kHitKeys), spare, keys[cur], perm, n, hit_bits, true, &hk, &hperm, st);
    SWR_HIP(hipEventRecord(clk.ev[4], st));
    zero_u32_kernel<<<1, 64, 0, st>>>(counters, 2);
    const int nxt = cur ^ 1;
    ShadeOut so;
    so.next = queue[nxt];
    so.next_keys = keys[nxt];
    so.shadow_rays = ws<float4>(ctx, kShadowRays);
    so.shadow_col = ws<float4>(ctx, kShadowCol);
    so.radiance = radiance;
    so.counters = reinterpret_cast<uint64_t*>(counters);
    shade_kernel<<<blocks(n, kShadeBlock), kShadeBlock, 0, st>>>(s, p, hperm, n, queue[cur], so);
    SWR_HIP(hipEventRecord(clk.ev[5], st));
    with_ring(ring, [&](auto kShort) {
      shadow_kernel<kShort.value><<<blocks(n, kTraceBlock), kTraceBlock, 0, st>>>(
          s.nodes, s.leaf_tris, n, so.shadow_rays, so.shadow_col, radiance, spill, spill_stride);
    });
    SWR_HIP(hipEventRecord(clk.ev[6], st));
    SWR_HIP(hipGetLastError());
    uint32_t host_counts[4];
    SWR_HIP(hipMemcpyAsync(host_counts, counters, 16, hipMemcpyDeviceToHost, st));
    SWR_HIP(hipStreamSynchronize(st));
    acc.ms_raygen += clk.ms(0, 1);
    acc.ms_sort_rays += clk.ms(1, 2);
    acc.ms_trace += clk.ms(2, 3);
    acc.ms_sort_hits += clk.ms(3, 4);
    acc.ms_shade += clk.ms(4, 5);
    acc.ms_shadow += clk.ms(5, 6);
    acc.camera_rays += k;
    acc.closest_rays += n;
    acc.shadow_rays += host_counts[1];
    acc.waves += 1;
    acc.max_wave = std::max(acc.max_wave, n);
    n_active = host_counts[0];
    next_sid += k;
    cur ^= 1;
  }
  SWR_HIP(hipEventRecord(clk.ev[1], st));
  dim3 fgrid(blocks(p.height, kFilterRows), blocks(p.width, kFilterCols));
  filter_kernel<<<fgrid, kFilterBlock, 0, st>>>(p, radiance, sxy, ws<float>(ctx, kPixels));
  SWR_HIP(hipGetLastError());
  SWR_HIP(hipEventRecord(clk.ev[2], st));
  SWR_HIP(hipEventRecord(ev_end, st));
  if (pixels_host)
    SWR_HIP(hipMemcpyAsync(pixels_host, ws<float>(ctx, kPixels), 12ull * p.width * p.height, hipMemcpyDeviceToHost,
                           st));
  if (samples_host) {
    DeviceBuffer staging;
    ensure(ctx, staging, 20ull * total);
    pack_samples_kernel<<<blocks(total), 256, 0, st>>>(p, total, sxy, radiance, static_cast<float*>(staging.ptr));
    SWR_HIP(hipMemcpyAsync(samples_host, staging.ptr, 20ull * total, hipMemcpyDeviceToHost, st));
    SWR_HIP(hipStreamSynchronize(st));
    release(ctx, staging);
  }
  SWR_HIP(hipStreamSynchronize(st));
  acc.ms_filter = clk.ms(1, 2);
  float total_ms = 0.f;
  SWR_HIP(hipEventElapsedTime(&total_ms, ev_start, ev_end));
  acc.ms_total = total_ms;
  if (stats) *stats = acc;
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
    auto time_it = [&](const swr::Tuning& t) {
      ctx->tuning = t;
      double best = 1e30;
      for (int rep = 0; rep < 2; ++rep) {
        swr_stats st{};
        swr::render_frame(ctx, scene, d, nullptr, nullptr, &st);
        best = std::min(best, st.ms_total);
      }
      return best;
    };
    swr::Tuning best = ctx->tuning;
    swr::render_frame(ctx, scene, d, nullptr, nullptr, nullptr);  // allocations outside the timings
    double best_ms = time_it(best);
    for (int kpt : {8, 12, 16, 20, 24}) {
      swr::Tuning t = best;
      t.sort_kpt = kpt;
      const double ms = time_it(t);
      if (ms < best_ms) best_ms = ms, best = t;
    }
    for (int rb : {32, 24, 20, 16}) {
      swr::Tuning t = best;
      t.ray_key_bits = rb;
      const double ms = time_it(t);
      if (ms < best_ms) best_ms = ms, best = t;
    }
    const int full = swr::bits_for((uint64_t)scene->dev.num_tris);
    for (int hb : {0, full - 6, full - 12}) {
      if (hb < 0 || (hb > 0 && hb < 4)) continue;
      swr::Tuning t = best;
      t.hit_key_bits = hb;
      const double ms = time_it(t);
      if (ms < best_ms) best_ms = ms, best = t;
    }
    for (int ring : {8, 16, 32, 64}) {
      swr::Tuning t = best;
      t.trace_stack = ring;
      const double ms = time_it(t);
      if (ms < best_ms) best_ms = ms, best = t;
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

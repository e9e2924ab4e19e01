// integrator.h -- device side of the path tracer shared by the wavefront kernels: frame
// constants of one render job, the BSDFs, quad lights, cameras and the queued ray record.
//
// Every function restates the reference's arithmetic (file:line at each one); FP contraction is
// off for the whole build (-ffp-contract=off) so that the CPU restatement and the GPU round the
// same way.
#pragma once

#include <cstdint>

#include "bands.h"
#include "device_util.h"
#include "scene.h"

namespace swr {

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

// Constants of one render job (one Pathtracer::process() iteration of one camera on one scene).
// The wavefront loop keeps a table of these on the device; a queued ray names its job.
struct FrameParams {
  uint32_t width, height, spp, base;
  int32_t sampler, filter, min_depth, max_depth;
  float threshold;
  uint32_t iteration;
  uint32_t seed_lo, seed_hi;
  uint32_t num_lights;
  float pick;              // 1 / number of lights
  float3 key_lo, key_scale;
  uint32_t ray_key_shift;  // the ray sort orders by the top (32 - shift) key bits
  uint32_t key_prefix;     // scene index in the top key_scene_bits bits of the ray key
  uint32_t key_scene_bits; // 0 when every ray of the run is on one scene
  uint32_t scene;          // index into the loop's scene table
  uint32_t job;            // job slot in the loop's job table
  uint32_t seq;            // admission number of the job in its loop
  V3 env;
  int32_t env_nonzero;
  uint64_t sample_base;    // first slot of this job in the per-sample arenas
  uint64_t pixel_base;     // first pixel of this job in the framebuffer arena
  // Column band: the job traces the samples of image columns [band_x0, band_x0 + band_cols)
  // (sample ids sid_base + [0, band_cols * height * spp), column-major as Camera.cpp:28 numbers
  // them) and its filter writes columns [out_x0, out_x1). A whole frame is the band 0..width.
  uint32_t sid_base, band_x0, band_cols, out_x0, out_x1;
  CameraRec cam;
};

// Sort key of a queued ray: the scene index on top, then the octant + Morton key of the ray on
// the job's scene bounds, truncated to the tuned number of bits.
__device__ inline uint32_t queue_key(const FrameParams& p, V3 o, V3 d) {
  const uint32_t k = ray_key(o.x, o.y, o.z, d.x, d.y, d.z, p.key_lo, p.key_scale);
  return (p.key_prefix | (k >> p.key_scene_bits)) >> p.ray_key_shift;
}

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

// ---- textures --------------------------------------------------------------------------------
// StandardTexture::initialize (StandardTexture.cpp:11-41) looks each hit's texture coordinate up
// with OpenImageIO's texture(): periodic wrap in s and t, zero derivatives. At zero filter width
// that samples the finest level with the bicubic filter: texel centres at (i + 0.5) / size and
// uniform cubic B-spline weights over the 4 x 4 texels around the point. t = 0 is the first row.
struct TexAxis {
  uint32_t i[4];  // texel indices, wrapped
  float w[4];     // B-spline weights
};
__device__ inline TexAxis tex_axis(float s, uint32_t n) {
  if (!isfinite(s)) s = 0.f;
  const float x = (s - floorf(s)) * (float)n - 0.5f;  // in [-0.5, n - 0.5]
  const float fx = floorf(x);
  const float f = x - fx;
  const int i0 = (int)fx - 1 + 2 * (int)n;            // >= 0 for n >= 1
  const float g = 1.f - f, f2 = f * f, f3 = f2 * f;
  TexAxis a;
  a.w[0] = g * g * g * (1.f / 6.f);
  a.w[1] = (3.f * f3 - 6.f * f2 + 4.f) * (1.f / 6.f);
  a.w[2] = (-3.f * f3 + 3.f * f2 + 3.f * f + 1.f) * (1.f / 6.f);
  a.w[3] = f3 * (1.f / 6.f);
#pragma unroll
  for (int k = 0; k < 4; ++k) a.i[k] = (uint32_t)(i0 + k) % n;
  return a;
}
__device__ inline V3 tex_lookup(const DeviceScene& sc, int32_t image, float s, float t) {
  const TexImage im = sc.tex_images[image];
  const TexAxis ax = tex_axis(s, im.width), ay = tex_axis(t, im.height);
  V3 acc = v3(0.f, 0.f, 0.f);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float4* row = sc.texels + im.offset + (uint64_t)ay.i[j] * im.width;
    V3 r = v3(0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 c = row[ax.i[i]];
      r = r + v3(c.x, c.y, c.z) * ax.w[i];
    }
    acc = acc + r * ay.w[j];
  }
  return acc;
}
// A shader's texture colour at (s, t): its post-order program run on a stack of colours.
// LayeredTexture::initialize (LayeredTexture.cpp:14-32): upper * mask + lower * (1 - mask).
__device__ inline V3 texture_colour(const DeviceScene& sc, const ShaderRec& sh, float s, float t) {
  V3 st[kTexStack];
  int sp = 0;
  for (int k = 0; k < sh.tex_len; ++k) {
    const TexOp op = sc.tex_ops[sh.tex_ofs + k];
    if (op.op == SWR_TEXTURE_LAYERED) {
      const V3 m = st[sp - 1], lo = st[sp - 2], up = st[sp - 3];
      sp -= 3;
      st[sp++] = v3(up.x * m.x + lo.x * (1.f - m.x), up.y * m.y + lo.y * (1.f - m.y), up.z * m.z + lo.z * (1.f - m.z));
    } else if (op.op == SWR_TEXTURE_IMAGE) {
      st[sp++] = tex_lookup(sc, op.image, s, t);
    } else {
      st[sp++] = v3(op.colour[0], op.colour[1], op.colour[2]);
    }
  }
  return st[0];
}

// ---- queue records ----------------------------------------------------------------------------
// One queued ray is one 64-byte line: origin + tmin, direction + tmax, throughput + lastPdf, and
// id = (sample id within its job, job << 16 | depth, scene index until traced then hit t, hit
// leaf). Traversal reads the scene and fills the last two words, so shading gathers each ray with
// a single line read.
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
// Slot of a sample in its job's arena range: the band's columns, from its first one.
__device__ inline uint32_t slot_of_sid(const FrameParams& p, uint32_t sid) {
  const uint32_t x = sid / (p.height * p.spp);
  const uint32_t rem = sid - x * p.height * p.spp;
  const uint32_t y = rem / p.spp;
  return sample_slot(x - p.band_x0, y, rem - y * p.spp, p.height, p.spp);
}

// Host side (render.hip): frame constants of one job on one scene.
CameraRec make_camera(const swr_camera_desc& d);
uint64_t spp_of(const swr_render_desc& d);
FrameParams make_params(const swr_render_desc& d, const DeviceScene* s);
// Restricts a job to the column band whose filter writes columns [x0, x1): the band's samples
// are those columns plus, for the tent filter, one neighbour column on each side inside the image.
void set_band(FrameParams* p, uint32_t x0, uint32_t x1);
// Samples of a job (its band's columns x height x spp).
inline uint64_t band_samples(const FrameParams& p) { return (uint64_t)p.band_cols * p.height * p.spp; }
int bits_for(uint64_t max_value);

}  // namespace swr

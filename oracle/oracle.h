/*
 * oracle.h -- CPU restatement of joshbainbridge/msc-project's sorted-wavefront path tracer.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness oracle and the CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the reported CPU baseline. The product library (libswr.so) never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates. Random numbers
 * come from Philox-4x32-10 (Salmon et al., SC'11) keyed per (seed, sample, iteration, depth,
 * slot) instead of the reference's per-thread mt19937 (src/core/RandomGenerator.cpp:6-9),
 * so a render is a pure function of its inputs and the GPU path can be compared stream for
 * stream.
 */
#ifndef SWR_ORACLE_H
#define SWR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene description records. Field meanings follow the reference's YAML decoders
 * (the convert<> specialisations in include/core), laid out identically to the product's swr.h records. */
typedef struct {
  int32_t type;          /* 0 PinHole (PinHoleCamera.cpp), 1 ThinLens (ThinLensCamera.cpp) */
  float translation[3];
  float rotation[3];     /* degrees; camera order T*Ry*Rx*Rz (PinHoleCamera.cpp:7-11) */
  float focal_length;    /* mm */
  float focal_distance;  /* ThinLens only */
  float aperture;        /* f-number, ThinLens only */
} oracle_camera_desc;

typedef struct {
  int32_t width, height;
  int32_t base;          /* sample base (Image.h:46) */
  int32_t spp;           /* 0 -> base*base (Image.h:65); otherwise explicit (extension) */
  int32_t sampler;       /* 0 Stratified, 1 Independent, 2 Grid */
  int32_t filter;        /* 0 Tent, 1 Box */
  int32_t min_depth, max_depth;
  float threshold;
  int32_t shading_size;  /* reference knob, unused by the path-at-a-time oracle */
  int32_t bin_exponent;  /* reference knob, unused by the path-at-a-time oracle */
  uint32_t iteration;    /* progressive iteration index, an RNG counter word */
  uint64_t seed;         /* Philox key */
  oracle_camera_desc camera;
} oracle_render_desc;

typedef struct {
  int32_t type;          /* 0 Lambert (LambertShader.cpp), 1 Null (NullShader.cpp), 2 Glossy (extension) */
  float colour[3];       /* Constant texture colour (ConstantTexture.cpp:22-25) */
  float reflectance;     /* LambertShader::continuation (LambertShader.cpp:22-25) */
  float roughness;       /* Glossy only: GGX alpha */
} oracle_shader_desc;

typedef struct {
  float translation[3];
  float rotation[3];     /* degrees; light order T*Rx*Ry*Rz*S (QuadLight.cpp:24-29) */
  float scale[2];
  float intensity;
} oracle_quad_light_desc;

typedef struct {
  uint64_t camera_rays;
  uint64_t closest_rays;   /* all closest-hit queries (camera + continuation) */
  uint64_t shadow_rays;
  double seconds;
} oracle_stats;

/* ---- primitives ---------------------------------------------------------------------- */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* u32 -> float in [0,1): top 24 bits times 2^-24. */
float oracle_u01(uint32_t x);
/* Octant(3 bits) << 29 | morton29(quantised origin) for n rays of 8 floats
 * (ox,oy,oz,tmin,dx,dy,dz,tmax). scale[i] = 2^bits_i / (hi_i - lo_i), computed by the caller. */
void oracle_ray_keys(const float* rays, size_t n, const float lo[3], const float scale[3],
                     uint32_t* keys);
uint32_t oracle_morton29(uint32_t qx, uint32_t qy, uint32_t qz);
/* Hit shading key: global triangle id + 1, 0 for a miss (RayUncompressed.h:63-69 order). */
void oracle_hit_keys(const uint32_t* prim, size_t n, uint32_t* keys);
/* Stable sort: perm[i] = original index of the i-th smallest key (std::stable_sort). */
void oracle_stable_sort_perm(const uint32_t* keys, size_t n, uint32_t* perm);

/* Texture nodes (the reference's TextureInterface implementations, decoded by the
 * convert<> specialisations of ConstantTexture.h, StandardTexture.h and LayeredTexture.h):
 * 0 Constant (colour), 1 Standard (an image), 2 Layered (indices of earlier nodes). */
typedef struct {
  int32_t type;
  float colour[3];
  int32_t image;
  int32_t upper, lower, mask;
} oracle_texture_desc;

/* ---- scene ----------------------------------------------------------------------------- */
typedef struct oracle_scene oracle_scene;
oracle_scene* oracle_scene_create(void);
void oracle_scene_destroy(oracle_scene* s);
int oracle_scene_add_shader(oracle_scene* s, const oracle_shader_desc* d);
/* Positions are float4 (x,y,z,w) per vertex as handed to rtcSetBuffer (Pathtracer.cpp:185-192),
 * already transformed. uv may be NULL. Returns the object index. */
int oracle_scene_add_mesh(oracle_scene* s, const float* pos4, uint32_t nverts,
                          const uint32_t* idx3, uint32_t ntris, const float* uv, int32_t shader);
int oracle_scene_add_quad_light(oracle_scene* s, const oracle_quad_light_desc* d);
/* Texture images (width x height RGB floats, row 0 at t = 0), texture nodes and the texture of
 * a shader (LambertShader's texture, LambertShader.h convert<>). Return an index or -1. */
int oracle_scene_add_image(oracle_scene* s, int32_t width, int32_t height, const float* rgb);
int oracle_scene_add_texture(oracle_scene* s, const oracle_texture_desc* d);
int oracle_scene_set_shader_texture(oracle_scene* s, int32_t shader, int32_t texture);
void oracle_scene_set_environment(oracle_scene* s, const float rgb[3]);
/* Lays out geomIDs as the reference does (objects, then one mesh per light; Pathtracer.cpp:166-272)
 * and builds a binned-SAH BVH. */
int oracle_scene_commit(oracle_scene* s);
uint32_t oracle_scene_num_triangles(const oracle_scene* s);
/* Triangle vertices in global triangle order (9 floats each), for test inspection. */
int oracle_scene_triangles(const oracle_scene* s, float* out9);

/* Closest hit (RayIntersect.cpp:7-12 -> rtcIntersect). rays: n x 8 floats. hits: n x 4 words
 * (t, u, v as float bits; prim = global triangle id or 0xFFFFFFFF). */
int oracle_intersect(const oracle_scene* s, const float* rays, size_t n, uint32_t mask, uint32_t* hits);
/* Same query by brute force over all triangles in double precision (independent pin). */
int oracle_intersect_brute(const oracle_scene* s, const float* rays, size_t n, uint32_t mask,
                           uint32_t* hits);
/* Any hit in (tmin, tmax) (Integrator.cpp:153-174 -> rtcOccluded): flags[i] = 1 if occluded. */
int oracle_occluded(const oracle_scene* s, const float* rays, size_t n, uint32_t mask, uint32_t* flags);

/* StandardTexture's image lookup (StandardTexture.cpp:11-41, OpenImageIO texture() with periodic
 * wrap and zero derivatives): cubic B-spline over the 4 x 4 texels around (s, t). */
void oracle_texture_lookup(const float* rgb, int32_t width, int32_t height, float s, float t, float out[3]);
/* Colour of a committed scene's shader at barycentrics (u, v) of global triangle prim: the texture
 * coordinate of PolygonObject::texture (PolygonObject.cpp:45-63), then the shader's texture
 * (its constant colour when it has none). */
int oracle_texture_colour(const oracle_scene* s, int32_t shader, uint32_t prim, float u, float v, float out[3]);

/* ---- render ---------------------------------------------------------------------------- */
/* Camera rays for sample ids [sid0, sid0+n) (Camera.cpp:6-59): rays n x 8, film xy n x 2. */
int oracle_raygen(const oracle_render_desc* d, uint32_t sid0, size_t n, float* rays, float* sample_xy);
/* One Pathtracer::process() iteration (Pathtracer.cpp:425-493) computed path by path.
 * pixels: W*H*3 (row y*W+x, as Pixel in Image.h:16-20), the filtered iteration result.
 * samples (optional): W*H*spp*5 (x, y, r, g, b) in reference sampleID order.
 * Sample ids outside [sid_begin, sid_end) are skipped (sid_end 0 = all); threads 0 = all. */
int oracle_render(const oracle_scene* s, const oracle_render_desc* d, float* pixels, float* samples,
                  uint32_t sid_begin, uint32_t sid_end, int threads, oracle_stats* stats);
/* Filter stage alone (TentFilter.cpp:5-63, BoxFilter.cpp:5-37): samples W*H*spp*5 -> pixels W*H*3
 * (written, not accumulated). */
void oracle_filter(int32_t filter, int32_t width, int32_t height, int32_t spp, const float* samples,
                   float* pixels);

/* ---- BSDFs alone (the shading kernel's lobes: LambertShader.cpp:27-73, NullShader.cpp:25-54,
 * and the GGX reflection lobe of the Glossy shader), vectorised over n direction pairs. ------ */
/* GGX normal distribution D(cos theta_h) and Smith masking G1(cos theta_v) for roughness alpha. */
void oracle_ggx_d_n(float alpha, size_t n, const float* cos_h, float* d);
void oracle_ggx_g1_n(float alpha, size_t n, const float* cos_v, float* g);
/* ShaderInterface::evaluate: in3/out3/n3 are n x 3 unit vectors (light direction, view direction,
 * shading normal); f3 n x 3, cos and pdf n each. */
void oracle_bsdf_eval_n(const oracle_shader_desc* sh, size_t n, const float* in3, const float* out3,
                        const float* n3, float* f3, float* cos, float* pdf);
/* ShaderInterface::sample from the uniforms u2 (n x 2): the sampled direction in in3 with its
 * evaluation; ok[i] = 0 when the sample carries no energy (the path ends there). */
void oracle_bsdf_sample_n(const oracle_shader_desc* sh, size_t n, const float* u2, const float* out3,
                          const float* n3, float* in3, float* f3, float* cos, float* pdf, uint8_t* ok);

#ifdef __cplusplus
}
#endif
#endif

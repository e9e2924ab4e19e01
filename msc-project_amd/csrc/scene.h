// scene.h -- device scene layout shared by the BVH build, traversal and shading kernels.
//
// The reference hands Embree one triangle mesh per object and light, float4 vertices and
// u32x3 indices (src/core/Pathtracer.cpp:178-201, :242-265), and keeps per-object shader and
// light records on the host (Scene.h:17-25). Here the whole scene lives in HBM:
//   geometry (global triangle order, geomID-major): pos float4 per vertex, tri uint4
//     (i0, i1, i2, geomID) per triangle, optional uv float2 per vertex;
//   records: GeomRec per geomID, ShaderRec per shader, LightRec per quad light;
//   BVH4 nodes (128 B each) and the leaf triangles in BVH order (48 B each: v0 with the global
//     triangle id in w, e1 = v0 - v1 with the geometry mask in w, e2 = v2 - v0 with the geomID
//     in w), so that shading gets Ng = e1 x e2 and the geometry from one 48-byte read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "swr_internal.h"

namespace swr {

// BVH4 node. Child c is internal when child[c] < 0x80000000 (node index), a leaf when bit 31 is
// set: bits 30..27 hold count-1 (1..16 triangles) and bits 26..0 the first leaf triangle.
// Unused slots hold kNoChild and a box far outside any ray's range.
struct alignas(16) Node4 {
  float4 lox, hix, loy, hiy, loz, hiz;
  uint4 child;
  uint4 meta;  // x: number of children, y: depth, z/w: unused
};
static_assert(sizeof(Node4) == 128, "BVH4 node must be 128 bytes");

// The same node with quantised child boxes (64 B, half a cache line): per axis the node box
// origin p and a power-of-two quantum s = 2^(e - 127) (stored as the float s), and per child its
// lo and hi planes as bytes q with plane = p + q * s, rounded outward by at least one quantum (so
// that the rounding of the traversal's slab arithmetic never culls a box). Rows: (px, py, pz, sz),
// (qlo_x, qhi_x, qlo_y, qhi_y), (qlo_z, qhi_z, sx, sy), child; byte c of a q word is child c. An
// unused slot has planes qlo 255, qhi 0 (an empty interval) and the null leaf as its child.
struct alignas(16) NodeQ {
  float px, py, pz, sz;
  uint32_t qlox, qhix, qloy, qhiy;
  uint32_t qloz, qhiz;
  float sx, sy;
  uint4 child;
};
static_assert(sizeof(NodeQ) == 64, "quantised node must be 64 bytes");

constexpr uint32_t kNoChild = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
// Material word of a leaf triangle (w of its third row): this bit and the light index for a light,
// else the shader index, so that shading needs no geometry-table read.
constexpr uint32_t kLeafLight = 0x80000000u;
constexpr float kEmptyBox = 1e30f;
__host__ __device__ inline uint32_t leaf_code(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1u) << 27) | first;
}
__host__ __device__ inline uint32_t leaf_first(uint32_t c) { return c & 0x07FFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(uint32_t c) { return ((c >> 27) & 15u) + 1u; }

struct GeomRec {
  int32_t shader;   // index into the shader table
  int32_t light;    // light index or -1 (Scene::shaders_to_lights, Pathtracer.cpp:228-230)
  uint32_t mask;    // 0xFFFFFFFF objects, 0xF0000000 lights (Pathtracer.cpp:267)
  uint32_t first_tri;
};

struct ShaderRec {
  int32_t type;     // SWR_SHADER_*
  float colour[3];
  float reflectance;
  float roughness;
  int32_t tex_ofs;  // texture program: tex_len ops from tex_ops[tex_ofs] (tex_len 0: constant colour)
  int32_t tex_len;
};

// A shader's texture graph compiled to a post-order stack program: CONSTANT and IMAGE push a
// colour, LAYERED pops mask, lower, upper and pushes upper * mask + lower * (1 - mask).
constexpr int kTexStack = 8;        // deepest stack a program may use
constexpr size_t kTexOpsMax = 4096;  // longest program (shared subgraphs are expanded)
struct TexOp {
  int32_t op;       // SWR_TEXTURE_*
  float colour[3];  // CONSTANT
  int32_t image;    // IMAGE
  int32_t pad[3];
};
struct TexImage {
  uint32_t width, height;
  uint64_t offset;  // first texel in the texel array (float4 per texel, rgb + unused)
};
static_assert(sizeof(ShaderRec) == 32 && sizeof(TexOp) == 32 && sizeof(TexImage) == 16,
              "scene records are read as 16-byte rows");

// QuadLight after construct() (QuadLight.cpp:5-39): transform T*R*S(1,sx,sy) and normal.
struct LightRec {
  float rot[9];     // row-major Rx*Ry*Rz
  float t[3];
  float sx, sy;
  float n[3];
  float intensity;
  float pad[2];
};

struct DeviceScene {
  const float4* pos = nullptr;
  const uint4* tri = nullptr;
  const float2* uv = nullptr;
  const GeomRec* geoms = nullptr;
  const ShaderRec* shaders = nullptr;
  const LightRec* lights = nullptr;
  const Node4* nodes = nullptr;
  const NodeQ* qnodes = nullptr;      // the same nodes quantised
  const float4* leaf_tris = nullptr;  // 3 float4 per triangle in BVH order
  const TexOp* tex_ops = nullptr;
  const TexImage* tex_images = nullptr;
  const float4* texels = nullptr;
  uint32_t num_tris = 0, num_geoms = 0, num_lights = 0, num_nodes = 0;
  uint32_t max_depth = 0;             // BVH4 depth (root = 0)
  uint32_t hit_base = 0;              // in a wavefront loop's scene table: the scene's hit keys are
                                      // hit_base + triangle + 1 (hit_base for a miss), so keys order
                                      // hits by (scene, triangle) in the fewest bits
  float env[3] = {0.f, 0.f, 0.f};
  float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};  // scene bounds
};

// ---- bvh.hip ----
// Builds the BVH4 of the scene's triangles on the device: 30-bit centroid Morton codes, radix
// sort, Karras hierarchy, bottom-up boxes, BVH4 collapse and the BVH-order triangle array.
struct BvhBuild {
  DeviceBuffer nodes, qnodes, leaf_tris;
  uint32_t num_nodes = 0, max_depth = 0;
  uint32_t num_tris = 0;  // leaf_tris holds num_tris triangles and the null leaf after them
};
// Code of the null leaf: one degenerate triangle (all zeros, geometry mask 0) after the scene's
// num_tris leaf triangles, which no ray hits; unused slots of quantised nodes point at it.
__host__ __device__ inline uint32_t null_leaf_code(uint32_t num_tris) { return leaf_code(num_tris, 1u); }
void build_bvh(swr_context* ctx, const float4* pos, const uint4* tri, const GeomRec* geoms, uint32_t num_tris,
               const float lo[3], const float hi[3], int leaf_max, BvhBuild* out);
// Writes out->qnodes, the quantised copy of out->nodes.
void quantize_bvh(swr_context* ctx, BvhBuild* out);
// Quantum exponent (biased) and outward-rounded plane bytes of one axis of a node whose valid
// children span [lo[c], hi[c]] (valid[c]); shared by the build and its validation.
struct QuantAxis {
  float p;
  uint32_t e, qlo, qhi;
};
__host__ __device__ inline QuantAxis quantize_axis(const float* lo, const float* hi, const bool* valid) {
  float l = 0.f, h = 0.f;
  bool any = false;
  for (int c = 0; c < 4; ++c)
    if (valid[c]) {
      l = any ? fminf(l, lo[c]) : lo[c];
      h = any ? fmaxf(h, hi[c]) : hi[c];
      any = true;
    }
  // s = 2^k: at least 2^-126, (h - l) / s <= 252, and s >= ulp(p) so that p = l - s loses
  // less than one quantum when rounded down to a float.
  const double E = (double)h - (double)l;
  int k = -126;
  if (E > 0.0 && E < 1e300) k = (int)ceil(log2(E / 252.0));
  int em = 0;
  frexp((double)fmaxf(fabsf(l), fabsf(h)), &em);  // |l|, |h| < 2^em
  k = k > em - 23 ? k : em - 23;
  k = k > -126 ? k : -126;
  k = k < 127 ? k : 127;
  const double sd = ldexp(1.0, k);
  double pd = (double)l - sd;
  float p = (float)pd;
  if ((double)p > pd) p = nextafterf(p, -INFINITY);  // round down
  QuantAxis a{p, (uint32_t)(k + 127), 0u, 0u};
  for (int c = 0; c < 4; ++c) {
    uint32_t ql = 255u, qh = 0u;  // unused slots: an empty interval (masked by the traversal anyway)
    if (valid[c]) {
      const double lq = floor(((double)lo[c] - (double)p) / sd) - 1.0;
      const double hq = ceil(((double)hi[c] - (double)p) / sd) + 1.0;
      ql = (uint32_t)(!(lq >= 0.0) ? 0.0 : (lq > 255.0 ? 255.0 : lq));
      qh = (uint32_t)(!(hq >= 0.0) ? 0.0 : (hq > 255.0 ? 255.0 : hq));
    }
    a.qlo |= ql << (8 * c);
    a.qhi |= qh << (8 * c);
  }
  return a;
}

// ---- trace.hip ----
// Closest hit (hits: t, u, v bits and global triangle id) and any-hit occlusion queries.
// keys (optional) receives the hit shading key (prim + 1, misses 0) of every ray.
void launch_intersect(swr_context* ctx, const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask,
                      uint4* hits, uint32_t* keys, hipStream_t stream);
void launch_occluded(swr_context* ctx, const DeviceScene& s, const float4* rays, uint32_t n, uint32_t mask,
                     uint32_t* flags, hipStream_t stream);
// Stack entries one traversal needs for a BVH4 of the given depth: every level leaves at most
// three siblings behind, and the deepest node may push four before popping one.
inline uint32_t stack_need(uint32_t max_depth) { return 3u * max_depth + 4u; }
uint32_t max_supported_depth();

}  // namespace swr

struct swr_scene {
  swr_context* ctx = nullptr;
  struct Mesh {
    std::vector<float> pos4;
    std::vector<uint32_t> idx;
    std::vector<float> uv;
    int32_t shader = 0;
  };
  std::vector<Mesh> objects;
  std::vector<swr_shader_desc> shaders;
  std::vector<int32_t> shader_texture;  // per shader: texture node or -1
  std::vector<swr_texture_desc> textures;
  struct Image {
    int32_t width = 0, height = 0;
    std::vector<float> rgb;
  };
  std::vector<Image> images;
  bool textured = false;  // committed: some shader has a texture
  std::vector<swr_quad_light_desc> light_descs;
  float env[3] = {0.f, 0.f, 0.f};
  bool committed = false;
  int leaf_max = 2;
  int tuned_ring = 0;  // traversal stack ring chosen for this scene by the batch autotuner (0 = none)

  // committed
  std::vector<swr::GeomRec> geoms;
  std::vector<swr::ShaderRec> all_shaders;
  std::vector<swr::LightRec> lights;
  swr::DeviceBuffer d_pos, d_tri, d_uv, d_geoms, d_shaders, d_lights, d_tex_ops, d_tex_images, d_texels;
  swr::BvhBuild bvh;
  swr::DeviceScene dev;
};

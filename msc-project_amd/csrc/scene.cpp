// scene.cpp -- scene assembly, upload and the Embree-replacement query entry points.
//
// Mirrors the scene half of Pathtracer::construct (src/core/Pathtracer.cpp:163-274): objects
// and shaders in document order, then one placeholder object + Null shader + quad mesh with
// mask 0xF0000000 per light, then rtcCommit. geomIDs are therefore objects first, lights after.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>

#include "scene.h"

namespace swr {
namespace {

const float kPi180 = 0.01745329251994329577f;  // Constants.h:12

// 3x3 rotation about one axis by deg degrees, row-major (Eigen::AngleAxisf).
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

void mat_vec(const float m[9], const float v[3], float out[3]) {
  const float x = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  const float y = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  const float z = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
  out[0] = x;
  out[1] = y;
  out[2] = z;
}

// QuadLight::construct (QuadLight.cpp:24-31): T * Rx * Ry * Rz * S(1, sx, sy); the normal is the
// inverse-transpose image of x-hat, which for this transform is R * x-hat.
LightRec make_light(const swr_quad_light_desc& d) {
  LightRec L{};
  float rx[9], ry[9], rz[9], t[9];
  axis_rotation(0, d.rotation[0], rx);
  axis_rotation(1, d.rotation[1], ry);
  axis_rotation(2, d.rotation[2], rz);
  mat_mul(rx, ry, t);
  mat_mul(t, rz, L.rot);
  std::memcpy(L.t, d.translation, sizeof L.t);
  L.sx = d.scale[0];
  L.sy = d.scale[1];
  const float xhat[3] = {1.f, 0.f, 0.f};
  float n[3];
  mat_vec(L.rot, xhat, n);
  const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int k = 0; k < 3; ++k) L.n[k] = len > 0.f ? n[k] * (1.f / len) : n[k];
  L.intensity = d.intensity;
  return L;
}

void light_point(const LightRec& L, float px, float py, float pz, float out[3]) {
  const float p[3] = {px, L.sx * py, L.sy * pz};
  mat_vec(L.rot, p, out);
  for (int k = 0; k < 3; ++k) out[k] += L.t[k];
}

template <class T>
void upload(swr_context* ctx, DeviceBuffer& b, const std::vector<T>& v) {
  ensure(ctx, b, std::max<size_t>(v.size() * sizeof(T), 16));
  if (!v.empty()) SWR_HIP(hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

}  // namespace

void commit_scene(swr_scene* s) {
  swr_context* ctx = s->ctx;
  SWR_HIP(hipSetDevice(ctx->device));
  for (const auto& m : s->objects)
    if (m.shader < 0 || m.shader >= (int32_t)s->shaders.size())
      fail(SWR_E_INVALID, "scene: object references shader " + std::to_string(m.shader) + " of " +
                              std::to_string(s->shaders.size()));
  s->geoms.clear();
  s->all_shaders.clear();
  s->lights.clear();
  // Texture programs: each textured shader's graph in post order (children before parents).
  std::vector<TexOp> ops;
  s->textured = false;
  for (size_t i = 0; i < s->shaders.size(); ++i) {
    const swr_shader_desc& d = s->shaders[i];
    ShaderRec r{};
    r.type = d.type;
    std::memcpy(r.colour, d.colour, sizeof r.colour);
    r.reflectance = d.reflectance;
    r.roughness = d.roughness;
    const int32_t root = i < s->shader_texture.size() ? s->shader_texture[i] : -1;
    if (root >= 0) {
      r.tex_ofs = (int32_t)ops.size();
      int depth = 0, max_depth = 0;
      // Post order by an explicit stack of (node, children visited).
      std::vector<std::pair<int32_t, bool>> todo{{root, false}};
      while (!todo.empty()) {
        auto [node, expanded] = todo.back();
        todo.pop_back();
        const swr_texture_desc& t = s->textures[node];
        if (t.type == SWR_TEXTURE_LAYERED && !expanded) {
          todo.push_back({node, true});
          todo.push_back({t.mask, false});
          todo.push_back({t.lower, false});
          todo.push_back({t.upper, false});
          continue;
        }
        TexOp op{};
        op.op = t.type;
        std::memcpy(op.colour, t.colour, sizeof op.colour);
        op.image = t.image;
        ops.push_back(op);
        if (ops.size() - (size_t)r.tex_ofs > kTexOpsMax)
          fail(SWR_E_INVALID, "scene: texture of shader " + std::to_string(i) + " expands to more than " +
                                  std::to_string(kTexOpsMax) + " nodes");
        depth += t.type == SWR_TEXTURE_LAYERED ? -2 : 1;
        max_depth = std::max(max_depth, depth);
      }
      if (max_depth > kTexStack)
        fail(SWR_E_INVALID, "scene: texture of shader " + std::to_string(i) + " nests layers too deeply");
      r.tex_len = (int32_t)ops.size() - r.tex_ofs;
      s->textured = true;
    }
    s->all_shaders.push_back(r);
  }
  std::vector<TexImage> images;
  std::vector<float4> texels;
  for (const auto& im : s->images) {
    images.push_back(TexImage{(uint32_t)im.width, (uint32_t)im.height, (uint64_t)texels.size()});
    for (size_t k = 0; k < (size_t)im.width * im.height; ++k)
      texels.push_back(make_float4(im.rgb[3 * k], im.rgb[3 * k + 1], im.rgb[3 * k + 2], 0.f));
  }
  std::vector<float4> pos;
  std::vector<uint4> tri;
  std::vector<float2> uv;
  bool any_uv = false;
  for (const auto& m : s->objects) any_uv |= !m.uv.empty();
  // Pass 1: objects (Pathtracer.cpp:166-217).
  for (size_t g = 0; g < s->objects.size(); ++g) {
    const auto& m = s->objects[g];
    GeomRec gr{m.shader, -1, SWR_MASK_ALL, (uint32_t)tri.size()};
    s->geoms.push_back(gr);
    const uint32_t base = (uint32_t)pos.size();
    const size_t nv = m.pos4.size() / 4;
    for (size_t v = 0; v < nv; ++v) {
      pos.push_back(make_float4(m.pos4[4 * v], m.pos4[4 * v + 1], m.pos4[4 * v + 2], 0.f));
      if (any_uv) uv.push_back(m.uv.empty() ? make_float2(0.f, 0.f) : make_float2(m.uv[2 * v], m.uv[2 * v + 1]));
    }
    for (size_t t = 0; t < m.idx.size() / 3; ++t)
      tri.push_back(make_uint4(base + m.idx[3 * t], base + m.idx[3 * t + 1], base + m.idx[3 * t + 2], (uint32_t)g));
  }
  // Pass 2: lights (Pathtracer.cpp:219-272; the quad of QuadLight.cpp:7-19).
  const float quad[6][3] = {{0.f, -0.5f, -0.5f}, {0.f, 0.5f, -0.5f}, {0.f, 0.5f, 0.5f},
                            {0.f, 0.5f, 0.5f},   {0.f, -0.5f, 0.5f}, {0.f, -0.5f, -0.5f}};
  for (size_t l = 0; l < s->light_descs.size(); ++l) {
    const LightRec L = make_light(s->light_descs[l]);
    const int32_t shader_id = (int32_t)s->all_shaders.size();
    ShaderRec null_shader{};
    null_shader.type = SWR_SHADER_NULL;
    s->all_shaders.push_back(null_shader);
    const uint32_t g = (uint32_t)s->geoms.size();
    s->geoms.push_back(GeomRec{shader_id, (int32_t)l, SWR_MASK_LIGHT, (uint32_t)tri.size()});
    const uint32_t base = (uint32_t)pos.size();
    for (int k = 0; k < 6; ++k) {
      float p[3];
      light_point(L, quad[k][0], quad[k][1], quad[k][2], p);
      pos.push_back(make_float4(p[0], p[1], p[2], 0.f));
      if (any_uv) uv.push_back(make_float2(0.f, 0.f));
    }
    tri.push_back(make_uint4(base + 0, base + 1, base + 2, g));
    tri.push_back(make_uint4(base + 3, base + 4, base + 5, g));
    s->lights.push_back(L);
  }
  if (tri.size() >= (size_t(1) << 27)) fail(SWR_E_INVALID, "scene: at most 2^27 triangles");

  // Scene bounds and centroid bounds.
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (const uint4& t : tri) {
    const float4 a = pos[t.x], b = pos[t.y], c = pos[t.z];
    const float cx = (a.x + b.x + c.x) * (1.f / 3.f), cy = (a.y + b.y + c.y) * (1.f / 3.f),
                cz = (a.z + b.z + c.z) * (1.f / 3.f);
    const float vx[3] = {a.x, b.x, c.x}, vy[3] = {a.y, b.y, c.y}, vz[3] = {a.z, b.z, c.z};
    for (int k = 0; k < 3; ++k) {
      lo[0] = std::min(lo[0], vx[k]); hi[0] = std::max(hi[0], vx[k]);
      lo[1] = std::min(lo[1], vy[k]); hi[1] = std::max(hi[1], vy[k]);
      lo[2] = std::min(lo[2], vz[k]); hi[2] = std::max(hi[2], vz[k]);
    }
    clo[0] = std::min(clo[0], cx); chi[0] = std::max(chi[0], cx);
    clo[1] = std::min(clo[1], cy); chi[1] = std::max(chi[1], cy);
    clo[2] = std::min(clo[2], cz); chi[2] = std::max(chi[2], cz);
  }
  if (tri.empty())
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = clo[k] = chi[k] = 0.f;

  upload(ctx, s->d_pos, pos);
  upload(ctx, s->d_tri, tri);
  if (any_uv) upload(ctx, s->d_uv, uv);
  upload(ctx, s->d_geoms, s->geoms);
  upload(ctx, s->d_shaders, s->all_shaders);
  upload(ctx, s->d_lights, s->lights);
  if (!ops.empty()) upload(ctx, s->d_tex_ops, ops);
  if (!images.empty()) upload(ctx, s->d_tex_images, images);
  if (!texels.empty()) upload(ctx, s->d_texels, texels);
  build_bvh(ctx, static_cast<const float4*>(s->d_pos.ptr), static_cast<const uint4*>(s->d_tri.ptr),
            static_cast<const GeomRec*>(s->d_geoms.ptr), (uint32_t)tri.size(), clo, chi, s->leaf_max, &s->bvh);
  quantize_bvh(ctx, &s->bvh);
  SWR_HIP(hipStreamSynchronize(ctx->stream));
  if (stack_need(s->bvh.max_depth) > 3 * max_supported_depth() + 4)
    fail(SWR_E_INVALID, "scene: BVH depth " + std::to_string(s->bvh.max_depth) + " exceeds the traversal stack");
  // Traversal addresses nodes by 32-bit byte offsets (128-byte nodes).
  if (s->bvh.num_nodes >= (1u << 25))
    fail(SWR_E_INVALID, "scene: " + std::to_string(s->bvh.num_nodes) + " BVH nodes; at most 2^25 - 1");

  DeviceScene& d = s->dev;
  d.pos = static_cast<const float4*>(s->d_pos.ptr);
  d.tri = static_cast<const uint4*>(s->d_tri.ptr);
  d.uv = any_uv ? static_cast<const float2*>(s->d_uv.ptr) : nullptr;
  d.geoms = static_cast<const GeomRec*>(s->d_geoms.ptr);
  d.shaders = static_cast<const ShaderRec*>(s->d_shaders.ptr);
  d.lights = static_cast<const LightRec*>(s->d_lights.ptr);
  d.nodes = static_cast<const Node4*>(s->bvh.nodes.ptr);
  d.qnodes = static_cast<const NodeQ*>(s->bvh.qnodes.ptr);
  d.leaf_tris = static_cast<const float4*>(s->bvh.leaf_tris.ptr);
  d.tex_ops = ops.empty() ? nullptr : static_cast<const TexOp*>(s->d_tex_ops.ptr);
  d.tex_images = images.empty() ? nullptr : static_cast<const TexImage*>(s->d_tex_images.ptr);
  d.texels = texels.empty() ? nullptr : static_cast<const float4*>(s->d_texels.ptr);
  d.num_tris = (uint32_t)tri.size();
  d.num_geoms = (uint32_t)s->geoms.size();
  d.num_lights = (uint32_t)s->lights.size();
  d.num_nodes = s->bvh.num_nodes;
  d.max_depth = s->bvh.max_depth;
  std::memcpy(d.env, s->env, sizeof d.env);
  std::memcpy(d.lo, lo, sizeof lo);
  std::memcpy(d.hi, hi, sizeof hi);
  s->committed = true;
}

}  // namespace swr

void swr::detach_scene(swr_scene* s) {
  for (DeviceBuffer* b : {&s->d_pos, &s->d_tri, &s->d_uv, &s->d_geoms, &s->d_shaders, &s->d_lights, &s->bvh.nodes,
                          &s->bvh.qnodes, &s->bvh.leaf_tris, &s->d_tex_ops, &s->d_tex_images, &s->d_texels})
    *b = DeviceBuffer{};
  s->dev = DeviceScene{};
  s->committed = false;
  s->ctx = nullptr;
}

using swr::Error;

template <class F>
static int scene_guard(swr_context* ctx, F&& f) {
  try {
    f();
    return SWR_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return SWR_E_INVALID;
  }
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

extern "C" {

int swr_scene_create(swr_context* ctx, swr_scene** out) {
  if (!ctx || !out) return SWR_E_INVALID;
  *out = new swr_scene();
  (*out)->ctx = ctx;
  ctx->scenes.insert(*out);
  return SWR_OK;
}

void swr_scene_destroy(swr_scene* s) {
  if (!s) return;
  if (swr_context* ctx = s->ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (swr::DeviceBuffer* b : {&s->d_pos, &s->d_tri, &s->d_uv, &s->d_geoms, &s->d_shaders, &s->d_lights,
                                 &s->bvh.nodes, &s->bvh.qnodes, &s->bvh.leaf_tris, &s->d_tex_ops, &s->d_tex_images, &s->d_texels}) {
      try {
        swr::release(ctx, *b);
      } catch (const Error&) {
      }
    }
    ctx->scenes.erase(s);
  }
  delete s;
}

int swr_scene_add_shader(swr_scene* s, const swr_shader_desc* d) {
  if (!s || !d || !s->ctx) return SWR_E_INVALID;
  if (s->committed || d->type < 0 || d->type > SWR_SHADER_GLOSSY) {
    s->ctx->err = "scene_add_shader: scene committed or unknown shader type";
    return SWR_E_INVALID;
  }
  s->shaders.push_back(*d);
  return (int)s->shaders.size() - 1;
}

int swr_scene_add_mesh(swr_scene* s, const float* pos4, uint32_t nverts, const uint32_t* idx3, uint32_t ntris,
                       const float* uv2, int32_t shader) {
  if (!s || !s->ctx) return SWR_E_INVALID;
  if (s->committed || (nverts && !pos4) || (ntris && !idx3)) {
    s->ctx->err = "scene_add_mesh: scene committed or null buffer";
    return SWR_E_INVALID;
  }
  for (size_t i = 0; i < 3 * (size_t)ntris; ++i)
    if (idx3[i] >= nverts) {
      s->ctx->err = "scene_add_mesh: index " + std::to_string(idx3[i]) + " out of range";
      return SWR_E_INVALID;
    }
  swr_scene::Mesh m;
  m.pos4.assign(pos4, pos4 + 4 * (size_t)nverts);
  m.idx.assign(idx3, idx3 + 3 * (size_t)ntris);
  if (uv2) m.uv.assign(uv2, uv2 + 2 * (size_t)nverts);
  m.shader = shader;
  s->objects.push_back(std::move(m));
  return (int)s->objects.size() - 1;
}

int swr_scene_add_image(swr_scene* s, int32_t width, int32_t height, const float* rgb) {
  if (!s || !s->ctx) return SWR_E_INVALID;
  if (s->committed || width <= 0 || height <= 0 || !rgb || (int64_t)width * height > (int64_t(1) << 28)) {
    s->ctx->err = "scene_add_image: scene committed, empty image, null texels or more than 2^28 texels";
    return SWR_E_INVALID;
  }
  swr_scene::Image im;
  im.width = width;
  im.height = height;
  im.rgb.assign(rgb, rgb + 3 * (size_t)width * height);
  for (float v : im.rgb)
    if (!std::isfinite(v)) {
      s->ctx->err = "scene_add_image: texels must be finite";
      return SWR_E_INVALID;
    }
  s->images.push_back(std::move(im));
  return (int)s->images.size() - 1;
}

int swr_scene_add_texture(swr_scene* s, const swr_texture_desc* d) {
  if (!s || !d || !s->ctx) return SWR_E_INVALID;
  const int32_t n = (int32_t)s->textures.size();
  const bool ok = !s->committed &&
                  ((d->type == SWR_TEXTURE_CONSTANT) ||
                   (d->type == SWR_TEXTURE_IMAGE && d->image >= 0 && d->image < (int32_t)s->images.size()) ||
                   (d->type == SWR_TEXTURE_LAYERED && d->upper >= 0 && d->upper < n && d->lower >= 0 && d->lower < n &&
                    d->mask >= 0 && d->mask < n));
  if (!ok) {
    s->ctx->err = "scene_add_texture: scene committed, unknown type, or an image / layer index out of range";
    return SWR_E_INVALID;
  }
  s->textures.push_back(*d);
  return n;
}

int swr_scene_set_shader_texture(swr_scene* s, int32_t shader, int32_t texture) {
  if (!s || !s->ctx) return SWR_E_INVALID;
  if (s->committed || shader < 0 || shader >= (int32_t)s->shaders.size() || texture < -1 ||
      texture >= (int32_t)s->textures.size()) {
    s->ctx->err = "scene_set_shader_texture: scene committed or shader / texture index out of range";
    return SWR_E_INVALID;
  }
  if (s->shader_texture.size() < s->shaders.size()) s->shader_texture.resize(s->shaders.size(), -1);
  s->shader_texture[shader] = texture;
  return SWR_OK;
}

int swr_scene_add_quad_light(swr_scene* s, const swr_quad_light_desc* d) {
  if (!s || !d || !s->ctx) return SWR_E_INVALID;
  if (s->committed) {
    s->ctx->err = "scene_add_quad_light: scene committed";
    return SWR_E_INVALID;
  }
  s->light_descs.push_back(*d);
  return (int)s->light_descs.size() - 1;
}

int swr_scene_set_environment(swr_scene* s, const float rgb[3]) {
  if (!s || !rgb || !s->ctx || s->committed) return SWR_E_INVALID;
  std::memcpy(s->env, rgb, 3 * sizeof(float));
  return SWR_OK;
}

int swr_scene_commit(swr_scene* s) {
  if (!s || !s->ctx) return SWR_E_INVALID;
  return scene_guard(s->ctx, [&] {
    if (s->committed) swr::fail(SWR_E_INVALID, "scene already committed");
    swr::commit_scene(s);
  });
}

uint32_t swr_scene_num_triangles(const swr_scene* s) { return s && s->committed ? s->dev.num_tris : 0; }
uint32_t swr_scene_num_geometries(const swr_scene* s) { return s && s->committed ? s->dev.num_geoms : 0; }

int swr_scene_bounds(const swr_scene* s, float out6[6]) {
  if (!s || !out6 || !s->committed) return SWR_E_INVALID;
  std::memcpy(out6, s->dev.lo, 3 * sizeof(float));
  std::memcpy(out6 + 3, s->dev.hi, 3 * sizeof(float));
  return SWR_OK;
}

int swr_scene_set_stack_ring(swr_scene* s, int32_t ring) {
  if (!s || !(ring == 0 || ring == 8 || ring == 16 || ring == 32 || ring == 64)) return SWR_E_INVALID;
  s->tuned_ring = ring;
  return SWR_OK;
}

int swr_scene_stack_ring(const swr_scene* s, int32_t* ring) {
  if (!s || !ring) return SWR_E_INVALID;
  *ring = s->tuned_ring;
  return SWR_OK;
}

int swr_scene_prim_to_geom(const swr_scene* s, uint32_t prim, uint32_t* geom_id, uint32_t* prim_id) {
  if (!s || !s->committed || !geom_id || !prim_id || prim >= s->dev.num_tris) return SWR_E_INVALID;
  auto it = std::upper_bound(s->geoms.begin(), s->geoms.end(), prim,
                             [](uint32_t p, const swr::GeomRec& g) { return p < g.first_tri; });
  const uint32_t g = (uint32_t)(it - s->geoms.begin()) - 1u;
  *geom_id = g;
  *prim_id = prim - s->geoms[g].first_tri;
  return SWR_OK;
}

// Validates the BVH on the host: every triangle in exactly one leaf, leaf boxes contain their
// triangles, child node boxes lie inside the parent's slot box, the leaf array is a permutation
// followed by the null leaf, the quantised copy conservative with the null leaf in unused slots.
int swr_scene_bvh_info(const swr_scene* s, uint32_t* nodes_out, uint32_t* depth_out, uint32_t* valid_out) {
  if (!s || !s->committed) return SWR_E_INVALID;
  swr_context* ctx = s->ctx;
  return scene_guard(ctx, [&] {
    const uint32_t nn = s->bvh.num_nodes, nt = s->dev.num_tris;
    std::vector<swr::Node4> nodes(nn);
    std::vector<swr::NodeQ> qnodes(nn);
    std::vector<float4> lt(3 * (size_t)nt);
    SWR_HIP(hipMemcpy(nodes.data(), s->bvh.nodes.ptr, nn * sizeof(swr::Node4), hipMemcpyDeviceToHost));
    SWR_HIP(hipMemcpy(qnodes.data(), s->bvh.qnodes.ptr, nn * sizeof(swr::NodeQ), hipMemcpyDeviceToHost));
    if (nt) SWR_HIP(hipMemcpy(lt.data(), s->bvh.leaf_tris.ptr, lt.size() * sizeof(float4), hipMemcpyDeviceToHost));
    bool ok = true;
    {  // the null leaf after the triangles: all zeros
      float4 null_tri[3];
      SWR_HIP(hipMemcpy(null_tri, static_cast<const float4*>(s->bvh.leaf_tris.ptr) + 3 * (size_t)nt, sizeof null_tri,
                        hipMemcpyDeviceToHost));
      const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
      for (const float4& v : null_tri)
        if (std::memcmp(&v, &zero, sizeof v) != 0) ok = false;
    }
    std::vector<uint8_t> seen_leaf(nt, 0), seen_id(nt, 0), seen_node(nn, 0);
    for (uint32_t j = 0; j < nt; ++j) {
      uint32_t id;
      std::memcpy(&id, &lt[3 * (size_t)j].w, 4);
      if (id >= nt || seen_id[id]++) ok = false;
    }
    struct Item {
      uint32_t node, depth;
    };
    std::vector<Item> stack{{0u, 0u}};
    uint32_t max_depth = 0;
    const float eps = 0.f;
    while (!stack.empty() && ok) {
      Item it = stack.back();
      stack.pop_back();
      if (it.node >= nn || seen_node[it.node]++) {
        ok = false;
        break;
      }
      max_depth = std::max(max_depth, it.depth);
      const swr::Node4& nd = nodes[it.node];
      const float* lx = &nd.lox.x; const float* hx = &nd.hix.x;
      const float* ly = &nd.loy.x; const float* hy = &nd.hiy.x;
      const float* lz = &nd.loz.x; const float* hz = &nd.hiz.x;
      const uint32_t* ch = &nd.child.x;
      // The quantised copy: same children (the null leaf in unused slots), planes at least one
      // quantum outside every child box.
      const swr::NodeQ& qn = qnodes[it.node];
      const uint32_t* qch = &qn.child.x;
      for (int c = 0; c < 4; ++c)
        if (qch[c] != (ch[c] == swr::kNoChild ? swr::null_leaf_code(nt) : ch[c])) ok = false;
      {
        const float* los[3] = {lx, ly, lz};
        const float* his[3] = {hx, hy, hz};
        const float ps[3] = {qn.px, qn.py, qn.pz};
        const uint32_t qlo[3] = {qn.qlox, qn.qloy, qn.qloz}, qhi[3] = {qn.qhix, qn.qhiy, qn.qhiz};
        for (int a = 0; a < 3; ++a) {
          const float sq[3] = {qn.sx, qn.sy, qn.sz};
          uint32_t e;
          std::memcpy(&e, &sq[a], 4);
          if (e & 0x807FFFFFu) ok = false;  // a power of two
          e >>= 23;
          if (e == 0u || e == 255u) ok = false;
          const double q = std::ldexp(1.0, (int)e - 127);
          for (int c = 0; c < 4; ++c) {
            if (ch[c] == swr::kNoChild) {  // an empty interval
              if (((qlo[a] >> (8 * c)) & 0xFFu) != 255u || ((qhi[a] >> (8 * c)) & 0xFFu) != 0u) ok = false;
              continue;
            }
            const double l = (double)ps[a] + (double)((qlo[a] >> (8 * c)) & 0xFFu) * q;
            const double h = (double)ps[a] + (double)((qhi[a] >> (8 * c)) & 0xFFu) * q;
            if (!(l <= (double)los[a][c] - q && h >= (double)his[a][c] + q)) ok = false;
          }
        }
      }
      for (int c = 0; c < 4; ++c) {
        if (ch[c] == swr::kNoChild) continue;
        if (ch[c] & swr::kLeafBit) {
          const uint32_t f = swr::leaf_first(ch[c]), k = swr::leaf_count(ch[c]);
          for (uint32_t j = f; j < f + k; ++j) {
            if (j >= nt || seen_leaf[j]++) {
              ok = false;
              break;
            }
            const float4 v0 = lt[3 * (size_t)j], e1 = lt[3 * (size_t)j + 1], e2 = lt[3 * (size_t)j + 2];
            // v1 = v0 - e1 and v2 = v0 + e2 within rounding
            const float px[3] = {v0.x, v0.x - e1.x, v0.x + e2.x}, py[3] = {v0.y, v0.y - e1.y, v0.y + e2.y},
                        pz[3] = {v0.z, v0.z - e1.z, v0.z + e2.z};
            for (int q = 0; q < 3; ++q) {
              const float tol = 1e-5f * (1.f + std::fabs(px[q]) + std::fabs(py[q]) + std::fabs(pz[q]));
              if (px[q] < lx[c] - tol || px[q] > hx[c] + tol || py[q] < ly[c] - tol || py[q] > hy[c] + tol ||
                  pz[q] < lz[c] - tol || pz[q] > hz[c] + tol)
                ok = false;
            }
          }
        } else {
          if (ch[c] >= nn) {
            ok = false;
            break;
          }
          const swr::Node4& cn = nodes[ch[c]];
          const float* clx = &cn.lox.x; const float* chx = &cn.hix.x;
          const float* cly = &cn.loy.x; const float* chy = &cn.hiy.x;
          const float* clz = &cn.loz.x; const float* chz = &cn.hiz.x;
          for (int q = 0; q < 4; ++q) {
            if (cn.child.x == swr::kNoChild && q == 0) continue;
            const uint32_t* cc = &cn.child.x;
            if (cc[q] == swr::kNoChild) continue;
            if (clx[q] < lx[c] - eps || chx[q] > hx[c] + eps || cly[q] < ly[c] - eps || chy[q] > hy[c] + eps ||
                clz[q] < lz[c] - eps || chz[q] > hz[c] + eps)
              ok = false;
          }
          stack.push_back({ch[c], it.depth + 1});
        }
      }
    }
    for (uint32_t j = 0; j < nt && ok; ++j)
      if (!seen_leaf[j]) ok = false;
    if (nodes_out) *nodes_out = nn;
    if (depth_out) *depth_out = max_depth;
    if (valid_out) *valid_out = ok ? 1u : 0u;
    if (ok && max_depth != s->bvh.max_depth) *valid_out = 0u;
  });
}

int swr_intersect(swr_context* ctx, const swr_scene* s, const float* rays_dev, size_t n, uint32_t mask,
                  uint32_t* hits_dev) {
  if (!ctx || !s || s->ctx != ctx) return SWR_E_INVALID;
  return scene_guard(ctx, [&] {
    if (!s->committed) swr::fail(SWR_E_INVALID, "intersect: scene not committed");
    if (n && (!rays_dev || !hits_dev || !aligned16(rays_dev) || !aligned16(hits_dev)))
      swr::fail(SWR_E_INVALID, "intersect: rays and hits must be 16-byte aligned device buffers");
    if (n > 0xFFFFFFFFull) swr::fail(SWR_E_INVALID, "intersect: too many rays");
    swr::Timer timer(ctx);
    swr::launch_intersect(ctx, s->dev, reinterpret_cast<const float4*>(rays_dev), (uint32_t)n, mask,
                          reinterpret_cast<uint4*>(hits_dev), nullptr, ctx->stream);
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

int swr_occluded(swr_context* ctx, const swr_scene* s, const float* rays_dev, size_t n, uint32_t mask,
                 uint32_t* flags_dev) {
  if (!ctx || !s || s->ctx != ctx) return SWR_E_INVALID;
  return scene_guard(ctx, [&] {
    if (!s->committed) swr::fail(SWR_E_INVALID, "occluded: scene not committed");
    if (n && (!rays_dev || !flags_dev || !aligned16(rays_dev)))
      swr::fail(SWR_E_INVALID, "occluded: rays must be a 16-byte aligned device buffer");
    if (n > 0xFFFFFFFFull) swr::fail(SWR_E_INVALID, "occluded: too many rays");
    swr::Timer timer(ctx);
    swr::launch_occluded(ctx, s->dev, reinterpret_cast<const float4*>(rays_dev), (uint32_t)n, mask, flags_dev,
                         ctx->stream);
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

}  // extern "C"

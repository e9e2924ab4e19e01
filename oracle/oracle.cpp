/*
 * oracle.cpp -- CPU restatement of joshbainbridge/msc-project's path tracer (test oracle and
 * CPU baseline; never part of the product path).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   philox      Random123 known-answer vectors for Philox-4x32-10.
 *   morton/keys closed-form bit interleave + brute force on small inputs.
 *   sort        std::stable_sort (textbook).
 *   intersect   FP64 brute force (oracle_intersect_brute), a linear solve independent of the
 *               float Moller-Trumbore used by the BVH path.
 *   raygen      closed form of the Grid sampler's pixel-centre rays; thin-lens focal-plane law.
 *   filter      closed forms on synthetic sample sets.
 *   integrator  white-furnace closed form; analytic rectangle-to-point irradiance;
 *               russian-roulette unbiasedness.
 *   ggx lobe    D normalisation and Smith's visible-area identity by quadrature; pdf against
 *               the sampler (acceptance integral, chi-square); the sampled-weight identity;
 *               reciprocity; albedo bounds; planted mistakes built into copies fail.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Constants.h:10-15
const float kPi = 3.14159265358979323846f;
const float kInvPi = 0.31830988618379067154f;
const float kPi180 = 0.01745329251994329577f;
const float kEps = 1.19209290e-07f;  // FLT_EPSILON (M_EPSILON)

struct V3 {
  float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
  float n = norm(a);
  return n > 0.f ? a * (1.f / n) : a;
}
inline float maxc(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }

// 3x3 rotation, row-major.
struct M3 {
  float m[3][3];
};
M3 mmul(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}
M3 rot_axis(int axis, float deg) {
  float a = deg * kPi180, c = std::cos(a), s = std::sin(a);
  M3 r{};
  if (axis == 0) {
    r.m[0][0] = 1; r.m[1][1] = c; r.m[1][2] = -s; r.m[2][1] = s; r.m[2][2] = c;
  } else if (axis == 1) {
    r.m[0][0] = c; r.m[0][2] = s; r.m[1][1] = 1; r.m[2][0] = -s; r.m[2][2] = c;
  } else {
    r.m[0][0] = c; r.m[0][1] = -s; r.m[1][0] = s; r.m[1][1] = c; r.m[2][2] = 1;
  }
  return r;
}
inline V3 apply(const M3& r, V3 v) {
  return v3(r.m[0][0] * v.x + r.m[0][1] * v.y + r.m[0][2] * v.z,
            r.m[1][0] * v.x + r.m[1][1] * v.y + r.m[1][2] * v.z,
            r.m[2][0] * v.x + r.m[2][1] * v.y + r.m[2][2] * v.z);
}

// ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----------------------------------
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}
inline void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    philox_round(c, k);
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}
inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Four uniforms for (seed, sample, iteration, word2, word3).
struct Rand4 {
  float r[4];
};
inline Rand4 rand4(uint64_t seed, uint32_t sid, uint32_t iter, uint32_t w2, uint32_t w3) {
  uint32_t ctr[4] = {sid, iter, w2, w3};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox10(ctr, key, o);
  return Rand4{{u01(o[0]), u01(o[1]), u01(o[2]), u01(o[3])}};
}
// Counter word 2: camera = 0; shading of a ray at depth d uses ((d+1) << 8) | slot.
inline uint32_t shade_word(int depth, int slot) { return ((uint32_t)(depth + 1) << 8) | (uint32_t)slot; }

// ---- scene -------------------------------------------------------------------------------
struct Light {
  M3 rot;          // Rx*Ry*Rz (QuadLight.cpp:24-28)
  V3 t;            // translation
  float sx, sy;    // AlignedScaling3f(1, sx, sy) (QuadLight.cpp:29)
  V3 n;            // inverse-transpose * x-hat, normalised (QuadLight.cpp:31)
  float intensity;
  // transform(p) = t + rot * (p.x, sx*p.y, sy*p.z)
  V3 xf(V3 p) const { return t + apply(rot, v3(p.x, sx * p.y, sy * p.z)); }
};

struct Tri {
  V3 v0, v1, v2;
};

struct Node {  // BVH2, leaves hold [first, first+count) in prim order
  float lo[3], hi[3];
  int32_t left;    // child index or -1 for a leaf
  int32_t right;
  uint32_t first, count;
};

}  // namespace

struct oracle_scene {
  struct Mesh {
    std::vector<float> pos4;
    std::vector<uint32_t> idx;
    std::vector<float> uv;
    int32_t shader;
  };
  std::vector<Mesh> objects;
  std::vector<oracle_shader_desc> shaders;
  std::vector<oracle_quad_light_desc> light_descs;
  struct Image {
    int32_t width, height;
    std::vector<float> rgb;
  };
  std::vector<Image> images;
  std::vector<oracle_texture_desc> textures;
  std::vector<int32_t> shader_texture;           // per shader: texture node or -1
  V3 env{0.f, 0.f, 0.f};
  bool committed = false;

  // committed layout (geomID order: objects then lights; Pathtracer.cpp:166-272)
  std::vector<Light> lights;
  std::vector<oracle_shader_desc> all_shaders;   // shaders, then one Null per light
  std::vector<int32_t> geom_shader;              // per geomID
  std::vector<uint32_t> geom_mask;               // 0xFFFFFFFF objects, 0xF0000000 lights
  std::vector<int32_t> geom_light;               // light index or -1
  std::vector<uint32_t> geom_first_tri;          // prefix of triangle counts
  std::vector<Tri> tris;                         // global triangle order
  std::vector<float> tri_uv;                     // 6 floats per triangle: uv of v0, v1, v2
  std::vector<int32_t> tri_geom;
  std::vector<Node> nodes;
  std::vector<uint32_t> order;                   // BVH leaf order -> global triangle id
};

namespace {

// QuadLight::construct (QuadLight.cpp:5-39)
Light make_light(const oracle_quad_light_desc& d) {
  Light L;
  L.rot = mmul(mmul(rot_axis(0, d.rotation[0]), rot_axis(1, d.rotation[1])), rot_axis(2, d.rotation[2]));
  L.t = v3(d.translation[0], d.translation[1], d.translation[2]);
  L.sx = d.scale[0];
  L.sy = d.scale[1];
  // inverse-transpose of rot*diag(1,sx,sy) applied to x-hat is rot*(1,0,0) (scale 1 on x).
  L.n = normalize(apply(L.rot, v3(1.f, 0.f, 0.f)));
  L.intensity = d.intensity;
  return L;
}

// ---- triangle tests ------------------------------------------------------------------------
// Float Moller-Trumbore in the arrangement of Embree 2's triangle intersector: e1 = v0-v1,
// e2 = v2-v0, Ng = e1 x e2, hit = (1-u-v) v0 + u v1 + v v2, accepted when tnear < t < tfar.
inline bool tri_hit(const Tri& tr, V3 o, V3 d, float tnear, float tfar, float* t, float* u, float* v) {
  V3 e1 = tr.v0 - tr.v1, e2 = tr.v2 - tr.v0;
  V3 ng = cross(e1, e2);
  V3 c = tr.v0 - o;
  V3 r = cross(d, c);
  float den = dot(ng, d);
  float aden = std::fabs(den);
  float sgn = den < 0.f ? -1.f : 1.f;
  float U = dot(r, e2) * sgn;
  float V = dot(r, e1) * sgn;
  if (!(den != 0.f && U >= 0.f && V >= 0.f && U + V <= aden)) return false;
  float T = dot(ng, c) * sgn;
  if (!(T > aden * tnear && T < aden * tfar)) return false;
  float inv = 1.f / aden;
  *t = T * inv;
  *u = U * inv;
  *v = V * inv;
  return true;
}

inline bool box_hit(const Node& n, V3 o, V3 inv, float tnear, float tfar) {
  float t0 = tnear, t1 = tfar;
  const float oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
  for (int a = 0; a < 3; ++a) {
    float ta = (n.lo[a] - oo[a]) * ii[a];
    float tb = (n.hi[a] - oo[a]) * ii[a];
    if (ta > tb) std::swap(ta, tb);
    if (std::isnan(ta)) ta = -INFINITY;  // 0*inf on a slab plane
    if (std::isnan(tb)) tb = INFINITY;
    t0 = std::max(t0, ta);
    t1 = std::min(t1, tb);
  }
  return t0 <= t1 * 1.00000024f;
}

// ---- binned SAH BVH build ------------------------------------------------------------------
struct BuildPrim {
  float lo[3], hi[3], c[3];
};

void grow(float lo[3], float hi[3], const float plo[3], const float phi[3]) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = std::min(lo[a], plo[a]);
    hi[a] = std::max(hi[a], phi[a]);
  }
}
float area(const float lo[3], const float hi[3]) {
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  if (dx < 0 || dy < 0 || dz < 0) return 0.f;
  return dx * dy + dy * dz + dz * dx;
}

void build_bvh(oracle_scene* s) {
  const uint32_t n = (uint32_t)s->tris.size();
  std::vector<BuildPrim> bp(n);
  for (uint32_t i = 0; i < n; ++i) {
    const Tri& t = s->tris[i];
    const V3 vs[3] = {t.v0, t.v1, t.v2};
    for (int a = 0; a < 3; ++a) {
      float lo = INFINITY, hi = -INFINITY;
      for (int k = 0; k < 3; ++k) {
        float c = a == 0 ? vs[k].x : (a == 1 ? vs[k].y : vs[k].z);
        lo = std::min(lo, c);
        hi = std::max(hi, c);
      }
      bp[i].lo[a] = lo;
      bp[i].hi[a] = hi;
      bp[i].c[a] = 0.5f * (lo + hi);
    }
  }
  s->order.resize(n);
  std::iota(s->order.begin(), s->order.end(), 0u);
  s->nodes.clear();
  s->nodes.reserve(n ? 2 * n : 1);
  struct Item {
    uint32_t node, first, count;
  };
  std::vector<Item> stack;
  Node root{};
  s->nodes.push_back(root);
  stack.push_back({0, 0, n});
  const int kBins = 16;
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    Node& nd = s->nodes[it.node];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = it.first; i < it.first + it.count; ++i) {
      const BuildPrim& p = bp[s->order[i]];
      grow(lo, hi, p.lo, p.hi);
      grow(clo, chi, p.c, p.c);
    }
    std::memcpy(nd.lo, lo, sizeof lo);
    std::memcpy(nd.hi, hi, sizeof hi);
    nd.left = nd.right = -1;
    nd.first = it.first;
    nd.count = it.count;
    if (it.count <= 4) continue;
    int best_axis = -1, best_split = 0;
    float best_cost = (float)it.count * area(lo, hi);
    for (int a = 0; a < 3; ++a) {
      float ext = chi[a] - clo[a];
      if (!(ext > 0.f)) continue;
      float blo[kBins][3], bhi[kBins][3];
      uint32_t bcnt[kBins] = {0};
      for (int b = 0; b < kBins; ++b)
        for (int k = 0; k < 3; ++k) { blo[b][k] = INFINITY; bhi[b][k] = -INFINITY; }
      for (uint32_t i = it.first; i < it.first + it.count; ++i) {
        const BuildPrim& p = bp[s->order[i]];
        int b = std::min(kBins - 1, (int)((p.c[a] - clo[a]) / ext * kBins));
        bcnt[b]++;
        grow(blo[b], bhi[b], p.lo, p.hi);
      }
      float rarea[kBins];
      uint32_t rcnt[kBins];
      float alo[3] = {INFINITY, INFINITY, INFINITY}, ahi[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t acc = 0;
      for (int b = kBins - 1; b > 0; --b) {
        grow(alo, ahi, blo[b], bhi[b]);
        acc += bcnt[b];
        rarea[b] = area(alo, ahi);
        rcnt[b] = acc;
      }
      float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t lc = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        grow(llo, lhi, blo[b], bhi[b]);
        lc += bcnt[b];
        if (lc == 0 || rcnt[b + 1] == 0) continue;
        float cost = lc * area(llo, lhi) + rcnt[b + 1] * rarea[b + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = a;
          best_split = b;
        }
      }
    }
    uint32_t mid;
    if (best_axis < 0) {
      // no useful split: median on the widest centroid axis
      int a = 0;
      for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[a] - clo[a]) a = k;
      mid = it.first + it.count / 2;
      std::nth_element(s->order.begin() + it.first, s->order.begin() + mid,
                       s->order.begin() + it.first + it.count,
                       [&](uint32_t x, uint32_t y) { return bp[x].c[a] < bp[y].c[a]; });
    } else {
      float ext = chi[best_axis] - clo[best_axis];
      auto pivot = std::partition(s->order.begin() + it.first, s->order.begin() + it.first + it.count,
                                  [&](uint32_t x) {
                                    int b = std::min(kBins - 1,
                                                     (int)((bp[x].c[best_axis] - clo[best_axis]) / ext * kBins));
                                    return b <= best_split;
                                  });
      mid = (uint32_t)(pivot - s->order.begin());
      if (mid == it.first || mid == it.first + it.count) mid = it.first + it.count / 2;
    }
    int32_t l = (int32_t)s->nodes.size();
    s->nodes.push_back(Node{});
    s->nodes.push_back(Node{});
    Node& nd2 = s->nodes[it.node];
    nd2.left = l;
    nd2.right = l + 1;
    nd2.count = 0;
    stack.push_back({(uint32_t)l, it.first, mid - it.first});
    stack.push_back({(uint32_t)l + 1, mid, it.first + it.count - mid});
  }
}

// Closest hit (any_hit = false) or occlusion (any_hit = true) through the BVH.
bool trace(const oracle_scene* s, V3 o, V3 d, float tnear, float tfar, uint32_t mask, bool any_hit,
           float* ot, float* ou, float* ov, uint32_t* oprim) {
  if (s->nodes.empty() || s->tris.empty()) return false;
  V3 inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);
  const float ray_tfar = tfar;
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  bool found = false;
  while (sp) {
    const Node& n = s->nodes[stack[--sp]];
    if (!box_hit(n, o, inv, tnear, tfar)) continue;
    if (n.left < 0) {
      for (uint32_t i = n.first; i < n.first + n.count; ++i) {
        uint32_t prim = s->order[i];
        if ((s->geom_mask[s->tri_geom[prim]] & mask) == 0) continue;
        float t, u, v;
        if (tri_hit(s->tris[prim], o, d, tnear, ray_tfar, &t, &u, &v)) {
          // closest first; at exactly equal t the lower triangle id wins (documented tie rule)
          if (!found || t < tfar || (t == tfar && prim < *oprim)) {
            tfar = t;
            *ot = t;
            *ou = u;
            *ov = v;
            *oprim = prim;
            found = true;
            if (any_hit) return true;
          }
        }
      }
    } else {
      stack[sp++] = (uint32_t)n.left;
      stack[sp++] = (uint32_t)n.right;
    }
  }
  return found;
}

// ---- BSDFs (LambertShader.cpp, NullShader.cpp; Glossy is the GGX extension) --------------
struct Eval {
  V3 f;
  float cos;
  float pdf;
};

// Local frame shared by every sampler (LambertShader.cpp:65-67).
inline void frame(V3 n, V3* s, V3* t) {
  V3 temp = (n.x > 0.99f) ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
  *s = normalize(cross(temp, n));
  *t = normalize(cross(n, *s));
}

inline float ggx_d(float a2, float ch) {
  float k = ch * ch * (a2 - 1.f) + 1.f;
  return a2 / (kPi * k * k);
}
inline float ggx_g1(float a2, float c) { return 2.f * c / (c + std::sqrt(a2 + (1.f - a2) * c * c)); }

Eval bsdf_eval(const oracle_shader_desc& sh, V3 in, V3 out, V3 n) {
  Eval e;
  if (sh.type == 1) {  // NullShader::evaluate (NullShader.cpp:25-38)
    e.f = v3(0.f, 0.f, 0.f);
    e.cos = 1.f;
    e.pdf = 1.f * kInvPi * 0.5f;
    return e;
  }
  float ci = dot(n, in), co = dot(n, out);
  V3 col = v3(sh.colour[0], sh.colour[1], sh.colour[2]);
  if (sh.type == 0) {  // LambertShader::evaluate (LambertShader.cpp:27-47)
    if (ci < 0.f || co < 0.f)
      e.f = v3(0.f, 0.f, 0.f);
    else
      e.f = col * sh.reflectance * kInvPi;
    e.cos = ci < 0.f ? 0.f : ci;
    e.pdf = e.cos * kInvPi;
    return e;
  }
  // Glossy: GGX reflection (Walter et al. 2007), F = colour * reflectance
  e.cos = ci < 0.f ? 0.f : ci;
  if (ci <= 0.f || co <= 0.f) {
    e.f = v3(0.f, 0.f, 0.f);
    e.pdf = 0.f;
    return e;
  }
  float a = sh.roughness, a2 = a * a;
  V3 h = normalize(in + out);
  float ch = dot(n, h), oh = dot(out, h);
  float D = ggx_d(a2, ch);
  float G = ggx_g1(a2, ci) * ggx_g1(a2, co);
  e.f = col * (sh.reflectance * D * G / (4.f * ci * co));
  e.pdf = oh > 0.f ? D * ch / (4.f * oh) : 0.f;
  return e;
}

// Returns false when the sampled direction carries no energy (the path ends).
bool bsdf_sample(const oracle_shader_desc& sh, float u1, float u2, V3 out, V3 n, V3* in, Eval* e) {
  V3 s, t;
  frame(n, &s, &t);
  V3 col = v3(sh.colour[0], sh.colour[1], sh.colour[2]);
  if (sh.type == 1) {  // NullShader::sample (NullShader.cpp:40-54): zero weight
    return false;
  }
  if (sh.type == 0) {  // LambertShader::sample (LambertShader.cpp:49-73)
    float azimuth = 2.f * kPi * u1;
    float zenith = std::acos(std::sqrt(u2));
    // sphericalToCartesianCoord (Common.h:28-33)
    float x = std::cos(azimuth) * std::sin(zenith);
    float y = std::sin(azimuth) * std::sin(zenith);
    float z = std::cos(zenith);
    *in = t * x + s * y + n * z;
    e->f = col * sh.reflectance * kInvPi;
    e->cos = dot(n, *in);
    e->pdf = e->cos * kInvPi;
    return e->pdf > 0.f;
  }
  // Glossy: sample the GGX half vector from D(h) cos(theta_h)
  float a = sh.roughness, a2 = a * a;
  float phi = 2.f * kPi * u1;
  float tan2 = a2 * u2 / (1.f - u2);
  float ch = 1.f / std::sqrt(1.f + tan2);
  float sh_ = std::sqrt(std::max(0.f, 1.f - ch * ch));
  V3 h = t * (std::cos(phi) * sh_) + s * (std::sin(phi) * sh_) + n * ch;
  float oh = dot(out, h);
  *in = h * (2.f * oh) - out;
  *e = bsdf_eval(sh, *in, out, n);
  return e->pdf > 0.f && maxc(e->f) > 0.f;
}

// QuadLight::illuminate (QuadLight.cpp:41-67)
void light_illuminate(const Light& L, float ux, float uy, V3 p, V3* dir, float* dist, V3* rad, float* pdfw) {
  V3 smp = L.xf(v3(0.f, ux - 0.5f, uy - 0.5f));
  float c = dot(normalize(p - smp), L.n);
  *rad = c < 0.f ? v3(0.f, 0.f, 0.f) : v3(L.intensity, L.intensity, L.intensity);
  c = c < 0.f ? 0.f : c;
  V3 dv = smp - p;
  *dir = normalize(dv);
  *dist = norm(dv);
  *pdfw = (dot(dv, dv) / c) / (L.sx * L.sy);
}

// ---- camera (Camera.cpp:6-59, PinHoleCamera.cpp, ThinLensCamera.cpp) ------------------------
struct CameraXf {
  M3 rot;  // Ry*Rx*Rz (PinHoleCamera.cpp:7-11)
  V3 t;
  V3 n;
};
CameraXf camera_xf(const oracle_camera_desc& c) {
  CameraXf x;
  x.rot = mmul(mmul(rot_axis(1, c.rotation[1]), rot_axis(0, c.rotation[0])), rot_axis(2, c.rotation[2]));
  x.t = v3(c.translation[0], c.translation[1], c.translation[2]);
  x.n = normalize(apply(x.rot, v3(1.f, 0.f, 0.f)));
  return x;
}

inline int spp_of(const oracle_render_desc* d) { return d->spp > 0 ? d->spp : (int)((int64_t)d->base * d->base); }

// One camera ray for sample id sid (Camera.cpp:15-38).
void camera_ray(const oracle_render_desc* d, const CameraXf& cx, uint32_t sid, V3* org, V3* dir, float* sx,
                float* sy) {
  const uint32_t W = (uint32_t)d->width, H = (uint32_t)d->height, spp = (uint32_t)spp_of(d);
  uint32_t x = sid / (H * spp);
  uint32_t rem = sid - x * H * spp;
  uint32_t y = rem / spp;
  uint32_t k = rem - y * spp;
  Rand4 r = rand4(d->seed, sid, d->iteration, 0u, 0u);
  float jx, jy;
  int base = d->base > 0 ? d->base : 1;
  float cell = 1.f / (float)base;
  uint32_t ix = k / (uint32_t)base, iy = k % (uint32_t)base;
  if (d->sampler == 1) {  // IndependentSampler.cpp:11-12
    jx = r.r[0];
    jy = r.r[1];
  } else if (d->sampler == 2) {  // GridSampler.cpp:13-14
    jx = (float)ix * cell + 0.5f * cell;
    jy = (float)iy * cell + 0.5f * cell;
  } else {  // StratifiedSampler.cpp:13-14
    jx = (float)ix * cell + r.r[0] * cell;
    jy = (float)iy * cell + r.r[1] * cell;
  }
  *sx = (float)x + jx;
  *sy = (float)y + jy;
  float fx = (((float)x + jx) * 2.f - (float)W) / (float)W * 36.f;   // Camera.cpp:34
  float fy = (((float)y + jy) * 2.f - (float)H) / (float)W * 36.f;   // Camera.cpp:35
  const oracle_camera_desc& c = d->camera;
  V3 nodal = cx.t + cx.n * c.focal_length;
  V3 film = cx.t + apply(cx.rot, v3(0.f, fy, fx));
  if (c.type == 0) {  // PinHoleCamera::sample (PinHoleCamera.cpp:16-32)
    *org = nodal;
    *dir = normalize(nodal - film);
    return;
  }
  // ThinLensCamera::sample (ThinLensCamera.cpp:16-56)
  float ratio = c.focal_distance / c.focal_length;
  V3 focal = film + (nodal - film) * ratio;
  float ax = 0.f, ay = 0.f;
  float cx0 = r.r[2] - 0.5f, cy0 = r.r[3] - 0.5f;
  for (uint32_t attempt = 0; attempt < 64; ++attempt) {
    if (attempt) {
      Rand4 q = rand4(d->seed, sid, d->iteration, 0u, attempt);
      cx0 = q.r[0] - 0.5f;
      cy0 = q.r[1] - 0.5f;
    }
    if (std::sqrt(cx0 * cx0 + cy0 * cy0) < 0.5f) {
      ax = cx0;
      ay = cy0;
      break;
    }
  }
  float scale = c.focal_length / c.aperture;
  V3 ap = cx.t + apply(cx.rot, v3(0.f, ax * scale, ay * scale));
  *org = ap + cx.n * c.focal_length;
  *dir = normalize(focal - *org);
}

// Tent/box resolve of one pixel (TentFilter.cpp:17-61, BoxFilter.cpp:17-36).
void filter_pixel(int32_t filter, int32_t W, int32_t H, int32_t spp, const float* samples, int px, int py,
                  float out[3]) {
  if (filter == 1) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int i = 0; i < spp; ++i) {
      size_t si = ((size_t)px * H * spp + (size_t)py * spp + i) * 5;
      s0 += samples[si + 2];
      s1 += samples[si + 3];
      s2 += samples[si + 4];
    }
    out[0] = s0 / (float)spp;
    out[1] = s1 / (float)spp;
    out[2] = s2 / (float)spp;
    return;
  }
  float sum[3] = {0.f, 0.f, 0.f};
  float cxp = (float)px + 0.5f, cyp = (float)py + 0.5f;
  float den = 0.f;
  for (int mx = -1; mx < 2; ++mx)
    for (int my = -1; my < 2; ++my) {
      int qx = px + mx, qy = py + my;
      if (qx < 0 || qx >= W || qy < 0 || qy >= H) continue;
      for (int i = 0; i < spp; ++i) {
        size_t si = ((size_t)qx * H * spp + (size_t)qy * spp + i) * 5;
        float dx = std::fmax(0.f, 1.f - std::fabs(samples[si + 0] - cxp));
        float dy = std::fmax(0.f, 1.f - std::fabs(samples[si + 1] - cyp));
        float w = dx * dy;
        den = den + w;
        sum[0] += samples[si + 2] * w;
        sum[1] += samples[si + 3] * w;
        sum[2] += samples[si + 4] * w;
      }
    }
  out[0] = sum[0] / den;
  out[1] = sum[1] / den;
  out[2] = sum[2] / den;
}

// ---- textures (StandardTexture.cpp:11-41, LayeredTexture.cpp:14-32, ConstantTexture.cpp:22-25) ----
// OpenImageIO's texture() with WrapPeriodic and zero derivatives samples the finest level with
// its bicubic filter: texel centres at (i + 0.5) / n, uniform cubic B-spline weights
//   (1-f)^3/6, (3f^3 - 6f^2 + 4)/6, (-3f^3 + 3f^2 + 3f + 1)/6, f^3/6
// over texels i-1 .. i+2 around x = s * n - 0.5 = i + f, indices taken modulo n. The coordinate
// is first reduced to [0, 1] (s - floor(s)), which the periodic wrap leaves unchanged.
void bspline_axis(float s, int32_t n, int32_t idx[4], float w[4]) {
  if (!std::isfinite(s)) s = 0.f;
  const float x = (s - std::floor(s)) * (float)n - 0.5f;
  const float fx = std::floor(x);
  const float f = x - fx;
  const float g = 1.f - f, f2 = f * f, f3 = f2 * f;
  w[0] = g * g * g * (1.f / 6.f);
  w[1] = (3.f * f3 - 6.f * f2 + 4.f) * (1.f / 6.f);
  w[2] = (-3.f * f3 + 3.f * f2 + 3.f * f + 1.f) * (1.f / 6.f);
  w[3] = f3 * (1.f / 6.f);
  const int32_t i = (int32_t)fx;
  for (int k = 0; k < 4; ++k) {
    int32_t j = (i - 1 + k) % n;
    idx[k] = j < 0 ? j + n : j;
  }
}

V3 image_lookup(const float* rgb, int32_t width, int32_t height, float s, float t) {
  int32_t xi[4], yi[4];
  float xw[4], yw[4];
  bspline_axis(s, width, xi, xw);
  bspline_axis(t, height, yi, yw);
  V3 acc = v3(0.f, 0.f, 0.f);
  for (int j = 0; j < 4; ++j) {
    V3 row = v3(0.f, 0.f, 0.f);
    for (int i = 0; i < 4; ++i) {
      const float* c = rgb + 3 * ((size_t)yi[j] * width + xi[i]);
      row = row + v3(c[0], c[1], c[2]) * xw[i];
    }
    acc = acc + row * yw[j];
  }
  return acc;
}

V3 texture_eval(const oracle_scene* s, int32_t node, float tu, float tv) {
  const oracle_texture_desc& t = s->textures[node];
  if (t.type == 0) return v3(t.colour[0], t.colour[1], t.colour[2]);
  if (t.type == 1) {
    const oracle_scene::Image& im = s->images[t.image];
    return image_lookup(im.rgb.data(), im.width, im.height, tu, tv);
  }
  const V3 up = texture_eval(s, t.upper, tu, tv);
  const V3 lo = texture_eval(s, t.lower, tu, tv);
  const V3 m = texture_eval(s, t.mask, tu, tv);
  const V3 inv = v3(1.f - m.x, 1.f - m.y, 1.f - m.z);
  return mul(up, m) + mul(lo, inv);
}

// The shader at a hit with its texture applied: the texture coordinate of PolygonObject::texture
// (PolygonObject.cpp:45-63) from the hit barycentrics, then the texture's colour replaces the
// shader's constant colour (LambertShader.cpp:43, :70).
oracle_shader_desc shade_with_texture(const oracle_scene* s, int32_t shader, uint32_t prim, float u, float v) {
  oracle_shader_desc sh = s->all_shaders[shader];
  const int32_t node = shader < (int32_t)s->shader_texture.size() ? s->shader_texture[shader] : -1;
  if (node < 0) return sh;
  const float* uv = &s->tri_uv[6 * (size_t)prim];
  const float w0 = 1.f - u - v;
  const float tu = w0 * uv[0] + u * uv[2] + v * uv[4];
  const float tv = w0 * uv[1] + u * uv[3] + v * uv[5];
  const V3 c = texture_eval(s, node, tu, tv);
  sh.colour[0] = c.x;
  sh.colour[1] = c.y;
  sh.colour[2] = c.z;
  return sh;
}

// One path (Integrator.cpp:12-255 restated per ray, path at a time).
V3 trace_path(const oracle_scene* s, const oracle_render_desc* d, uint32_t sid, V3 org, V3 dir, uint64_t* n_closest,
              uint64_t* n_shadow) {
  V3 L = v3(0.f, 0.f, 0.f);
  V3 w = v3(1.f, 1.f, 1.f);
  float last_pdf = 1.f;
  int depth = 0;
  const size_t nl = s->lights.size();
  const float pick = nl ? 1.f / (float)nl : 0.f;
  for (;;) {
    float t, u, v;
    uint32_t prim = 0xFFFFFFFFu;
    ++*n_closest;
    // RayDecompress.cpp:15-20: tnear 0.001, tfar 1e5, mask all
    if (!trace(s, org, dir, 0.001f, 100000.f, 0xFFFFFFFFu, false, &t, &u, &v, &prim)) {
      // Integrator.cpp:26-27: misses add nothing; a constant environment is the extension.
      L = L + mul(w, s->env);
      break;
    }
    int32_t g = s->tri_geom[prim];
    int32_t li = s->geom_light[g];
    V3 dn = normalize(dir);
    if (li >= 0) {  // Integrator.cpp:37-75
      const Light& lt = s->lights[li];
      float c = dot(lt.n, dn * -1.f);  // QuadLight::radiance (QuadLight.cpp:76-83)
      V3 le = c < 0.f ? v3(0.f, 0.f, 0.f) : v3(lt.intensity, lt.intensity, lt.intensity);
      float pdfa = 1.f / (lt.sx * lt.sy);
      float mis = 1.f;
      if (depth > 0) {
        float pdfw = pdfa * (t * t) / std::fabs(c);  // Common.h:35-38
        mis = last_pdf / (last_pdf + pdfw * pick);  // Common.h:45-48
      }
      if (maxc(le) > kEps) L = L + v3(le.x * mis * w.x, le.y * mis * w.y, le.z * mis * w.z);
      break;
    }
    const oracle_shader_desc sh = shade_with_texture(s, s->geom_shader[g], prim, u, v);
    const Tri& tr = s->tris[prim];
    V3 ng = cross(tr.v0 - tr.v1, tr.v2 - tr.v0);
    V3 n = normalize(ng) * -1.f;  // Integrator.cpp:121-125
    V3 p = org + dn * t;          // Integrator.cpp:126
    V3 out = dn * -1.f;
    Rand4 r = rand4(d->seed, sid, d->iteration, shade_word(depth, 1), 0u);
    // Next event estimation (Integrator.cpp:102-178)
    if (nl) {
      size_t l = std::min(nl - 1, (size_t)((float)nl * r.r[0]));
      V3 in;
      float dist, pdfw;
      V3 le;
      light_illuminate(s->lights[l], r.r[1], r.r[2], p, &in, &dist, &le, &pdfw);
      if (maxc(le) > kEps && std::isfinite(pdfw)) {
        Eval e = bsdf_eval(sh, in, out, n);
        if (maxc(e.f) > kEps) {
          float mis = (pdfw * pick) / (pdfw * pick + e.pdf);
          float k = mis * e.cos / (pdfw * pick);
          V3 contrib = v3(k * (le.x * e.f.x), k * (le.y * e.f.y), k * (le.z * e.f.z));
          ++*n_shadow;
          float st, su, sv;
          uint32_t sp;
          // shadow ray: tnear 0.001, tfar distance, mask 0x0FFFFFFF (Integrator.cpp:153-167)
          if (!trace(s, p, in, 0.001f, dist, 0x0FFFFFFFu, true, &st, &su, &sv, &sp))
            L = L + mul(contrib, w);
        }
      }
    }
    // Continuation (Integrator.cpp:181-246)
    if (depth >= d->max_depth) break;
    float cont = sh.type == 1 ? 1.f : sh.reflectance;
    float q = std::fmin(1.f, cont / d->threshold);
    if (depth < d->min_depth) q = 1.f;
    if (r.r[3] > q) break;
    Rand4 r2 = rand4(d->seed, sid, d->iteration, shade_word(depth, 2), 0u);
    V3 in;
    Eval e;
    if (!bsdf_sample(sh, r2.r[0], r2.r[1], out, n, &in, &e)) break;
    float k = e.cos / e.pdf;
    w = v3(w.x * e.f.x * k / q, w.y * e.f.y * k / q, w.z * e.f.z * k / q);
    if (!(maxc(w) > 0.f)) break;
    last_pdf = e.pdf;
    depth += 1;
    org = p;
    dir = in;
  }
  return L;
}

}  // namespace

// ===================================== C API ===================================================
extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox10(ctr, key, out); }
float oracle_u01(uint32_t x) { return u01(x); }

uint32_t oracle_morton29(uint32_t qx, uint32_t qy, uint32_t qz) {
  // bits from the top: x9 y9 z8 x8 y8 z7 ... x1 y1 z0 x0 y0
  uint32_t code = 0;
  for (int i = 9; i >= 0; --i) {
    code = (code << 1) | ((qx >> i) & 1u);
    code = (code << 1) | ((qy >> i) & 1u);
    if (i >= 1) code = (code << 1) | ((qz >> (i - 1)) & 1u);
  }
  return code;
}

void oracle_ray_keys(const float* rays, size_t n, const float lo[3], const float scale[3], uint32_t* keys) {
  const float qmax[3] = {1023.f, 1023.f, 511.f};
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + 8 * i;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      float f = std::floor((r[a] - lo[a]) * scale[a]);
      f = std::fmin(std::fmax(f, 0.f), qmax[a]);
      q[a] = (uint32_t)f;
    }
    uint32_t oct = (r[4] < 0.f ? 1u : 0u) | (r[5] < 0.f ? 2u : 0u) | (r[6] < 0.f ? 4u : 0u);
    keys[i] = (oct << 29) | oracle_morton29(q[0], q[1], q[2]);
  }
}

void oracle_hit_keys(const uint32_t* prim, size_t n, uint32_t* keys) {
  for (size_t i = 0; i < n; ++i) keys[i] = prim[i] + 1u;  // 0xFFFFFFFF (miss) wraps to 0
}

void oracle_stable_sort_perm(const uint32_t* keys, size_t n, uint32_t* perm) {
  std::iota(perm, perm + n, 0u);
  std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
}

oracle_scene* oracle_scene_create(void) { return new oracle_scene(); }
void oracle_scene_destroy(oracle_scene* s) { delete s; }

int oracle_scene_add_shader(oracle_scene* s, const oracle_shader_desc* d) {
  if (!s || !d || s->committed) return -1;
  s->shaders.push_back(*d);
  return (int)s->shaders.size() - 1;
}

int oracle_scene_add_mesh(oracle_scene* s, const float* pos4, uint32_t nverts, const uint32_t* idx3, uint32_t ntris,
                          const float* uv, int32_t shader) {
  if (!s || s->committed || (nverts && !pos4) || (ntris && !idx3)) return -1;
  for (uint32_t i = 0; i < 3 * ntris; ++i)
    if (idx3[i] >= nverts) return -1;
  oracle_scene::Mesh m;
  m.pos4.assign(pos4, pos4 + 4 * (size_t)nverts);
  m.idx.assign(idx3, idx3 + 3 * (size_t)ntris);
  if (uv) m.uv.assign(uv, uv + 2 * (size_t)nverts);
  m.shader = shader;
  s->objects.push_back(std::move(m));
  return (int)s->objects.size() - 1;
}

int oracle_scene_add_quad_light(oracle_scene* s, const oracle_quad_light_desc* d) {
  if (!s || !d || s->committed) return -1;
  s->light_descs.push_back(*d);
  return (int)s->light_descs.size() - 1;
}

void oracle_scene_set_environment(oracle_scene* s, const float rgb[3]) { s->env = v3(rgb[0], rgb[1], rgb[2]); }

int oracle_scene_add_image(oracle_scene* s, int32_t width, int32_t height, const float* rgb) {
  if (!s || s->committed || width <= 0 || height <= 0 || !rgb) return -1;
  oracle_scene::Image im;
  im.width = width;
  im.height = height;
  im.rgb.assign(rgb, rgb + 3 * (size_t)width * height);
  s->images.push_back(std::move(im));
  return (int)s->images.size() - 1;
}

int oracle_scene_add_texture(oracle_scene* s, const oracle_texture_desc* d) {
  if (!s || !d || s->committed) return -1;
  const int32_t n = (int32_t)s->textures.size();
  if (d->type == 1 && (d->image < 0 || d->image >= (int32_t)s->images.size())) return -1;
  if (d->type == 2 && (d->upper < 0 || d->upper >= n || d->lower < 0 || d->lower >= n || d->mask < 0 || d->mask >= n))
    return -1;
  if (d->type < 0 || d->type > 2) return -1;
  s->textures.push_back(*d);
  return n;
}

int oracle_scene_set_shader_texture(oracle_scene* s, int32_t shader, int32_t texture) {
  if (!s || s->committed || shader < 0 || shader >= (int32_t)s->shaders.size() || texture < -1 ||
      texture >= (int32_t)s->textures.size())
    return -1;
  if (s->shader_texture.size() < s->shaders.size()) s->shader_texture.resize(s->shaders.size(), -1);
  s->shader_texture[shader] = texture;
  return 0;
}

void oracle_texture_lookup(const float* rgb, int32_t width, int32_t height, float s, float t, float out[3]) {
  const V3 c = image_lookup(rgb, width, height, s, t);
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
}

int oracle_texture_colour(const oracle_scene* s, int32_t shader, uint32_t prim, float u, float v, float out[3]) {
  if (!s || !s->committed || shader < 0 || shader >= (int32_t)s->all_shaders.size() || prim >= s->tris.size())
    return -1;
  const oracle_shader_desc sh = shade_with_texture(s, shader, prim, u, v);
  out[0] = sh.colour[0];
  out[1] = sh.colour[1];
  out[2] = sh.colour[2];
  return 0;
}

int oracle_scene_commit(oracle_scene* s) {
  if (!s || s->committed) return -1;
  for (const auto& m : s->objects)
    if (m.shader < 0 || m.shader >= (int32_t)s->shaders.size()) return -1;
  s->all_shaders = s->shaders;
  s->geom_shader.clear();
  s->geom_mask.clear();
  s->geom_light.clear();
  s->geom_first_tri.clear();
  s->tris.clear();
  s->tri_uv.clear();
  s->tri_geom.clear();
  // Pass 1: objects (Pathtracer.cpp:166-217)
  for (size_t g = 0; g < s->objects.size(); ++g) {
    const auto& m = s->objects[g];
    s->geom_first_tri.push_back((uint32_t)s->tris.size());
    s->geom_shader.push_back(m.shader);
    s->geom_mask.push_back(0xFFFFFFFFu);
    s->geom_light.push_back(-1);
    for (size_t i = 0; i < m.idx.size() / 3; ++i) {
      Tri t;
      const float* a = &m.pos4[4 * m.idx[3 * i + 0]];
      const float* b = &m.pos4[4 * m.idx[3 * i + 1]];
      const float* c = &m.pos4[4 * m.idx[3 * i + 2]];
      t.v0 = v3(a[0], a[1], a[2]);
      t.v1 = v3(b[0], b[1], b[2]);
      t.v2 = v3(c[0], c[1], c[2]);
      s->tris.push_back(t);
      s->tri_geom.push_back((int32_t)g);
      for (int k = 0; k < 3; ++k) {
        const uint32_t vi = m.idx[3 * i + k];
        s->tri_uv.push_back(m.uv.empty() ? 0.f : m.uv[2 * vi]);
        s->tri_uv.push_back(m.uv.empty() ? 0.f : m.uv[2 * vi + 1]);
      }
    }
  }
  // Pass 2: lights, each with a placeholder object, a Null shader and mask 0xF0000000
  // (Pathtracer.cpp:219-272; quad geometry QuadLight.cpp:7-38)
  const float quad[6][3] = {{0.f, -0.5f, -0.5f}, {0.f, 0.5f, -0.5f}, {0.f, 0.5f, 0.5f},
                            {0.f, 0.5f, 0.5f},   {0.f, -0.5f, 0.5f}, {0.f, -0.5f, -0.5f}};
  s->lights.clear();
  for (size_t l = 0; l < s->light_descs.size(); ++l) {
    Light L = make_light(s->light_descs[l]);
    int32_t shader_id = (int32_t)s->all_shaders.size();
    oracle_shader_desc null_shader{};
    null_shader.type = 1;
    s->all_shaders.push_back(null_shader);
    int32_t g = (int32_t)s->geom_shader.size();
    s->geom_first_tri.push_back((uint32_t)s->tris.size());
    s->geom_shader.push_back(shader_id);
    s->geom_mask.push_back(0xF0000000u);
    s->geom_light.push_back((int32_t)l);
    for (int k = 0; k < 2; ++k) {
      Tri t;
      t.v0 = L.xf(v3(quad[3 * k][0], quad[3 * k][1], quad[3 * k][2]));
      t.v1 = L.xf(v3(quad[3 * k + 1][0], quad[3 * k + 1][1], quad[3 * k + 1][2]));
      t.v2 = L.xf(v3(quad[3 * k + 2][0], quad[3 * k + 2][1], quad[3 * k + 2][2]));
      s->tris.push_back(t);
      s->tri_geom.push_back(g);
      s->tri_uv.insert(s->tri_uv.end(), 6, 0.f);
    }
    s->lights.push_back(L);
  }
  build_bvh(s);
  s->committed = true;
  return 0;
}

uint32_t oracle_scene_num_triangles(const oracle_scene* s) { return (uint32_t)s->tris.size(); }

int oracle_scene_triangles(const oracle_scene* s, float* out9) {
  for (size_t i = 0; i < s->tris.size(); ++i) {
    const Tri& t = s->tris[i];
    float* o = out9 + 9 * i;
    o[0] = t.v0.x; o[1] = t.v0.y; o[2] = t.v0.z;
    o[3] = t.v1.x; o[4] = t.v1.y; o[5] = t.v1.z;
    o[6] = t.v2.x; o[7] = t.v2.y; o[8] = t.v2.z;
  }
  return 0;
}

int oracle_intersect(const oracle_scene* s, const float* rays, size_t n, uint32_t mask, uint32_t* hits) {
  if (!s || !s->committed) return -1;
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < (long long)n; ++i) {
    const float* r = rays + 8 * i;
    float t = 0.f, u = 0.f, v = 0.f;
    uint32_t prim = 0xFFFFFFFFu;
    bool h = trace(s, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7], mask, false, &t, &u, &v, &prim);
    uint32_t* o = hits + 4 * i;
    if (!h) {
      float nan = NAN;
      std::memcpy(&o[0], &r[7], 4);
      std::memcpy(&o[1], &nan, 4);
      std::memcpy(&o[2], &nan, 4);
      o[3] = 0xFFFFFFFFu;
    } else {
      std::memcpy(&o[0], &t, 4);
      std::memcpy(&o[1], &u, 4);
      std::memcpy(&o[2], &v, 4);
      o[3] = prim;
    }
  }
  return 0;
}

int oracle_intersect_brute(const oracle_scene* s, const float* rays, size_t n, uint32_t mask, uint32_t* hits) {
  if (!s || !s->committed) return -1;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long i = 0; i < (long long)n; ++i) {
    const float* r = rays + 8 * i;
    double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    double best = r[7], bu = 0, bv = 0;
    uint32_t bp = 0xFFFFFFFFu;
    for (size_t k = 0; k < s->tris.size(); ++k) {
      if ((s->geom_mask[s->tri_geom[k]] & mask) == 0) continue;
      const Tri& t = s->tris[k];
      // Solve o + t d = v0 + u (v1 - v0) + v (v2 - v0) by Cramer's rule in double.
      double a[3] = {t.v1.x - (double)t.v0.x, t.v1.y - (double)t.v0.y, t.v1.z - (double)t.v0.z};
      double b[3] = {t.v2.x - (double)t.v0.x, t.v2.y - (double)t.v0.y, t.v2.z - (double)t.v0.z};
      double c[3] = {o[0] - t.v0.x, o[1] - t.v0.y, o[2] - t.v0.z};
      // columns: a, b, -d ; unknowns u, v, t
      double m[3][3] = {{a[0], b[0], -d[0]}, {a[1], b[1], -d[1]}, {a[2], b[2], -d[2]}};
      auto det3 = [](double q[3][3]) {
        return q[0][0] * (q[1][1] * q[2][2] - q[1][2] * q[2][1]) - q[0][1] * (q[1][0] * q[2][2] - q[1][2] * q[2][0]) +
               q[0][2] * (q[1][0] * q[2][1] - q[1][1] * q[2][0]);
      };
      double D = det3(m);
      if (D == 0.0) continue;
      double mu[3][3], mv[3][3], mt[3][3];
      std::memcpy(mu, m, sizeof m);
      std::memcpy(mv, m, sizeof m);
      std::memcpy(mt, m, sizeof m);
      for (int row = 0; row < 3; ++row) {
        mu[row][0] = c[row];
        mv[row][1] = c[row];
        mt[row][2] = c[row];
      }
      double u = det3(mu) / D, v = det3(mv) / D, tt = det3(mt) / D;
      if (u < 0 || v < 0 || u + v > 1) continue;
      if (!(tt > r[3] && tt < best)) continue;
      best = tt;
      bu = u;
      bv = v;
      bp = (uint32_t)k;
    }
    uint32_t* out = hits + 4 * i;
    float ft = (float)best, fu = (float)bu, fv = (float)bv;
    if (bp == 0xFFFFFFFFu) fu = fv = NAN;
    std::memcpy(&out[0], &ft, 4);
    std::memcpy(&out[1], &fu, 4);
    std::memcpy(&out[2], &fv, 4);
    out[3] = bp;
  }
  return 0;
}

int oracle_occluded(const oracle_scene* s, const float* rays, size_t n, uint32_t mask, uint32_t* flags) {
  if (!s || !s->committed) return -1;
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < (long long)n; ++i) {
    const float* r = rays + 8 * i;
    float t, u, v;
    uint32_t prim = 0xFFFFFFFFu;
    flags[i] = trace(s, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7], mask, true, &t, &u, &v, &prim) ? 1u : 0u;
  }
  return 0;
}

int oracle_raygen(const oracle_render_desc* d, uint32_t sid0, size_t n, float* rays, float* sample_xy) {
  if (!d) return -1;
  CameraXf cx = camera_xf(d->camera);
  for (size_t i = 0; i < n; ++i) {
    V3 o, dir;
    float sx, sy;
    camera_ray(d, cx, sid0 + (uint32_t)i, &o, &dir, &sx, &sy);
    float* r = rays + 8 * i;
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = 0.001f;
    r[4] = dir.x; r[5] = dir.y; r[6] = dir.z; r[7] = 100000.f;
    if (sample_xy) {
      sample_xy[2 * i] = sx;
      sample_xy[2 * i + 1] = sy;
    }
  }
  return 0;
}

void oracle_filter(int32_t filter, int32_t width, int32_t height, int32_t spp, const float* samples, float* pixels) {
#pragma omp parallel for schedule(static)
  for (int py = 0; py < height; ++py)
    for (int px = 0; px < width; ++px) {
      float o[3];
      filter_pixel(filter, width, height, spp, samples, px, py, o);
      float* p = pixels + 3 * ((size_t)py * width + px);
      p[0] = o[0];
      p[1] = o[1];
      p[2] = o[2];
    }
}

int oracle_render(const oracle_scene* s, const oracle_render_desc* d, float* pixels, float* samples, uint32_t sid_begin,
                  uint32_t sid_end, int threads, oracle_stats* stats) {
  if (!s || !s->committed || !d || d->width <= 0 || d->height <= 0) return -1;
  const int spp = spp_of(d);
  if (spp <= 0) return -1;
  const size_t total = (size_t)d->width * d->height * spp;
  if (total > 0xFFFFFFFFull) return -1;
  if (sid_end == 0 || sid_end > total) sid_end = (uint32_t)total;
#ifdef _OPENMP
  int prev_threads = omp_get_max_threads();
  if (threads > 0) omp_set_num_threads(threads);
#endif
  // Sample records are only kept when the caller wants them or the image is filtered.
  std::vector<float> own;
  float* smp = samples;
  if (!smp && pixels) {
    own.assign(total * 5, 0.f);
    smp = own.data();
  }
  if (smp) {
    for (size_t i = 0; i < (size_t)sid_begin * 5; ++i) smp[i] = 0.f;
    for (size_t i = (size_t)sid_end * 5; i < total * 5; ++i) smp[i] = 0.f;
  }
  CameraXf cx = camera_xf(d->camera);
  uint64_t n_cam = 0, n_closest = 0, n_shadow = 0;
  double checksum = 0.0;
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : n_cam, n_closest, n_shadow, checksum)
  for (long long sid = sid_begin; sid < (long long)sid_end; ++sid) {
    V3 org, dir;
    float sx, sy;
    camera_ray(d, cx, (uint32_t)sid, &org, &dir, &sx, &sy);
    ++n_cam;
    V3 L = trace_path(s, d, (uint32_t)sid, org, dir, &n_closest, &n_shadow);
    checksum += (double)L.x + L.y + L.z;
    if (smp) {
      float* o = smp + 5 * sid;
      o[0] = sx;
      o[1] = sy;
      o[2] = L.x;
      o[3] = L.y;
      o[4] = L.z;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  if (pixels) oracle_filter(d->filter, d->width, d->height, spp, smp, pixels);
#ifdef _OPENMP
  omp_set_num_threads(prev_threads);
#endif
  if (stats) {
    stats->camera_rays = n_cam;
    stats->closest_rays = n_closest;
    stats->shadow_rays = n_shadow;
    stats->seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  return 0;
}

// ---- the BSDFs alone, vectorised over direction pairs (for the lobes' property tests) ------
void oracle_ggx_d_n(float alpha, size_t n, const float* cos_h, float* d) {
  for (size_t i = 0; i < n; ++i) d[i] = ggx_d(alpha * alpha, cos_h[i]);
}

void oracle_ggx_g1_n(float alpha, size_t n, const float* cos_v, float* g) {
  for (size_t i = 0; i < n; ++i) g[i] = ggx_g1(alpha * alpha, cos_v[i]);
}

void oracle_bsdf_eval_n(const oracle_shader_desc* sh, size_t n, const float* in3, const float* out3,
                        const float* n3, float* f3, float* cos, float* pdf) {
  for (size_t i = 0; i < n; ++i) {
    V3 in = v3(in3[3 * i], in3[3 * i + 1], in3[3 * i + 2]);
    V3 out = v3(out3[3 * i], out3[3 * i + 1], out3[3 * i + 2]);
    V3 nn = v3(n3[3 * i], n3[3 * i + 1], n3[3 * i + 2]);
    Eval e = bsdf_eval(*sh, in, out, nn);
    f3[3 * i] = e.f.x;
    f3[3 * i + 1] = e.f.y;
    f3[3 * i + 2] = e.f.z;
    cos[i] = e.cos;
    pdf[i] = e.pdf;
  }
}

void oracle_bsdf_sample_n(const oracle_shader_desc* sh, size_t n, const float* u2, const float* out3,
                          const float* n3, float* in3, float* f3, float* cos, float* pdf, uint8_t* ok) {
  for (size_t i = 0; i < n; ++i) {
    V3 out = v3(out3[3 * i], out3[3 * i + 1], out3[3 * i + 2]);
    V3 nn = v3(n3[3 * i], n3[3 * i + 1], n3[3 * i + 2]);
    V3 in = v3(0.f, 0.f, 0.f);
    Eval e{v3(0.f, 0.f, 0.f), 0.f, 0.f};
    ok[i] = bsdf_sample(*sh, u2[2 * i], u2[2 * i + 1], out, nn, &in, &e) ? 1 : 0;
    in3[3 * i] = in.x;
    in3[3 * i + 1] = in.y;
    in3[3 * i + 2] = in.z;
    f3[3 * i] = e.f.x;
    f3[3 * i + 1] = e.f.y;
    f3[3 * i + 2] = e.f.z;
    cos[i] = e.cos;
    pdf[i] = e.pdf;
  }
}

}  // extern "C"

// traverse.h -- device-side BVH4 traversal of one ray (closest hit and any hit).
//
// Replaces Embree's rtcIntersect (src/core/RayIntersect.cpp:7-12) and rtcOccluded
// (src/core/Integrator.cpp:153-174). One lane per ray, wave64; each lane keeps the top of its
// traversal stack in an LDS ring and spills deeper entries to HBM (ShortStack below).
// Closest-hit traversal visits the children of a node nearest first and pushes the rest far
// to near.
//
// Arithmetic follows the CPU restatement exactly (no contraction): slabs are (lo - o) * (1/d)
// with zero direction components replaced by +-2^-80, and the triangle test is Moller-Trumbore
// in Embree 2's arrangement (e1 = v0 - v1, e2 = v2 - v0, Ng = e1 x e2; hit accepted when
// tmin < t < tmax). The closest hit is the smallest t, and at exactly equal t the smaller global
// triangle id, so the result does not depend on the tree or the visiting order.
#pragma once

#include <algorithm>
#include <type_traits>

#include "scene.h"

namespace swr {

constexpr uint32_t kMiss = 0xFFFFFFFFu;
// Boxes are accepted when entry <= exit * (1 + 2^-22): a margin for rounding in the slab
// arithmetic so that a triangle lying on its box face is never culled.
constexpr float kBoxSlack = 1.00000024f;

struct RayState {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
  float ix, iy, iz;
};

__device__ inline float safe_inv(float d) {
  const float tiny = 8.27180613e-25f;  // 2^-80
  return 1.f / (fabsf(d) > tiny ? d : copysignf(tiny, d));
}

__device__ inline RayState make_ray(float4 o, float4 d) {
  RayState r;
  // tmin and tmax canonicalised once here, so that the min/max of every box test need not.
  r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = __builtin_canonicalizef(o.w);
  r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = __builtin_canonicalizef(d.w);
  r.ix = safe_inv(d.x);
  r.iy = safe_inv(d.y);
  r.iz = safe_inv(d.z);
  return r;
}

// Slab entry distance of one child box, or +inf when the box is missed.
__device__ inline float slab(const RayState& r, float lx, float hx, float ly, float hy, float lz, float hz, float tfar) {
  const float tx0 = (lx - r.ox) * r.ix, tx1 = (hx - r.ox) * r.ix;
  const float ty0 = (ly - r.oy) * r.iy, ty1 = (hy - r.oy) * r.iy;
  const float tz0 = (lz - r.oz) * r.iz, tz1 = (hz - r.oz) * r.iz;
  const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tmin));
  const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tfar));
  return tn <= tf * kBoxSlack ? tn : __int_as_float(0x7F800000);
}

// Triangle test; returns true with t, u, v for a hit in (tmin, tmax).
__device__ inline bool tri_test(const RayState& r, float4 v0, float4 e1, float4 e2, float* t, float* u, float* v) {
  const float ngx = e1.y * e2.z - e1.z * e2.y;
  const float ngy = e1.z * e2.x - e1.x * e2.z;
  const float ngz = e1.x * e2.y - e1.y * e2.x;
  const float cx = v0.x - r.ox, cy = v0.y - r.oy, cz = v0.z - r.oz;
  const float rx = r.dy * cz - r.dz * cy;
  const float ry = r.dz * cx - r.dx * cz;
  const float rz = r.dx * cy - r.dy * cx;
  const float den = ngx * r.dx + ngy * r.dy + ngz * r.dz;
  const float aden = fabsf(den);
  const float sgn = den < 0.f ? -1.f : 1.f;
  const float U = (rx * e2.x + ry * e2.y + rz * e2.z) * sgn;
  const float V = (rx * e1.x + ry * e1.y + rz * e1.z) * sgn;
  if (!(den != 0.f && U >= 0.f && V >= 0.f && U + V <= aden)) return false;
  const float T = (ngx * cx + ngy * cy + ngz * cz) * sgn;
  if (!(T > aden * r.tmin && T < aden * r.tmax)) return false;
  const float inv = 1.f / aden;
  *t = T * inv;
  *u = U * inv;
  *v = V * inv;
  return true;
}

// The (u, v) that tri_test returns for a ray known to hit the triangle (same arithmetic, so the
// same bits): u weights vertex 1, v vertex 2.
__device__ inline void tri_uv(float4 o, float4 d, float4 v0, float4 e1, float4 e2, float* u, float* v) {
  const float ngx = e1.y * e2.z - e1.z * e2.y;
  const float ngy = e1.z * e2.x - e1.x * e2.z;
  const float ngz = e1.x * e2.y - e1.y * e2.x;
  const float cx = v0.x - o.x, cy = v0.y - o.y, cz = v0.z - o.z;
  const float rx = d.y * cz - d.z * cy;
  const float ry = d.z * cx - d.x * cz;
  const float rz = d.x * cy - d.y * cx;
  const float den = ngx * d.x + ngy * d.y + ngz * d.z;
  const float sgn = den < 0.f ? -1.f : 1.f;
  const float inv = 1.f / fabsf(den);
  *u = ((rx * e2.x + ry * e2.y + rz * e2.z) * sgn) * inv;
  *v = ((rx * e1.x + ry * e1.y + rz * e1.z) * sgn) * inv;
}

// Node and triangle pointers come out of the scene table in memory, so the compiler cannot tell
// their address space and would issue flat loads; flat loads also count against the LDS wait
// counter, so every stack access would wait for the node fetches in flight. These casts make
// them global loads.
// (Loads go through clang's native 4-vectors: HIP's float4 class cannot be copied from a
// global-address-space reference.)
#define SWR_GLOBAL __attribute__((address_space(1)))
typedef float gf4 __attribute__((ext_vector_type(4)));
typedef uint32_t gu4 __attribute__((ext_vector_type(4)));
__device__ inline float4 gload(const SWR_GLOBAL gf4* p) {
  const gf4 v = *p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline const SWR_GLOBAL gf4* global_f4(const float4* p) { return (const SWR_GLOBAL gf4*)p; }
// A scene record (a multiple of 16 bytes, 16-byte aligned in its array) as global loads.
template <typename T>
__device__ inline T gload_rec(const T* p) {
  static_assert(sizeof(T) % 16 == 0, "record size must be a multiple of 16 bytes");
  T v;
  const SWR_GLOBAL gu4* src = (const SWR_GLOBAL gu4*)p;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 16); ++i) {
    const gu4 c = src[i];
    __builtin_memcpy(reinterpret_cast<char*>(&v) + 16 * i, &c, 16);
  }
  return v;
}
__device__ inline const SWR_GLOBAL gf4* global_nodes(const Node4* p) { return (const SWR_GLOBAL gf4*)p; }

// Byte offsets, within a node, of the near and far plane rows of each axis for one ray: the lo
// row when the ray's inverse direction on that axis is >= 0, else the hi row. (lo - o) * inv and
// (hi - o) * inv are ordered like lo and hi (or reversed for a negative inverse), since both steps
// are monotone, so the near/far distances equal the min/max of the two: the same box test with
// the per-axis min and max replaced by a per-ray choice of row.
struct PlaneRows {
  uint32_t nx, fx, ny, fy, nz, fz;
};
__device__ inline PlaneRows plane_rows(const RayState& r) {
  const uint32_t sx = r.ix < 0.f ? 16u : 0u, sy = r.iy < 0.f ? 16u : 0u, sz = r.iz < 0.f ? 16u : 0u;
  return PlaneRows{sx, 16u - sx, 32u + sy, 48u - sy, 64u + sz, 80u - sz};
}

// Slab entry distance of one child box from its near and far planes, or +inf when missed.
__device__ inline float slab_nf(const RayState& r, float nx, float fx, float ny, float fy, float nz, float fz,
                                float tfar) {
  const float tnx = (nx - r.ox) * r.ix, tfx = (fx - r.ox) * r.ix;
  const float tny = (ny - r.oy) * r.iy, tfy = (fy - r.oy) * r.iy;
  const float tnz = (nz - r.oz) * r.iz, tfz = (fz - r.oz) * r.iz;
  const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
  const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tfar));
  return tn <= tf * kBoxSlack ? tn : __int_as_float(0x7F800000);
}

// Entry distances of the four children of a node; +inf for a missed box or an unused slot.
// All four box tests run unconditionally and unused slots are masked afterwards, so the node's
// rows are fetched in one round of loads rather than behind per-slot branches. The node is the
// 128 bytes at base + node_off (rows lox, hix, loy, hiy, loz, hiz, child, unused); the offsets
// are 32-bit, so the loads take the form with a scalar base and a per-lane offset.
// kNearFar: each lane loads its ray's near and far rows (six per-lane offsets, fewer VALU
// operations, more registers); otherwise all six rows at fixed offsets and per-axis min/max.
template <bool kNearFar>
__device__ inline void node_entries(const RayState& r, const PlaneRows& pr, const SWR_GLOBAL char* base,
                                    uint32_t node_off, float tfar, uint4* ch, float* t0, float* t1, float* t2,
                                    float* t3) {
  auto row = [&](uint32_t off) { return gload((const SWR_GLOBAL gf4*)(base + (node_off + off))); };
  const gu4 c = *(const SWR_GLOBAL gu4*)(base + (node_off + 96u));
  *ch = make_uint4(c.x, c.y, c.z, c.w);
  const float inf = __int_as_float(0x7F800000);
  float s0, s1, s2, s3;
  if (kNearFar) {
    const float4 nx = row(pr.nx), fx = row(pr.fx), ny = row(pr.ny), fy = row(pr.fy), nz = row(pr.nz),
                 fz = row(pr.fz);
    s0 = slab_nf(r, nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, tfar);
    s1 = slab_nf(r, nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, tfar);
    s2 = slab_nf(r, nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, tfar);
    s3 = slab_nf(r, nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, tfar);
  } else {
    const float4 lx = row(0), hx = row(16), ly = row(32), hy = row(48), lz = row(64), hz = row(80);
    s0 = slab(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, tfar);
    s1 = slab(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, tfar);
    s2 = slab(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, tfar);
    s3 = slab(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, tfar);
  }
  *t0 = ch->x == kNoChild ? inf : s0;
  *t1 = ch->y == kNoChild ? inf : s1;
  *t2 = ch->z == kNoChild ? inf : s2;
  *t3 = ch->w == kNoChild ? inf : s3;
}

// Entry distances of the four children of a quantised node (NodeQ, 64 bytes at base + node_off,
// four 16-byte loads). A plane's distance is fma(q, s / d, (p - o) / d), which rounds outside the
// exact (p + q s - o) / d by a few ulps of the terms; the planes lie at least one quantum outside
// the child boxes, and the exit distance gets a slack of 2^-21, so no box that the ray meets is
// culled. Near and far planes are picked per axis by the sign of the direction, as plane_rows
// does for full nodes.
// kOct >= 0: every ray of the wave has the direction octant kOct (bit 0: x < 0, bit 1: y < 0,
// bit 2: z < 0), so the near and far planes are fixed at compile time; -1: chosen per lane.
constexpr float kBoxSlackQ = 1.00000048f;
__device__ inline float qplane(uint32_t w, int c, float a, float b) {
  return fmaf((float)((w >> (8 * c)) & 0xFFu), a, b);
}
template <int kOct = -1>
__device__ inline void node_entries_q(const RayState& r, const SWR_GLOBAL char* base, uint32_t node_off, float tfar,
                                      uint4* ch, float* t0, float* t1, float* t2, float* t3) {
  const gf4 h = *(const SWR_GLOBAL gf4*)(base + node_off);
  const gu4 qa = *(const SWR_GLOBAL gu4*)(base + (node_off + 16u));
  const gu4 qb = *(const SWR_GLOBAL gu4*)(base + (node_off + 32u));
  const gu4 c = *(const SWR_GLOBAL gu4*)(base + (node_off + 48u));
  *ch = make_uint4(c.x, c.y, c.z, c.w);
  const float ax = __uint_as_float(qb.z) * r.ix, ay = __uint_as_float(qb.w) * r.iy, az = h.w * r.iz;
  const float bx = (h.x - r.ox) * r.ix, by = (h.y - r.oy) * r.iy, bz = (h.z - r.oz) * r.iz;
  const bool sx = kOct >= 0 ? (kOct & 1) != 0 : r.ix < 0.f;
  const bool sy = kOct >= 0 ? (kOct & 2) != 0 : r.iy < 0.f;
  const bool sz = kOct >= 0 ? (kOct & 4) != 0 : r.iz < 0.f;
  const uint32_t nx = sx ? qa.y : qa.x, fx = sx ? qa.x : qa.y;
  const uint32_t ny = sy ? qa.w : qa.z, fy = sy ? qa.z : qa.w;
  const uint32_t nz = sz ? qb.y : qb.x, fz = sz ? qb.x : qb.y;
  const float inf = __int_as_float(0x7F800000);
  float s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float tn = fmaxf(fmaxf(qplane(nx, k, ax, bx), qplane(ny, k, ay, by)), fmaxf(qplane(nz, k, az, bz), r.tmin));
    const float tf = fminf(fminf(qplane(fx, k, ax, bx), qplane(fy, k, ay, by)), fminf(qplane(fz, k, az, bz), tfar));
    s[k] = tn <= tf * kBoxSlackQ ? tn : inf;
  }
  // Unused slots need no test: their empty interval is missed, or the null leaf is entered.
  *t0 = s[0];
  *t1 = s[1];
  *t2 = s[2];
  *t3 = s[3];
}

__device__ inline void cswap(uint32_t& a, uint32_t& b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}

struct HitResult {
  float t, u, v;
  uint32_t id;    // global triangle id or kMiss
  uint32_t leaf;  // position of that triangle in the BVH-order leaf array
};

// Traversal stack of one lane: the top kShort entries live in an LDS ring (entry k at
// lds[(k % kShort) * kBlock], so a wave's accesses hit 64 consecutive words); entries pushed
// deeper than that spill, oldest first, to the lane's column in HBM and come back on pop.
// A small ring keeps occupancy high; the spill makes any BVH depth correct.
template <int kBlock, int kShort>
struct ShortStack {
  static_assert((kShort & (kShort - 1)) == 0, "ring size must be a power of two");
  uint32_t* lds;    // this lane's ring column (stride kBlock)
  uint32_t* spill;  // this lane's spill column (entries [0, base))
  int sp, base;
  __device__ ShortStack(uint32_t* l, uint32_t* s) : lds(l), spill(s), sp(0), base(0) {}
  __device__ void push(uint32_t v) {
    if (sp - base == kShort) {
      spill[base] = lds[(base & (kShort - 1)) * kBlock];
      ++base;
    }
    lds[(sp & (kShort - 1)) * kBlock] = v;
    ++sp;
  }
  // Makes room for n unchecked pushes: spills the oldest ring entries while the ring would
  // overflow, exactly the spills n checked pushes would make.
  __device__ void reserve(int n) {
    while (sp - base + n > kShort) {
      spill[base] = lds[(base & (kShort - 1)) * kBlock];
      ++base;
    }
  }
  __device__ void push_reserved(uint32_t v) {
    lds[(sp & (kShort - 1)) * kBlock] = v;
    ++sp;
  }
  __device__ bool empty() const { return sp == 0; }
  __device__ uint32_t pop() {
    --sp;
    if (sp < base) {
      base = sp;
      return spill[sp];
    }
    return lds[(sp & (kShort - 1)) * kBlock];
  }
};

// Tests the triangles of leaf code `node` against the ray and updates best (the closest hit, at
// equal t the smaller triangle id). kAny: returns true at the first hit (best holds it).
// kPairTris: two triangles per step, loaded together; with an odd count the second load repeats
// the first triangle's address and its test is skipped. Tests run in leaf order, so the first hit
// of an any-hit query and the tie order of a closest-hit query are those of one at a time.
template <bool kAny, bool kPairTris>
__device__ inline bool leaf_tests(const RayState& r, const SWR_GLOBAL gf4* gtris, uint32_t node, uint32_t mask,
                                  HitResult* best) {
  const uint32_t first = leaf_first(node), count = leaf_count(node);
  for (uint32_t k = 0; k < count; k += kPairTris ? 2 : 1) {
    const bool two = kPairTris && k + 1 < count;
    const SWR_GLOBAL gf4* ta = gtris + 3 * (size_t)(first + k);
    const SWR_GLOBAL gf4* tb = two ? ta + 3 : ta;
    const float4 va = gload(ta), ea = gload(ta + 1), fa = gload(ta + 2);
    float4 vb = va, eb = ea, fb = fa;
    if (kPairTris) vb = gload(tb), eb = gload(tb + 1), fb = gload(tb + 2);
#pragma unroll
    for (int s = 0; s < (kPairTris ? 2 : 1); ++s) {
      if (s == 1 && !two) break;
      const float4 v0 = s ? vb : va, e1 = s ? eb : ea, e2 = s ? fb : fa;
      float t, u, v;
      // The mask is checked after the test, so that the triangle's rows load together.
      if (tri_test(r, v0, e1, e2, &t, &u, &v) && (__float_as_uint(e1.w) & mask) != 0u) {
        const uint32_t id = __float_as_uint(v0.w);
        if (kAny) {
          best->t = t;
          best->u = u;
          best->v = v;
          best->id = id;
          best->leaf = first + k + s;
          return true;
        }
        if (best->id == kMiss || t < best->t || (t == best->t && id < best->id)) {
          best->t = t;
          best->u = u;
          best->v = v;
          best->id = id;
          best->leaf = first + k + s;
        }
      }
    }
  }
  return false;
}

// Traverses one ray. kAny: returns on the first hit (id of that triangle), else the closest.
// lds_col: this lane's LDS ring column (kShort entries, stride kBlock); spill_col: this lane's
// HBM spill column (stack_need(depth) entries; unused when the ring is deep enough).
// kPairTris: leaf triangles are tested two per step with their loads together (fewer memory
// round trips, more registers: worth it where the kernel has registers to spare). kNearFar: box
// tests from the ray's near and far planes (see node_entries; same results, fewer operations,
// more registers).
// Popped nodes are fetched and their children tested against the closest hit so far; keeping
// each entry's distance to skip far ones unfetched cost more in stack traffic than it saved.
// kQuant: nodes is the quantised node array (NodeQ) instead of the full one (Node4).
// kChildLds (closest hit only): child_col is this lane's column of four LDS words (stride
// kBlock). A node with children to push stores its four child codes there, and the pushes read
// them back by slot: an LDS read per push instead of a chain of selects over the four registers
// (the traversal is bound by vector-instruction issue; LDS has room).
template <int kBlock, int kShort, bool kAny, bool kPairTris = false, bool kNearFar = false, bool kQuant = false,
          bool kChildLds = false, int kOct = -1>
__device__ inline HitResult traverse(const void* __restrict__ nodes, const float4* __restrict__ leaf_tris,
                                     const RayState& r, uint32_t mask, uint32_t* lds_col, uint32_t* spill_col,
                                     uint32_t* child_col = nullptr) {
  HitResult best;
  best.t = r.tmax;
  best.u = best.v = __int_as_float(0x7FC00000);
  best.id = kMiss;
  best.leaf = kMiss;
  const float inf = __int_as_float(0x7F800000);
  ShortStack<kBlock, kShort> st(lds_col, spill_col);
  const SWR_GLOBAL char* gnodes = (const SWR_GLOBAL char*)nodes;
  const SWR_GLOBAL gf4* gtris = global_f4(leaf_tris);
  const PlaneRows pr = plane_rows(r);
  uint32_t node = 0;  // root
  for (;;) {
    if (!(node & kLeafBit)) {
      // Unused slots (kNoChild) are never entered, whatever their box.
      uint4 ch;
      float t0, t1, t2, t3;
      if (kQuant)
        node_entries_q<kOct>(r, gnodes, node * 64u, best.t, &ch, &t0, &t1, &t2, &t3);
      else
        node_entries<kNearFar>(r, pr, gnodes, node * 128u, best.t, &ch, &t0, &t1, &t2, &t3);
      if (kAny) {
        // Occlusion only asks whether anything is hit, so the order is free: descend into the
        // first child whose box is hit and push the others (no stack round trip for it).
        const bool h0 = t0 != inf, h1 = t1 != inf, h2 = t2 != inf, h3 = t3 != inf;
        if (h1 && h0) st.push(ch.y);
        if (h2 && (h0 || h1)) st.push(ch.z);
        if (h3 && (h0 || h1 || h2)) st.push(ch.w);
        if (h0 || h1 || h2 || h3) {
          node = h0 ? ch.x : (h1 ? ch.y : (h2 ? ch.z : ch.w));
          continue;
        }
      } else {
        // Entry distances are >= tmin >= 0, so their bit patterns order like the floats; the low
        // two bits carry the child slot.
        uint32_t k0 = (__float_as_uint(t0) & ~3u) | 0u;
        uint32_t k1 = (__float_as_uint(t1) & ~3u) | 1u;
        uint32_t k2 = (__float_as_uint(t2) & ~3u) | 2u;
        uint32_t k3 = (__float_as_uint(t3) & ~3u) | 3u;
        cswap(k0, k1);
        cswap(k2, k3);
        cswap(k0, k2);
        cswap(k1, k3);
        cswap(k1, k2);
        const uint32_t miss_key = 0x7F800000u;
        const uint32_t cs[4] = {ch.x, ch.y, ch.z, ch.w};
        // Push hit children far to near except the nearest (keys ascend, so a hit k3 implies
        // hits k0..k2), then descend into the nearest.
        const bool h1 = (k1 & ~3u) < miss_key, h2 = (k2 & ~3u) < miss_key, h3 = (k3 & ~3u) < miss_key;
        if (h1) {
          st.reserve((int)h1 + (int)h2 + (int)h3);
          if (kChildLds) {
            child_col[0] = ch.x;
            child_col[kBlock] = ch.y;
            child_col[2 * kBlock] = ch.z;
            child_col[3 * kBlock] = ch.w;
            if (h3) st.push_reserved(child_col[(k3 & 3u) * kBlock]);
            if (h2) st.push_reserved(child_col[(k2 & 3u) * kBlock]);
            st.push_reserved(child_col[(k1 & 3u) * kBlock]);
          } else {
            if (h3) st.push_reserved(cs[k3 & 3u]);
            if (h2) st.push_reserved(cs[k2 & 3u]);
            st.push_reserved(cs[k1 & 3u]);
          }
        }
        if ((k0 & ~3u) < miss_key) {
          node = cs[k0 & 3u];
          continue;
        }
      }
    } else if (leaf_tests<kAny, kPairTris>(r, gtris, node, mask, &best)) {
      return best;
    }
    if (st.empty()) break;
    node = st.pop();
  }
  return best;
}

// traverse() with leaves postponed (Aila and Laine, "Understanding the efficiency of ray
// traversal on GPUs", HPG 2009, "while-while" with postponing): a lane that reaches a leaf keeps
// it aside and goes on through internal nodes until every lane of the wave still walking internal
// nodes holds a leaf; then the wave tests its held leaves together. Internal-node steps and leaf
// steps then run with more of the wave's lanes on the same branch. Same results as traverse():
// the closest hit (by t, then triangle id) does not depend on the visiting order, nor does
// whether anything is hit. Measured on the 1080p terrain frame: shadow rays (any hit) 2% faster,
// and 9% on the 256-job batch; closest-hit rays 2% slower (a held leaf cannot shorten the walk
// that continues past it), so the closest-hit kernels keep traverse().
template <int kBlock, int kShort, bool kAny, bool kPairTris = false, bool kNearFar = false, bool kQuant = false>
__device__ inline HitResult traverse_postponed(const void* __restrict__ nodes, const float4* __restrict__ leaf_tris,
                                               const RayState& r, uint32_t mask, uint32_t* lds_col,
                                               uint32_t* spill_col) {
  HitResult best;
  best.t = r.tmax;
  best.u = best.v = __int_as_float(0x7FC00000);
  best.id = kMiss;
  best.leaf = kMiss;
  const float inf = __int_as_float(0x7F800000);
  ShortStack<kBlock, kShort> st(lds_col, spill_col);
  const SWR_GLOBAL char* gnodes = (const SWR_GLOBAL char*)nodes;
  const SWR_GLOBAL gf4* gtris = global_f4(leaf_tris);
  const PlaneRows pr = plane_rows(r);
  constexpr uint32_t kDone = 0x7FFFFFFFu;  // the stack is empty (no node has this index)
  uint32_t node = 0;                       // root
  uint32_t leaf = 0;                       // the held leaf code, 0 for none
  for (;;) {
    while (node != kDone && !(node & kLeafBit)) {
      uint4 ch;
      float t0, t1, t2, t3;
      if (kQuant)
        node_entries_q(r, gnodes, node * 64u, best.t, &ch, &t0, &t1, &t2, &t3);
      else
        node_entries<kNearFar>(r, pr, gnodes, node * 128u, best.t, &ch, &t0, &t1, &t2, &t3);
      if (kAny) {
        const bool h0 = t0 != inf, h1 = t1 != inf, h2 = t2 != inf, h3 = t3 != inf;
        if (h1 && h0) st.push(ch.y);
        if (h2 && (h0 || h1)) st.push(ch.z);
        if (h3 && (h0 || h1 || h2)) st.push(ch.w);
        if (h0 || h1 || h2 || h3)
          node = h0 ? ch.x : (h1 ? ch.y : (h2 ? ch.z : ch.w));
        else
          node = st.empty() ? kDone : st.pop();
      } else {
        uint32_t k0 = (__float_as_uint(t0) & ~3u) | 0u;
        uint32_t k1 = (__float_as_uint(t1) & ~3u) | 1u;
        uint32_t k2 = (__float_as_uint(t2) & ~3u) | 2u;
        uint32_t k3 = (__float_as_uint(t3) & ~3u) | 3u;
        cswap(k0, k1);
        cswap(k2, k3);
        cswap(k0, k2);
        cswap(k1, k3);
        cswap(k1, k2);
        const uint32_t miss_key = 0x7F800000u;
        const uint32_t cs[4] = {ch.x, ch.y, ch.z, ch.w};
        const bool h1 = (k1 & ~3u) < miss_key, h2 = (k2 & ~3u) < miss_key, h3 = (k3 & ~3u) < miss_key;
        st.reserve((int)h1 + (int)h2 + (int)h3);
        if (h3) st.push_reserved(cs[k3 & 3u]);
        if (h2) st.push_reserved(cs[k2 & 3u]);
        if (h1) st.push_reserved(cs[k1 & 3u]);
        if ((k0 & ~3u) < miss_key)
          node = cs[k0 & 3u];
        else
          node = st.empty() ? kDone : st.pop();
      }
      // The first leaf reached is held and the walk goes on with the next stack entry.
      if ((node & kLeafBit) && node != kDone && leaf == 0u) {
        leaf = node;
        node = st.empty() ? kDone : st.pop();
      }
      // Every lane still in this loop holds a leaf: test them together.
      if (!__ballot(leaf == 0u)) break;
    }
    // The held leaf, then any leaf the walk stopped at.
    while (leaf != 0u) {
      if (leaf_tests<kAny, kPairTris>(r, gtris, leaf, mask, &best)) return best;
      leaf = 0u;
      if (node != kDone && (node & kLeafBit)) {
        leaf = node;
        node = st.empty() ? kDone : st.pop();
      }
    }
    if (node == kDone) break;
  }
  return best;
}

// Traversal state of one lane, advanced one node (internal or leaf) per step() so that a lane
// whose ray finishes can take another while the rest of its wave is still walking (see
// trace_chunk below). kAny: finish on the first hit, else the closest.
template <int kBlock, int kShort, bool kAny>
struct Walker {
  RayState r;
  HitResult best;
  uint32_t node;
  const Node4* nodes;
  const float4* leaf_tris;
  ShortStack<kBlock, kShort> st;

  __device__ Walker(uint32_t* lds_col, uint32_t* spill_col) : st(lds_col, spill_col) {}

  __device__ void start(const RayState& ray, const Node4* n, const float4* t) {
    r = ray;
    nodes = n;
    leaf_tris = t;
    best.t = ray.tmax;
    best.u = best.v = __int_as_float(0x7FC00000);
    best.id = kMiss;
    best.leaf = kMiss;
    node = 0;  // root
    st.sp = st.base = 0;
  }

  // One node; returns true when the ray is finished (result in best).
  __device__ bool step(uint32_t mask) {
    const float inf = __int_as_float(0x7F800000);
    if (!(node & kLeafBit)) {
      // Unused slots (kNoChild) are never entered, whatever their box.
      uint4 ch;
      float t0, t1, t2, t3;
      node_entries<true>(r, plane_rows(r), (const SWR_GLOBAL char*)global_nodes(nodes), node * 128u, best.t, &ch,
                         &t0, &t1, &t2, &t3);
      if (kAny) {
        // As in traverse: descend into the first hit child, push the others.
        const bool h0 = t0 != inf, h1 = t1 != inf, h2 = t2 != inf, h3 = t3 != inf;
        if (h1 && h0) st.push(ch.y);
        if (h2 && (h0 || h1)) st.push(ch.z);
        if (h3 && (h0 || h1 || h2)) st.push(ch.w);
        if (h0 || h1 || h2 || h3) {
          node = h0 ? ch.x : (h1 ? ch.y : (h2 ? ch.z : ch.w));
          return false;
        }
      } else {
        // Entry distances are >= tmin >= 0, so their bit patterns order like the floats; the low
        // two bits carry the child slot.
        uint32_t k0 = (__float_as_uint(t0) & ~3u) | 0u;
        uint32_t k1 = (__float_as_uint(t1) & ~3u) | 1u;
        uint32_t k2 = (__float_as_uint(t2) & ~3u) | 2u;
        uint32_t k3 = (__float_as_uint(t3) & ~3u) | 3u;
        cswap(k0, k1);
        cswap(k2, k3);
        cswap(k0, k2);
        cswap(k1, k3);
        cswap(k1, k2);
        const uint32_t miss_key = 0x7F800000u;
        const uint32_t cs[4] = {ch.x, ch.y, ch.z, ch.w};
        // Push hit children far to near except the nearest (keys ascend, so a hit k3 implies
        // hits k0..k2), then descend into the nearest.
        if ((k3 & ~3u) < miss_key) st.push(cs[k3 & 3u]);
        if ((k2 & ~3u) < miss_key) st.push(cs[k2 & 3u]);
        if ((k1 & ~3u) < miss_key) st.push(cs[k1 & 3u]);
        if ((k0 & ~3u) < miss_key) {
          node = cs[k0 & 3u];
          return false;
        }
      }
    } else {
      const uint32_t first = leaf_first(node), count = leaf_count(node);
      for (uint32_t k = 0; k < count; ++k) {
        const SWR_GLOBAL gf4* tp = global_f4(leaf_tris) + 3 * (size_t)(first + k);
        // The mask is checked after the test, so that the triangle's three rows load together.
        const float4 v0 = gload(tp), e1 = gload(tp + 1), e2 = gload(tp + 2);
        float t, u, v;
        if (tri_test(r, v0, e1, e2, &t, &u, &v) && (__float_as_uint(e1.w) & mask) != 0u) {
          const uint32_t id = __float_as_uint(v0.w);
          if (kAny) {
            best.t = t;
            best.u = u;
            best.v = v;
            best.id = id;
            best.leaf = first + k;
            return true;
          }
          if (best.id == kMiss || t < best.t || (t == best.t && id < best.id)) {
            best.t = t;
            best.u = u;
            best.v = v;
            best.id = id;
            best.leaf = first + k;
          }
        }
      }
    }
    if (st.empty()) return true;
    node = st.pop();
    return false;
  }
};

// Rays per wave of a traversal launch of n rays over `waves` waves: a multiple of 64, so that
// every wave but the last gets whole groups.
__host__ __device__ inline uint32_t chunk_rays(uint32_t n, uint32_t waves) {
  const uint32_t c = (n + waves - 1) / waves;
  return (c + 63u) & ~63u;
}

// Traces the rays [begin, end) of a wave with lane refill: whenever lanes are free, they take
// the next rays of the range in order (one ballot, no atomics), and every busy lane advances its
// ray one node per iteration. A wave then costs about the sum of its rays' work over 64 plus the
// longest single ray, instead of the longest ray of every group of 64 (traversal work per ray is
// heavy-tailed: a few grazing rays take several times the mean).
//   fetch(j, walker) -> bool: loads ray j and calls walker.start, or returns false to skip it;
//   finish(j, walker): stores the result of ray j.
template <int kBlock, int kShort, bool kAny, class Fetch, class Finish>
__device__ inline void trace_chunk(uint32_t begin, uint32_t end, uint32_t mask, uint32_t* lds_col,
                                   uint32_t* spill_col, Fetch&& fetch, Finish&& finish) {
  Walker<kBlock, kShort, kAny> wk(lds_col, spill_col);
  const uint32_t lane = __lane_id();
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t next = begin;  // the same on every lane
  uint32_t j = 0;
  bool busy = false;
  for (;;) {
    if (next < end) {
      const uint64_t idle = __ballot(!busy);
      if (idle) {
        if (!busy) {
          const uint32_t jj = next + (uint32_t)__popcll(idle & lt);
          if (jj < end) {
            j = jj;
            busy = fetch(j, wk);
          }
        }
        next += (uint32_t)__popcll(idle);
      }
    }
    if (!__ballot(busy)) {
      if (next >= end) break;
      continue;
    }
    if (busy && wk.step(mask)) {
      finish(j, wk);
      busy = false;
    }
  }
}

// Launch shape of the traversal kernels: workgroups of 128 lanes and an LDS ring of `ring`
// entries per lane (8, 16, 32 or 64).
//
// Queue kernels trace one ray per lane: their rays are sorted, and lanes of a wave walking the
// same nodes in step share each node fetch. Measured on the terrain frame, that beats lane refill
// (trace_chunk) at every chunk size: 40 ms per frame with one ray per lane, 44 / 53 / 65 ms with
// chunks of 64 / 128 / 256 rays per wave. Unsorted rays (the stand-alone queries) gain from
// refill instead: 8 M random rays take 2.46 ms with 64-ray chunks and 1.86 ms with 256.
constexpr int kTraceBlock = 128;
// Build switches (both on; off gives the variants they were measured against): the one-scene
// traversal reads the quantised nodes, and multi-scene loops trace scene-uniform tiles with the
// one-scene traversal. The per-lane multi-scene traversal keeps the full nodes: quantised, it
// needs more registers (79 against 76) and was slower.
#ifndef SWR_QUANT_NODES
#define SWR_QUANT_NODES 1
#endif
constexpr bool kQuantNodes = SWR_QUANT_NODES != 0;
// The one-scene closest-hit traversal reads pushed children back from LDS (traverse's child_col).
#ifndef SWR_LDS_CHILDREN
#define SWR_LDS_CHILDREN 1
#endif
constexpr bool kLdsChildren = SWR_LDS_CHILDREN != 0;
// The one-scene shadow-ray traversal holds leaves back (traverse_postponed).
#ifndef SWR_POSTPONE_SHADOW
#define SWR_POSTPONE_SHADOW 1
#endif
constexpr bool kPostponeShadow = SWR_POSTPONE_SHADOW != 0;
#ifndef SWR_SCENE_TILES
#define SWR_SCENE_TILES 1
#endif
constexpr bool kSceneTiles = SWR_SCENE_TILES != 0;
constexpr uint32_t kRefillChunk = 256;  // rays per wave of the stand-alone queries
inline uint32_t refill_waves(uint32_t n) { return std::max(1u, (n + kRefillChunk - 1) / kRefillChunk); }
inline uint32_t trace_blocks(uint32_t waves) { return (waves + kTraceBlock / 64 - 1) / (kTraceBlock / 64); }

// The ring size to use: the tuned one, or else the smallest power of two >= the stack the BVH
// can need, capped at 16 (deeper walks spill to HBM). Both autotuners pick 16 on every measured
// scene (the 2.1 M-triangle terrain, whose BVH4 is 12 levels deep, and the batch's 8 scenes);
// sizing the ring to the whole stack (64 entries there) cost occupancy and made the batched
// loop's closest-hit traversal 2.3x slower and its shadow rays 2.6x slower.
inline int choose_ring(int tuned, uint32_t max_depth) {
  if (tuned > 0) return tuned;
  const uint32_t need = stack_need(max_depth);
  int r = 8;
  while ((uint32_t)r < need && r < 16) r <<= 1;
  return r;
}

template <class F>
inline void with_ring(int ring, F&& f) {
  switch (ring) {
    case 8: f(std::integral_constant<int, 8>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    case 32: f(std::integral_constant<int, 32>()); break;
    case 64: f(std::integral_constant<int, 64>()); break;
    default: fail(SWR_E_INVALID, "traversal: stack ring must be 8, 16, 32 or 64 entries");
  }
}

}  // namespace swr

// bvh.hip -- GPU-built LBVH collapsed to a BVH4 (gfx950).
//
// Replaces Embree's acceleration build, rtcCommit (src/core/Pathtracer.cpp:274), over the
// triangle meshes the reference registers per object and light (:178-201, :242-265).
// Steps (Karras, "Maximizing parallelism in the construction of BVHs, octrees and k-d trees",
// HPG 2012; Karras and Aila's treelet restructuring, HPG 2013; a surface-area-heuristic
// BVH2 -> BVH4 collapse):
//   1. 30-bit Morton code of every triangle centroid over the centroid bounds;
//   2. stable radix sort of the codes (sort.hip), ties broken by triangle index;
//   3. one thread per internal node finds its key range and split -> binary radix tree;
//   4. leaves walk up with per-node arrival counters, the second arrival unions the boxes;
//   5. rounds of treelet restructuring (bottom-up the same way) lower the tree's SAH cost; the
//      triangles are then renumbered in the new tree's depth-first order;
//   6. step 4 again over that order, the second arrival also picking the node's least-cost BVH4
//      layout from its children's (dynamic programming);
//   7. level by level, each BVH4 node adopts the up to four BVH2 descendants those choices name;
//      subtrees the choices make leaves (<= leaf_max triangles) are contiguous runs of the
//      triangle order;
//   8. the leaf triangles are written in that order as (v0|id, e1|mask, e2|material).
#include <algorithm>
#include <cfloat>

#include "scene.h"

namespace swr {
namespace {

constexpr uint32_t kBvh2Leaf = 0x80000000u;

__device__ inline uint32_t spread10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ inline void tri_verts(const float4* pos, uint4 t, float3* a, float3* b, float3* c) {
  const float4 p0 = pos[t.x], p1 = pos[t.y], p2 = pos[t.z];
  *a = make_float3(p0.x, p0.y, p0.z);
  *b = make_float3(p1.x, p1.y, p1.z);
  *c = make_float3(p2.x, p2.y, p2.z);
}

__global__ void morton_kernel(const float4* __restrict__ pos, const uint4* __restrict__ tri, uint32_t n, float3 lo,
                              float3 inv_ext, uint32_t* __restrict__ codes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float3 a, b, c;
  tri_verts(pos, tri[i], &a, &b, &c);
  const float cx = (a.x + b.x + c.x) * (1.f / 3.f), cy = (a.y + b.y + c.y) * (1.f / 3.f),
              cz = (a.z + b.z + c.z) * (1.f / 3.f);
  const uint32_t qx = (uint32_t)fminf(fmaxf((cx - lo.x) * inv_ext.x * 1024.f, 0.f), 1023.f);
  const uint32_t qy = (uint32_t)fminf(fmaxf((cy - lo.y) * inv_ext.y * 1024.f, 0.f), 1023.f);
  const uint32_t qz = (uint32_t)fminf(fmaxf((cz - lo.z) * inv_ext.z * 1024.f, 0.f), 1023.f);
  codes[i] = (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
}

// Length of the common prefix of sorted keys i and j, with the index as tie breaker.
__device__ inline int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

// Internal node i: children (bit 31 = BVH2 leaf) and the sorted-primitive range it covers.
__global__ void karras_kernel(const uint32_t* __restrict__ keys, int n, uint2* __restrict__ children,
                              uint2* __restrict__ range, uint32_t* __restrict__ parent_internal,
                              uint32_t* __restrict__ parent_leaf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + min(d, 0);
  const int first = min(i, j), last = max(i, j);
  uint32_t left, right;
  if (first == gamma) {
    left = kBvh2Leaf | (uint32_t)gamma;
    parent_leaf[gamma] = (uint32_t)i;
  } else {
    left = (uint32_t)gamma;
    parent_internal[gamma] = (uint32_t)i;
  }
  if (last == gamma + 1) {
    right = kBvh2Leaf | (uint32_t)(gamma + 1);
    parent_leaf[gamma + 1] = (uint32_t)i;
  } else {
    right = (uint32_t)(gamma + 1);
    parent_internal[gamma + 1] = (uint32_t)i;
  }
  children[i] = make_uint2(left, right);
  range[i] = make_uint2((uint32_t)first, (uint32_t)last);
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Leaf boxes, then unions up the tree: the second child to arrive at a node computes its box.
// The thread that writes a node's box is the one that then arrives at its parent, so its arrival
// (an acquire-release atomic on the parent's counter) publishes that box to the parent's second
// arrival. A node whose leaf range lies inside this workgroup's 1024 leaves has both arrivals in
// the workgroup, so its counter is a workgroup-scope atomic; only nodes whose range crosses a
// workgroup boundary (the top of the tree) need device scope, whose release and acquire write
// back and invalidate the XCD's L2 (12.4 ms for 2.1 M triangles with device scope throughout,
// 4.4 ms with 256-leaf workgroups, 1.9 ms with 1024).
__device__ inline float half_area(float4 lo, float4 hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}

// Surface-area-heuristic collapse of the BVH2 into a BVH4 (Ylitie, Karras and Laine, "Efficient
// incoherent ray traversal on GPUs through compressed wide BVHs", HPG 2017, for four slots):
// C(x, i) is the least cost of the subtree of BVH2 node x placed in at most i child slots of a
// BVH4 node, a BVH4 node of box area A costing A * kSahNode and a leaf of t triangles A * t *
// kSahTri. With L, R the children of x and S(i) = min over j of C(L, j) + C(R, i - j):
//   C(x, 1) = min(leaf cost if x has at most leaf_max triangles, A * kSahNode + S(4)),
//   C(x, i) = min(C(x, 1), S(i))  for i = 2, 3, 4.
// A single triangle (a BVH2 leaf) costs A * kSahTri in any number of slots. The choices are
// packed into one word per node (BvhChoice) for the top-down collapse. Only the ratio of the two
// costs matters: on the 2.1 M-triangle terrain (leaves of <= 2 triangles) closest-hit traversal
// takes 25.8 ms per frame with a triangle costing 0.25 or 0.5 of a node and 26.3 ms at 1.0, against
// 26.45 ms for a greedy collapse that expands the child of largest area x triangle count first.
constexpr float kSahNode = 1.0f;
constexpr float kSahTri = 0.5f;
// bits 0-1: the slots of L when x is a BVH4 node (1..3; R takes 4 minus that); bit 2: x is a leaf
// in one slot; bits 4-5, 6-7, 8-9: for i = 2, 3, 4 slots, 0 = x takes one slot, else the slots
// of L (R takes the rest).
struct BvhChoice {
  static __host__ __device__ uint32_t node_split(uint32_t c) { return c & 3u; }
  static __host__ __device__ bool leaf(uint32_t c) { return (c >> 2) & 1u; }
  static __host__ __device__ uint32_t split(uint32_t c, uint32_t slots) { return (c >> (2u * slots)) & 3u; }
};

// The costs C(x, 1..4) and the choice word of internal node `node` from its children's costs.
__device__ inline void sah_choose(uint32_t node, uint2 ch, float4 lo, float4 hi, uint32_t size, uint32_t leaf_max,
                                  const float4* __restrict__ leaf_lo, const float4* __restrict__ leaf_hi,
                                  float4* __restrict__ cost, uint32_t* __restrict__ choice) {
  float cl[5], cr[5];
  auto child_costs = [&](uint32_t c, float* out) {
    if (c & kBvh2Leaf) {
      const float t = half_area(leaf_lo[c & ~kBvh2Leaf], leaf_hi[c & ~kBvh2Leaf]) * kSahTri;
      out[1] = out[2] = out[3] = out[4] = t;
    } else {
      const float4 k = cost[c];
      out[1] = k.x;
      out[2] = k.y;
      out[3] = k.z;
      out[4] = k.w;
    }
  };
  child_costs(ch.x, cl);
  child_costs(ch.y, cr);
  float split_cost[5];
  uint32_t split_left[5];
  for (uint32_t i = 2; i <= 4; ++i) {
    split_cost[i] = cl[1] + cr[i - 1];
    split_left[i] = 1;
    for (uint32_t j = 2; j < i; ++j) {
      const float c = cl[j] + cr[i - j];
      if (c < split_cost[i]) {
        split_cost[i] = c;
        split_left[i] = j;
      }
    }
  }
  const float area = half_area(lo, hi);
  const float as_node = area * kSahNode + split_cost[4];
  const float as_leaf = size <= leaf_max ? area * kSahTri * (float)size : __int_as_float(0x7F800000);
  uint32_t c = split_left[4];
  float one = as_node;
  if (as_leaf <= as_node) {
    one = as_leaf;
    c |= 4u;
  }
  float out[5];
  out[1] = one;
  for (uint32_t i = 2; i <= 4; ++i) {
    if (split_cost[i] < one) {
      out[i] = split_cost[i];
      c |= split_left[i] << (2u * i);
    } else {
      out[i] = one;
    }
  }
  cost[node] = make_float4(out[1], out[2], out[3], out[4]);
  choice[node] = c;
}

__global__ __launch_bounds__(1024) void refit_kernel(const float4* __restrict__ pos, const uint4* __restrict__ tri,
                                                     const uint32_t* __restrict__ perm, int n,
                                                     const uint2* __restrict__ children,
                                                     const uint2* __restrict__ range,
                                                     const uint32_t* __restrict__ parent_internal,
                                                     const uint32_t* __restrict__ parent_leaf,
                                                     uint32_t* __restrict__ visits, float4* __restrict__ leaf_lo,
                                                     float4* __restrict__ leaf_hi, float4* __restrict__ node_lo,
                                                     float4* __restrict__ node_hi, uint32_t leaf_max,
                                                     float4* __restrict__ cost, uint32_t* __restrict__ choice) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float3 a, b, c;
  tri_verts(pos, tri[perm[i]], &a, &b, &c);
  leaf_lo[i] = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)), 0.f);
  leaf_hi[i] = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)), 0.f);
  if (n == 1) return;
  const uint32_t wg_first = blockIdx.x * blockDim.x, wg_last = wg_first + blockDim.x - 1u;
  uint32_t node = parent_leaf[i];
  for (;;) {
    const uint2 rg = range[node];
    const bool local = rg.x >= wg_first && rg.y <= wg_last;
    const uint32_t prior = local ? __hip_atomic_fetch_add(&visits[node], 1u, __ATOMIC_ACQ_REL,
                                                          __HIP_MEMORY_SCOPE_WORKGROUP)
                                 : __hip_atomic_fetch_add(&visits[node], 1u, __ATOMIC_ACQ_REL,
                                                          __HIP_MEMORY_SCOPE_AGENT);
    if (prior == 0u) return;  // the sibling finishes this node
    const uint2 ch = children[node];
    float4 l0, h0, l1, h1;
    if (ch.x & kBvh2Leaf) {
      l0 = leaf_lo[ch.x & ~kBvh2Leaf];
      h0 = leaf_hi[ch.x & ~kBvh2Leaf];
    } else {
      l0 = node_lo[ch.x];
      h0 = node_hi[ch.x];
    }
    if (ch.y & kBvh2Leaf) {
      l1 = leaf_lo[ch.y & ~kBvh2Leaf];
      h1 = leaf_hi[ch.y & ~kBvh2Leaf];
    } else {
      l1 = node_lo[ch.y];
      h1 = node_hi[ch.y];
    }
    // w: the triangle count under the node (bits) and the node's binary-tree SAH cost, the sum of
    // the box areas of the internal nodes of its subtree (a triangle's box area is paid once in
    // any tree, so it is left out; leaf boxes carry 0 in w).
    const float4 lo = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z),
                                  __uint_as_float(rg.y - rg.x + 1u));
    const float4 hi = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.f);
    node_lo[node] = lo;
    node_hi[node] = make_float4(hi.x, hi.y, hi.z, half_area(lo, hi) + h0.w + h1.w);
    // the collapse's costs and choices, from the children's (published like their boxes)
    sah_choose(node, ch, lo, hi, rg.y - rg.x + 1u, leaf_max, leaf_lo, leaf_hi, cost, choice);
    if (node == 0) return;
    node = parent_internal[node];
  }
}

// Treelet restructuring of the binary tree (Karras and Aila, "Fast parallel construction of
// high-quality bounding volume hierarchies", HPG 2013). A bottom-up pass like the refit's: the
// second arrival at a node of at least `min_size` triangles grows a treelet under it (starting
// from the node's two children, the treelet leaf of largest box area is replaced by its children
// until there are kTreeletLeaves leaves) and rebuilds the treelet's internal nodes in the topology
// that minimises the sum of internal box areas, by dynamic programming over the 2^9 subsets of
// treelet leaves. The treelet's subtrees are left as they are; its internal nodes reuse their
// indices. The wave that finds such a node works the subsets lane-parallel; the node's own lane
// does every global load and store of the treelet, so that its acquiring arrival orders the
// loads and its next (releasing) arrival publishes the stores.
//
// Node records while restructuring: node_lo.w = triangle count (bits), node_hi.w = subtree cost,
// range = (lowest, highest) sorted leaf index under the node, which is no longer a contiguous run;
// treelet_layout_kernel renumbers the leaves afterwards.
constexpr int kTreeletLeaves = 9;
constexpr uint32_t kTreeletSets = 1u << kTreeletLeaves;
constexpr int kTreeletBlock = 1024;
// Rounds, each with its least treelet-root size in triangles. On the 2.1 M-triangle terrain
// (closest-hit traversal per 1080p frame; scene commit), with treelets of 7 leaves: none 25.8 ms,
// 61 ms; one round at 7 triangles 23.4 ms, 70 ms; three 22.0 ms, 82 ms; five at 3 triangles
// 21.6 ms, 92 ms; eight or twelve rounds no faster. Treelets of 8 leaves, five rounds at 3:
// 21.4 ms, 97 ms; of 9 leaves: 0.25% less frame time, 123 ms.
constexpr uint32_t kTreeletMinSize[] = {3u, 3u, 3u, 3u, 3u};
// Deepest binary tree the leaf renumbering walks (one walk per node and triangle, so a chain-like
// tree would cost depth x triangles). The radix tree is at most 61 levels deep (30-bit codes, ties
// by a 31-bit index); restructuring keeps it shallow (even splits win ties, and the terrain's
// BVH4 stays 12 levels deep). A tree deeper than this fails the build.
constexpr uint32_t kMaxBvh2Depth = 4096;

// The bit masks of 0..511 ordered by bit count, then by value: those of one count that lie below
// 2^k (the subsets of k treelet leaves) come first, count[k][s] of them.
struct SubsetTable {
  uint16_t mask[512];
  uint16_t start[11];
  uint8_t count[10][10];
};
constexpr SubsetTable make_subset_table() {
  SubsetTable t{};
  uint32_t n = 0;
  for (uint32_t s = 0; s <= 9; ++s) {
    t.start[s] = (uint16_t)n;
    for (uint32_t m = 0; m < 512; ++m) {
      uint32_t bits = 0;
      for (uint32_t v = m; v; v &= v - 1u) ++bits;
      if (bits == s) t.mask[n++] = (uint16_t)m;
    }
  }
  t.start[10] = (uint16_t)n;
  for (uint32_t k = 0; k <= 9; ++k)
    for (uint32_t m = 0; m < (1u << k); ++m) {
      uint32_t bits = 0;
      for (uint32_t v = m; v; v &= v - 1u) ++bits;
      ++t.count[k][bits];
    }
  return t;
}
__constant__ SubsetTable kSubsets = make_subset_table();
static_assert(kTreeletLeaves <= 9, "the subset table covers treelets of up to 9 leaves");

struct TreeletWave {
  float cost[kTreeletSets];   // least cost of a subset of treelet leaves joined under one node
  uint16_t part[kTreeletSets];  // the subset's best split: the part holding its lowest leaf
  float4 lo[kTreeletLeaves], hi[kTreeletLeaves];  // treelet leaves: box, lo.w = triangles, hi.w = cost
  uint2 rng[kTreeletLeaves];
  uint32_t ref[kTreeletLeaves];    // treelet leaves (bit 31 = triangle)
  uint32_t inner[kTreeletLeaves];  // the treelet's internal node indices, reused by the new topology
  uint32_t q_sub[kTreeletLeaves], q_idx[kTreeletLeaves];
  uint32_t k;
};

// Orders this wave's shared-memory writes before its later reads in other lanes.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline void treelet_load(TreeletWave& t, uint32_t j, uint32_t r, const uint2* range,
                                    const float4* leaf_lo, const float4* leaf_hi, const float4* node_lo,
                                    const float4* node_hi) {
  t.ref[j] = r;
  if (r & kBvh2Leaf) {
    const uint32_t l = r & ~kBvh2Leaf;
    const float4 lo = leaf_lo[l], hi = leaf_hi[l];
    t.lo[j] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(1u));
    t.hi[j] = make_float4(hi.x, hi.y, hi.z, 0.f);
    t.rng[j] = make_uint2(l, l);
  } else {
    t.lo[j] = node_lo[r];
    t.hi[j] = node_hi[r];
    t.rng[j] = range[r];
  }
}

// Restructures the treelet under `root`; called by the whole wave, `own` on the root's lane only.
__device__ inline void treelet_restructure(TreeletWave& t, uint32_t root, bool own, uint32_t lane,
                                           uint2* children, uint2* range, uint32_t* parent_internal,
                                           uint32_t* parent_leaf, const float4* leaf_lo, const float4* leaf_hi,
                                           float4* node_lo, float4* node_hi) {
  if (own) {
    const uint2 ch = children[root];
    t.inner[0] = root;
    treelet_load(t, 0, ch.x, range, leaf_lo, leaf_hi, node_lo, node_hi);
    treelet_load(t, 1, ch.y, range, leaf_lo, leaf_hi, node_lo, node_hi);
    uint32_t k = 2;
    while (k < (uint32_t)kTreeletLeaves) {
      int best = -1;
      float best_area = -1.f;
      uint32_t best_tris = 0;
      for (uint32_t j = 0; j < k; ++j) {
        if (t.ref[j] & kBvh2Leaf) continue;
        // the largest box; on equal boxes (coincident triangles) the most triangles, so that the
        // treelet's leaves, and the even splits that win ties over them, stay balanced
        const float a = half_area(t.lo[j], t.hi[j]);
        const uint32_t tris = __float_as_uint(t.lo[j].w);
        if (a > best_area || (a == best_area && tris > best_tris)) {
          best_area = a;
          best_tris = tris;
          best = (int)j;
        }
      }
      if (best < 0) break;
      const uint32_t x = t.ref[best];
      t.inner[k - 1] = x;
      const uint2 g = children[x];
      treelet_load(t, (uint32_t)best, g.x, range, leaf_lo, leaf_hi, node_lo, node_hi);
      treelet_load(t, k, g.y, range, leaf_lo, leaf_hi, node_lo, node_hi);
      ++k;
    }
    t.k = k;
  }
  wave_sync();
  const uint32_t k = t.k;
  if (k <= 2u) {  // two triangles: nothing to rearrange, the cost is unchanged
    wave_sync();
    return;
  }
  const uint32_t full = (1u << k) - 1u;
  for (uint32_t x = lane + 1u; x <= full; x += 64u)
    if ((x & (x - 1u)) == 0u) t.cost[x] = t.hi[__builtin_ctz(x)].w;
  // Subsets of one size at a time, each subset's splits shared by a group of g lanes (a power of
  // two, as many as the subsets of this size leave free), then the group's least cost by xor
  // shuffles. Equal costs (leaves with one box, as coincident triangles have) keep the most even
  // split, then the one of smaller bit mask: a total order, so the tree does not depend on the
  // grouping, and ties do not chain the leaves one below the other.
  for (uint32_t size = 2; size <= k; ++size) {
    wave_sync();
    const uint32_t ns = kSubsets.count[k][size], splits = (1u << (size - 1u)) - 1u;
    uint32_t g = 1;
    while (2u * g * ns <= 64u && 2u * g <= splits) g *= 2u;
    const uint32_t c = lane % g;
    const uint32_t r0 = c * splits / g, r1 = (c + 1u) * splits / g;
    for (uint32_t first = 0; first < ns; first += 64u / g) {
      const uint32_t i = first + lane / g;
      float best = FLT_MAX;
      uint32_t best_p = 0xFFFFu, best_u = 0xFFu;  // split, its unevenness |2 |p| - size|
      uint32_t x = 0;
      if (i < ns) {
        x = kSubsets.mask[kSubsets.start[size] + i];
        // splits (p, x \ p) with the lowest leaf of x in p; split r puts bit b of r on the b-th
        // remaining leaf of x, and the scan steps through the subsets of `rest` in that order
        const uint32_t low = x & (0u - x), rest = x ^ low;
        uint32_t q = 0;
        for (uint32_t m = rest, b = r0; m && b; m &= m - 1u, b >>= 1)
          if (b & 1u) q |= m & (0u - m);
        for (uint32_t r = r0; r < r1; ++r) {
          const uint32_t p = low | q;
          const float cp = t.cost[p] + t.cost[x ^ p];
          const uint32_t u = (uint32_t)abs(2 * __builtin_popcount(p) - (int)size);
          if (cp < best || (cp == best && u < best_u)) {  // r ascends with p: the smaller mask wins
            best = cp;
            best_p = p;
            best_u = u;
          }
          q = (q - rest) & rest;
        }
      }
      for (uint32_t o = g / 2u; o; o >>= 1) {
        const float ob = __shfl_xor(best, (int)o);
        const uint32_t op = (uint32_t)__shfl_xor((int)best_p, (int)o);
        const uint32_t ou = (uint32_t)__shfl_xor((int)best_u, (int)o);
        if (ob < best || (ob == best && (ou < best_u || (ou == best_u && op < best_p)))) {
          best = ob;
          best_p = op;
          best_u = ou;
        }
      }
      if (i < ns && c == 0u) {
        float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX, hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
        for (uint32_t m = x; m; m &= m - 1u) {
          const uint32_t j = __builtin_ctz(m);
          const float4 a = t.lo[j], b = t.hi[j];
          lx = fminf(lx, a.x); ly = fminf(ly, a.y); lz = fminf(lz, a.z);
          hx = fmaxf(hx, b.x); hy = fmaxf(hy, b.y); hz = fmaxf(hz, b.z);
        }
        const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
        t.cost[x] = dx * dy + dy * dz + dz * dx + best;
        t.part[x] = (uint16_t)best_p;
      }
    }
  }
  wave_sync();
  if (own) {
    // Rebuild breadth first from the full set; internal nodes take the treelet's old indices.
    uint32_t head = 0, tail = 1, next_inner = 1;
    t.q_sub[0] = full;
    t.q_idx[0] = root;
    auto lowest = [&](uint32_t x) {
      uint32_t r = 0xFFFFFFFFu;
      for (uint32_t m = x; m; m &= m - 1u) r = min(r, t.rng[__builtin_ctz(m)].x);
      return r;
    };
    while (head < tail) {
      const uint32_t x = t.q_sub[head], idx = t.q_idx[head];
      ++head;
      uint32_t a = t.part[x], b = x ^ a;
      if (lowest(a) > lowest(b)) {  // keep the lower leaves on the left (Morton order)
        const uint32_t w = a;
        a = b;
        b = w;
      }
      auto child = [&](uint32_t y) {
        uint32_t r;
        if ((y & (y - 1u)) == 0u) {
          r = t.ref[__builtin_ctz(y)];
        } else {
          r = t.inner[next_inner++];
          t.q_sub[tail] = y;
          t.q_idx[tail] = r;
          ++tail;
        }
        if (r & kBvh2Leaf) parent_leaf[r & ~kBvh2Leaf] = idx;
        else parent_internal[r] = idx;
        return r;
      };
      const uint32_t ra = child(a);
      const uint32_t rb = child(b);
      children[idx] = make_uint2(ra, rb);
      float4 lo = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.f), hi = make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.f);
      uint32_t tris = 0, rmin = 0xFFFFFFFFu, rmax = 0;
      for (uint32_t m = x; m; m &= m - 1u) {
        const uint32_t j = __builtin_ctz(m);
        const float4 l = t.lo[j], h = t.hi[j];
        lo = make_float4(fminf(lo.x, l.x), fminf(lo.y, l.y), fminf(lo.z, l.z), 0.f);
        hi = make_float4(fmaxf(hi.x, h.x), fmaxf(hi.y, h.y), fmaxf(hi.z, h.z), 0.f);
        tris += __float_as_uint(l.w);
        rmin = min(rmin, t.rng[j].x);
        rmax = max(rmax, t.rng[j].y);
      }
      node_lo[idx] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(tris));
      node_hi[idx] = make_float4(hi.x, hi.y, hi.z, t.cost[x]);
      range[idx] = make_uint2(rmin, rmax);
    }
  }
  wave_sync();
}

__global__ __launch_bounds__(kTreeletBlock) void treelet_kernel(uint2* children, uint2* range,
                                                                uint32_t* parent_internal, uint32_t* parent_leaf,
                                                                uint32_t* visits, const float4* leaf_lo,
                                                                const float4* leaf_hi, float4* node_lo,
                                                                float4* node_hi, int n, uint32_t min_size) {
  __shared__ TreeletWave s_wave[kTreeletBlock / 64];
  TreeletWave& t = s_wave[threadIdx.x / 64u];
  const uint32_t lane = threadIdx.x % 64u;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t wg_first = blockIdx.x * blockDim.x, wg_last = wg_first + blockDim.x - 1u;
  bool active = i < n;
  uint32_t node = active ? parent_leaf[i] : 0u;
  while (__ballot(active) != 0ull) {
    bool second = false;
    if (active) {
      // The arrival counters work as in the refit: a node whose leaves all lie in this workgroup's
      // range sees both arrivals here (the range bounds the leaves, contiguous or not).
      const uint2 rg = range[node];
      const bool local = rg.x >= wg_first && rg.y <= wg_last;
      const uint32_t prior = local ? __hip_atomic_fetch_add(&visits[node], 1u, __ATOMIC_ACQ_REL,
                                                            __HIP_MEMORY_SCOPE_WORKGROUP)
                                   : __hip_atomic_fetch_add(&visits[node], 1u, __ATOMIC_ACQ_REL,
                                                            __HIP_MEMORY_SCOPE_AGENT);
      if (prior == 0u) active = false;
      else second = true;
    }
    const bool big = second && __float_as_uint(node_lo[node].w) >= min_size;
    for (uint64_t m = __ballot(big); m; m &= m - 1ull) {
      const uint32_t leader = (uint32_t)__builtin_ctzll(m);
      const uint32_t root = (uint32_t)__shfl((int)node, (int)leader);
      treelet_restructure(t, root, lane == leader, lane, children, range, parent_internal, parent_leaf, leaf_lo,
                          leaf_hi, node_lo, node_hi);
    }
    if (second) {
      if (node == 0u) active = false;
      else node = parent_internal[node];
    }
  }
}

// Leaf order after restructuring: a node's leaves become the contiguous run of a depth-first
// (left child first) order. Element e < n is triangle leaf e, e >= n internal node e - n; each
// walks to the root adding the triangle counts of the left siblings on its path.
__global__ void treelet_layout_kernel(const uint2* __restrict__ children, const uint32_t* __restrict__ parent_internal,
                                      const uint32_t* __restrict__ parent_leaf, const float4* __restrict__ node_lo,
                                      uint32_t n, uint32_t* __restrict__ leaf_pos, uint2* __restrict__ range,
                                      uint32_t* __restrict__ error) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2u * n - 1u) return;
  const bool leaf = e < n;
  const uint32_t x = leaf ? e : e - n;
  if (!leaf && x == 0u) {
    range[0] = make_uint2(0u, n - 1u);
    return;
  }
  uint32_t cur = leaf ? (kBvh2Leaf | x) : x;
  uint32_t p = leaf ? parent_leaf[x] : parent_internal[x];
  uint32_t off = 0;
  for (uint32_t step = 0;; ++step) {
    if (step == kMaxBvh2Depth) {  // too deep (or not a tree): the host fails the build
      atomicOr(error, 2u);
      return;
    }
    const uint2 ch = children[p];
    if (ch.y == cur) off += (ch.x & kBvh2Leaf) ? 1u : __float_as_uint(node_lo[ch.x].w);
    if (p == 0u) break;
    cur = p;
    p = parent_internal[p];
  }
  if (leaf) leaf_pos[x] = off;
  else range[x] = make_uint2(off, off + __float_as_uint(node_lo[x].w) - 1u);
}

// Renumbers the triangle leaves by their new positions: the sorted-order permutation, the leaves'
// parents and the internal nodes' leaf children.
__global__ void treelet_renumber_kernel(uint2* __restrict__ children, const uint32_t* __restrict__ parent_leaf,
                                        const uint32_t* __restrict__ perm, const uint32_t* __restrict__ leaf_pos,
                                        uint32_t n, uint32_t* __restrict__ perm_out,
                                        uint32_t* __restrict__ parent_leaf_out) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) {
    const uint32_t p = leaf_pos[e];
    perm_out[p] = perm[e];
    parent_leaf_out[p] = parent_leaf[e];
  } else if (e < 2u * n - 1u) {
    const uint32_t x = e - n;
    uint2 ch = children[x];
    if (ch.x & kBvh2Leaf) ch.x = kBvh2Leaf | leaf_pos[ch.x & ~kBvh2Leaf];
    if (ch.y & kBvh2Leaf) ch.y = kBvh2Leaf | leaf_pos[ch.y & ~kBvh2Leaf];
    children[x] = ch;
  }
}

struct Bvh2View {
  const uint2* children;
  const uint2* range;
  const float4 *leaf_lo, *leaf_hi, *node_lo, *node_hi;
  const uint32_t* choice;  // BvhChoice words of the internal nodes
};

// One frontier entry (BVH4 node index, BVH2 internal node) becomes a filled BVH4 node.
__global__ void collapse_kernel(Bvh2View v, const uint2* __restrict__ frontier, const uint32_t* __restrict__ count_in,
                                uint2* __restrict__ next, uint32_t* __restrict__ count_next,
                                uint32_t* __restrict__ node_alloc, Node4* __restrict__ nodes, uint32_t depth,
                                uint32_t capacity, uint32_t* __restrict__ overflow) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *count_in) return;
  const uint2 item = frontier[i];
  // The node's four slots, filled as the surface-area-heuristic choices of the refit say: the
  // BVH2 node's children share the slots, and a descendant given more than one slot passes them
  // on to its own children (left subtree first). At most four entries are pending at a time.
  uint32_t ref[4], pend_x[4], pend_slots[4];
  int cnt = 0, np = 0;
  {
    const uint2 ch = v.children[item.y];
    const uint32_t j = BvhChoice::node_split(v.choice[item.y]);
    pend_x[np] = ch.y;
    pend_slots[np++] = 4u - j;
    pend_x[np] = ch.x;
    pend_slots[np++] = j;
  }
  while (np > 0) {
    --np;
    const uint32_t x = pend_x[np], slots = pend_slots[np];
    const uint32_t d = (x & kBvh2Leaf) || slots == 1u ? 0u : BvhChoice::split(v.choice[x], slots);
    if (d == 0u) {
      ref[cnt++] = x;
      continue;
    }
    const uint2 g = v.children[x];
    pend_x[np] = g.y;
    pend_slots[np++] = slots - d;
    pend_x[np] = g.x;
    pend_slots[np++] = d;
  }
  // A child in one slot is a leaf when it is one triangle or the choices made it one.
  auto is_leaf = [&](uint32_t r) { return (r & kBvh2Leaf) || BvhChoice::leaf(v.choice[r]); };
  Node4 nd;
  float lo[3][4], hi[3][4];
  uint32_t cc[4];
  // The internal children of a node get consecutive indices from an even one (allocations come in
  // pairs; the allocator starts at 2), so that quantised siblings (64 B) share a 128-byte line.
  uint32_t internal = 0;
  for (int c = 0; c < cnt; ++c)
    if (!is_leaf(ref[c])) ++internal;
  uint32_t next_idx = internal ? atomicAdd(node_alloc, (internal + 1u) & ~1u) : 0u;
  if (internal && next_idx + internal > capacity) {  // no room: the host fails the build
    atomicOr(overflow, 1u);
    return;
  }
  for (int c = 0; c < 4; ++c) {
    if (c >= cnt) {
      lo[0][c] = lo[1][c] = lo[2][c] = kEmptyBox;
      hi[0][c] = hi[1][c] = hi[2][c] = kEmptyBox;
      cc[c] = kNoChild;
      continue;
    }
    const uint32_t r = ref[c];
    float4 l, h;
    if (r & kBvh2Leaf) {
      l = v.leaf_lo[r & ~kBvh2Leaf];
      h = v.leaf_hi[r & ~kBvh2Leaf];
    } else {
      l = v.node_lo[r];
      h = v.node_hi[r];
    }
    lo[0][c] = l.x; lo[1][c] = l.y; lo[2][c] = l.z;
    hi[0][c] = h.x; hi[1][c] = h.y; hi[2][c] = h.z;
    if (r & kBvh2Leaf) {
      cc[c] = leaf_code(r & ~kBvh2Leaf, 1u);
    } else if (is_leaf(r)) {
      const uint2 g = v.range[r];
      cc[c] = leaf_code(g.x, g.y - g.x + 1u);
    } else {
      const uint32_t idx = next_idx++;
      cc[c] = idx;
      next[atomicAdd(count_next, 1u)] = make_uint2(idx, r);
    }
  }
  nd.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
  nd.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
  nd.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
  nd.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
  nd.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
  nd.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
  nd.child = make_uint4(cc[0], cc[1], cc[2], cc[3]);
  nd.meta = make_uint4((uint32_t)cnt, depth, 0u, 0u);
  nodes[item.x] = nd;
}

// Leaf triangles in sorted order: v0 (w = global id), e1 = v0 - v1 (w = geometry mask), e2 = v2 - v0
// (w = the geometry's material word: kLeafLight | light index for a light, else its shader index).
__global__ void leaf_tris_kernel(const float4* __restrict__ pos, const uint4* __restrict__ tri,
                                 const GeomRec* __restrict__ geoms, const uint32_t* __restrict__ perm, uint32_t n,
                                 float4* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t t = perm[j];
  const uint4 ix = tri[t];
  float3 a, b, c;
  tri_verts(pos, ix, &a, &b, &c);
  const GeomRec g = geoms[ix.w];
  const uint32_t material = g.light >= 0 ? kLeafLight | (uint32_t)g.light : (uint32_t)g.shader;
  out[3 * (size_t)j + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
  out[3 * (size_t)j + 1] = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, __uint_as_float(g.mask));
  out[3 * (size_t)j + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, __uint_as_float(material));
}

inline unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// One node per thread: the quantised copy (NodeQ) of a BVH4 node. Unused slots get an empty plane
// interval and the code of the null leaf (one degenerate triangle that no ray hits), so that the
// traversal needs no per-slot test: such a slot's box is missed, and if rounding ever lets a ray
// into it, the leaf test finds nothing.
__global__ void quantize_kernel(const Node4* __restrict__ nodes, uint32_t n, uint32_t null_leaf,
                                NodeQ* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Node4 nd = nodes[i];
  const bool valid[4] = {nd.child.x != kNoChild, nd.child.y != kNoChild, nd.child.z != kNoChild,
                         nd.child.w != kNoChild};
  const float lx[4] = {nd.lox.x, nd.lox.y, nd.lox.z, nd.lox.w}, hx[4] = {nd.hix.x, nd.hix.y, nd.hix.z, nd.hix.w};
  const float ly[4] = {nd.loy.x, nd.loy.y, nd.loy.z, nd.loy.w}, hy[4] = {nd.hiy.x, nd.hiy.y, nd.hiy.z, nd.hiy.w};
  const float lz[4] = {nd.loz.x, nd.loz.y, nd.loz.z, nd.loz.w}, hz[4] = {nd.hiz.x, nd.hiz.y, nd.hiz.z, nd.hiz.w};
  const QuantAxis ax = quantize_axis(lx, hx, valid), ay = quantize_axis(ly, hy, valid),
                  az = quantize_axis(lz, hz, valid);
  NodeQ q;
  q.px = ax.p;
  q.py = ay.p;
  q.pz = az.p;
  q.sx = __uint_as_float(ax.e << 23);
  q.sy = __uint_as_float(ay.e << 23);
  q.sz = __uint_as_float(az.e << 23);
  q.qlox = ax.qlo;
  q.qhix = ax.qhi;
  q.qloy = ay.qlo;
  q.qhiy = ay.qhi;
  q.qloz = az.qlo;
  q.qhiz = az.qhi;
  q.child = make_uint4(valid[0] ? nd.child.x : null_leaf, valid[1] ? nd.child.y : null_leaf,
                       valid[2] ? nd.child.z : null_leaf, valid[3] ? nd.child.w : null_leaf);
  out[i] = q;
}

void quantize_bvh(swr_context* ctx, BvhBuild* out) {
  const uint32_t n = std::max<uint32_t>(out->num_nodes, 1);
  ensure(ctx, out->qnodes, (size_t)n * sizeof(NodeQ));
  quantize_kernel<<<blocks(n), 256, 0, ctx->stream>>>(static_cast<const Node4*>(out->nodes.ptr), n,
                                                       null_leaf_code(out->num_tris),
                                                       static_cast<NodeQ*>(out->qnodes.ptr));
  SWR_HIP(hipGetLastError());
}

void build_bvh(swr_context* ctx, const float4* pos, const uint4* tri, const GeomRec* geoms, uint32_t n,
               const float lo[3], const float hi[3], int leaf_max, BvhBuild* out) {
  hipStream_t s = ctx->stream;
  leaf_max = std::max(1, std::min(leaf_max, 16));
  const size_t nn = std::max<uint32_t>(n, 1);
  ensure(ctx, out->nodes, nn * sizeof(Node4));
  // The leaf triangles plus the null leaf at index n (zeros: a degenerate triangle, mask 0).
  ensure(ctx, out->leaf_tris, (nn + 1) * 3 * sizeof(float4));
  SWR_HIP(hipMemsetAsync(static_cast<float4*>(out->leaf_tris.ptr) + 3 * (size_t)n, 0, 3 * sizeof(float4), s));
  out->num_tris = n;
  Node4* nodes = static_cast<Node4*>(out->nodes.ptr);

  if (n == 0) {  // empty scene: a root without children
    Node4 root;
    root.lox = root.hix = root.loy = root.hiy = root.loz = root.hiz =
        make_float4(kEmptyBox, kEmptyBox, kEmptyBox, kEmptyBox);
    root.child = make_uint4(kNoChild, kNoChild, kNoChild, kNoChild);
    root.meta = make_uint4(0, 0, 0, 0);
    SWR_HIP(hipMemcpyAsync(nodes, &root, sizeof root, hipMemcpyHostToDevice, s));
    SWR_HIP(hipStreamSynchronize(s));
    out->num_nodes = 1;
    out->max_depth = 0;
    return;
  }

  // Temporary buffers for the build, released at the end.
  DeviceBuffer codes, perm, alt_k, alt_v, children, range, par_i, par_l, visits, leaf_lo, leaf_hi, node_lo, node_hi,
      fr_a, fr_b, counters, sah_cost, sah_choice, leaf_pos, perm_r, par_l_r;
  auto cleanup = [&] {
    (void)hipStreamSynchronize(s);
    for (DeviceBuffer* b : {&codes, &perm, &alt_k, &alt_v, &children, &range, &par_i, &par_l, &visits, &leaf_lo,
                            &leaf_hi, &node_lo, &node_hi, &fr_a, &fr_b, &counters, &sah_cost, &sah_choice,
                            &leaf_pos, &perm_r, &par_l_r})
      release(ctx, *b);
  };
  try {
    ensure(ctx, codes, nn * 4);
    ensure(ctx, perm, nn * 4);
    ensure(ctx, alt_k, nn * 4);
    ensure(ctx, alt_v, nn * 4);
    // One scale for all three axes (the largest centroid extent): a thin axis then contributes
    // few Morton bits, so the hierarchy does not split flat geometry (a height field) along its
    // thin direction.
    float3 flo = make_float3(lo[0], lo[1], lo[2]);
    const float ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const float inv = ext > 0.f ? 1.f / ext : 0.f;
    float3 inv_ext = make_float3(inv, inv, inv);
    morton_kernel<<<blocks(n), 256, 0, s>>>(pos, tri, n, flo, inv_ext, static_cast<uint32_t*>(codes.ptr));
    uint32_t *sk = nullptr, *sv = nullptr;
    sort_pairs(ctx, static_cast<uint32_t*>(codes.ptr), static_cast<uint32_t*>(perm.ptr),
               static_cast<uint32_t*>(alt_k.ptr), static_cast<uint32_t*>(alt_v.ptr), n, 30, true, &sk, &sv, s);

    ensure(ctx, leaf_lo, nn * 16);
    ensure(ctx, leaf_hi, nn * 16);
    const size_t ni = n > 1 ? n - 1 : 1;
    ensure(ctx, children, ni * 8);
    ensure(ctx, range, ni * 8);
    ensure(ctx, par_i, ni * 4);
    ensure(ctx, par_l, nn * 4);
    ensure(ctx, visits, ni * 4);
    ensure(ctx, node_lo, ni * 16);
    ensure(ctx, node_hi, ni * 16);
    ensure(ctx, sah_cost, ni * 16);
    ensure(ctx, sah_choice, ni * 4);
    if (n > 1)
      karras_kernel<<<blocks(n - 1), 256, 0, s>>>(sk, (int)n, static_cast<uint2*>(children.ptr),
                                                  static_cast<uint2*>(range.ptr), static_cast<uint32_t*>(par_i.ptr),
                                                  static_cast<uint32_t*>(par_l.ptr));
    fill_u32_kernel<<<blocks(ni), 256, 0, s>>>(static_cast<uint32_t*>(visits.ptr), (uint32_t)ni, 0u);
    refit_kernel<<<blocks(n, 1024), 1024, 0, s>>>(pos, tri, sv, (int)n, static_cast<uint2*>(children.ptr),
                                           static_cast<uint2*>(range.ptr), static_cast<uint32_t*>(par_i.ptr),
                                           static_cast<uint32_t*>(par_l.ptr),
                                           static_cast<uint32_t*>(visits.ptr), static_cast<float4*>(leaf_lo.ptr),
                                           static_cast<float4*>(leaf_hi.ptr), static_cast<float4*>(node_lo.ptr),
                                           static_cast<float4*>(node_hi.ptr), (uint32_t)leaf_max,
                                           static_cast<float4*>(sah_cost.ptr),
                                           static_cast<uint32_t*>(sah_choice.ptr));
    SWR_HIP(hipGetLastError());

    if ((int)n <= leaf_max || n == 1) {
      leaf_tris_kernel<<<blocks(n), 256, 0, s>>>(pos, tri, geoms, sv, n, static_cast<float4*>(out->leaf_tris.ptr));
      SWR_HIP(hipGetLastError());
      // Everything fits one leaf: root with a single child spanning all triangles.
      float4 l, h;
      SWR_HIP(hipMemcpyAsync(&l, leaf_lo.ptr, 16, hipMemcpyDeviceToHost, s));
      SWR_HIP(hipMemcpyAsync(&h, leaf_hi.ptr, 16, hipMemcpyDeviceToHost, s));
      if (n > 1) {
        SWR_HIP(hipMemcpyAsync(&l, node_lo.ptr, 16, hipMemcpyDeviceToHost, s));
        SWR_HIP(hipMemcpyAsync(&h, node_hi.ptr, 16, hipMemcpyDeviceToHost, s));
      }
      SWR_HIP(hipStreamSynchronize(s));
      Node4 root;
      root.lox = make_float4(l.x, kEmptyBox, kEmptyBox, kEmptyBox);
      root.hix = make_float4(h.x, kEmptyBox, kEmptyBox, kEmptyBox);
      root.loy = make_float4(l.y, kEmptyBox, kEmptyBox, kEmptyBox);
      root.hiy = make_float4(h.y, kEmptyBox, kEmptyBox, kEmptyBox);
      root.loz = make_float4(l.z, kEmptyBox, kEmptyBox, kEmptyBox);
      root.hiz = make_float4(h.z, kEmptyBox, kEmptyBox, kEmptyBox);
      root.child = make_uint4(leaf_code(0, n), kNoChild, kNoChild, kNoChild);
      root.meta = make_uint4(1, 0, 0, 0);
      SWR_HIP(hipMemcpyAsync(nodes, &root, sizeof root, hipMemcpyHostToDevice, s));
      SWR_HIP(hipStreamSynchronize(s));
      out->num_nodes = 1;
      out->max_depth = 0;
      cleanup();
      return;
    }

    // Treelet restructuring rounds, then the leaves renumbered in the new depth-first order and
    // the boxes, counts and collapse choices recomputed over it.
    ensure(ctx, leaf_pos, nn * 4);
    ensure(ctx, perm_r, nn * 4);
    ensure(ctx, par_l_r, nn * 4);
    ensure(ctx, counters, 16);
    uint32_t* cnt = static_cast<uint32_t*>(counters.ptr);
    SWR_HIP(hipMemsetAsync(cnt + 3, 0, 4, s));
    for (uint32_t min_size : kTreeletMinSize) {
      fill_u32_kernel<<<blocks(ni), 256, 0, s>>>(static_cast<uint32_t*>(visits.ptr), (uint32_t)ni, 0u);
      treelet_kernel<<<blocks(n, kTreeletBlock), kTreeletBlock, 0, s>>>(
          static_cast<uint2*>(children.ptr), static_cast<uint2*>(range.ptr), static_cast<uint32_t*>(par_i.ptr),
          static_cast<uint32_t*>(par_l.ptr), static_cast<uint32_t*>(visits.ptr), static_cast<float4*>(leaf_lo.ptr),
          static_cast<float4*>(leaf_hi.ptr), static_cast<float4*>(node_lo.ptr), static_cast<float4*>(node_hi.ptr),
          (int)n, min_size);
    }
    treelet_layout_kernel<<<blocks(2 * (size_t)n - 1), 256, 0, s>>>(
        static_cast<uint2*>(children.ptr), static_cast<uint32_t*>(par_i.ptr), static_cast<uint32_t*>(par_l.ptr),
        static_cast<float4*>(node_lo.ptr), n, static_cast<uint32_t*>(leaf_pos.ptr), static_cast<uint2*>(range.ptr),
        cnt + 3);
    treelet_renumber_kernel<<<blocks(2 * (size_t)n - 1), 256, 0, s>>>(
        static_cast<uint2*>(children.ptr), static_cast<uint32_t*>(par_l.ptr), sv,
        static_cast<uint32_t*>(leaf_pos.ptr), n, static_cast<uint32_t*>(perm_r.ptr),
        static_cast<uint32_t*>(par_l_r.ptr));
    sv = static_cast<uint32_t*>(perm_r.ptr);
    fill_u32_kernel<<<blocks(ni), 256, 0, s>>>(static_cast<uint32_t*>(visits.ptr), (uint32_t)ni, 0u);
    refit_kernel<<<blocks(n, 1024), 1024, 0, s>>>(pos, tri, sv, (int)n, static_cast<uint2*>(children.ptr),
                                           static_cast<uint2*>(range.ptr), static_cast<uint32_t*>(par_i.ptr),
                                           static_cast<uint32_t*>(par_l_r.ptr),
                                           static_cast<uint32_t*>(visits.ptr), static_cast<float4*>(leaf_lo.ptr),
                                           static_cast<float4*>(leaf_hi.ptr), static_cast<float4*>(node_lo.ptr),
                                           static_cast<float4*>(node_hi.ptr), (uint32_t)leaf_max,
                                           static_cast<float4*>(sah_cost.ptr),
                                           static_cast<uint32_t*>(sah_choice.ptr));
    leaf_tris_kernel<<<blocks(n), 256, 0, s>>>(pos, tri, geoms, sv, n, static_cast<float4*>(out->leaf_tris.ptr));
    SWR_HIP(hipGetLastError());

    // Level-by-level collapse. counters: [0] node allocator, [1] frontier A count, [2] B count.
    ensure(ctx, fr_a, nn * 8);
    ensure(ctx, fr_b, nn * 8);
    uint32_t init[3] = {2u, 1u, 0u};  // node 0 is the root; node 1 is left unused
    uint2 root_item = make_uint2(0u, 0u);
    SWR_HIP(hipMemcpyAsync(counters.ptr, init, 12, hipMemcpyHostToDevice, s));
    SWR_HIP(hipMemcpyAsync(fr_a.ptr, &root_item, 8, hipMemcpyHostToDevice, s));
    Bvh2View view{static_cast<uint2*>(children.ptr),  static_cast<uint2*>(range.ptr),
                  static_cast<float4*>(leaf_lo.ptr),  static_cast<float4*>(leaf_hi.ptr),
                  static_cast<float4*>(node_lo.ptr),  static_cast<float4*>(node_hi.ptr),
                  static_cast<uint32_t*>(sah_choice.ptr)};
    uint2* cur = static_cast<uint2*>(fr_a.ptr);
    uint2* nxt = static_cast<uint2*>(fr_b.ptr);
    int cur_slot = 1, nxt_slot = 2;
    uint32_t frontier = 1, depth = 0;
    while (frontier) {
      SWR_HIP(hipMemsetAsync(cnt + nxt_slot, 0, 4, s));
      collapse_kernel<<<blocks(frontier), 256, 0, s>>>(view, cur, cnt + cur_slot, nxt, cnt + nxt_slot, cnt, nodes, depth,
                                                       (uint32_t)nn, cnt + 3);
      SWR_HIP(hipGetLastError());
      SWR_HIP(hipMemcpyAsync(&frontier, cnt + nxt_slot, 4, hipMemcpyDeviceToHost, s));
      SWR_HIP(hipStreamSynchronize(s));
      std::swap(cur, nxt);
      std::swap(cur_slot, nxt_slot);
      if (frontier) ++depth;
      if (depth > 256) fail(SWR_E_INVALID, "bvh: collapse did not terminate");
    }
    uint32_t num_nodes = 0, overflow = 0;
    SWR_HIP(hipMemcpyAsync(&num_nodes, cnt, 4, hipMemcpyDeviceToHost, s));
    SWR_HIP(hipMemcpyAsync(&overflow, cnt + 3, 4, hipMemcpyDeviceToHost, s));
    SWR_HIP(hipStreamSynchronize(s));
    if (overflow & 2u)
      fail(SWR_E_INVALID, "bvh: restructured binary tree deeper than " + std::to_string(kMaxBvh2Depth) + " levels");
    if (overflow) fail(SWR_E_INVALID, "bvh: node array too small for the collapsed hierarchy");
    out->num_nodes = num_nodes;
    out->max_depth = depth;
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace swr

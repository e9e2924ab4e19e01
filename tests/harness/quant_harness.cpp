// quant_harness.cpp -- CPU check of the BVH node quantisation (quantize_axis in scene.h) and of
// the quantised slab test the traversal runs (node_entries_q in traverse.h), restated on the host
// with the same float operations (fmaf, the 2^-21 exit slack).
//
// For random nodes of four child boxes (sizes 1e-4 .. 1e3, some flat, at coordinates up to
// +-100) and random rays (origins inside and far outside, some direction components zero):
//  * every quantised plane lies at least one quantum outside its child box, in exact arithmetic;
//  * the quantised test never rejects a box that the exact slab test (in double) accepts, and
//    never reports an entry distance later than the exact one.
// Prints the counts; exits non-zero on any failure.
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <random>
#include "scene.h"
using namespace swr;
static float bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static float safe_inv(float d) { const float tiny = 8.27180613e-25f; return 1.f / (std::fabs(d) > tiny ? d : std::copysign(tiny, d)); }
int main(int argc, char** argv) {
  const int iterations = argc > 1 ? std::atoi(argv[1]) : 200000;
  std::mt19937_64 rng(1);
  std::uniform_real_distribution<double> U(0, 1);
  long fails = 0, cons_fail = 0, tests = 0, loose = 0;
  for (int it = 0; it < iterations; ++it) {
    // a node: 4 child boxes around a centre at scale
    const double scale = std::pow(10.0, -4 + 7 * U(rng));
    const double cx = (U(rng) - 0.5) * 200, cy = (U(rng) - 0.5) * 200, cz = (U(rng) - 0.5) * 200;
    float lo[3][4], hi[3][4];
    bool valid[4];
    for (int c = 0; c < 4; ++c) {
      valid[c] = U(rng) < 0.85 || c == 0;
      const double cen[3] = {cx, cy, cz};
      for (int a = 0; a < 3; ++a) {
        double l = cen[a] + (U(rng) - 0.5) * scale, h = l + (U(rng) < 0.2 ? 0.0 : U(rng) * scale);
        lo[a][c] = (float)l; hi[a][c] = std::max((float)h, lo[a][c]);
      }
    }
    QuantAxis q[3];
    for (int a = 0; a < 3; ++a) q[a] = quantize_axis(lo[a], hi[a], valid);
    // conservativeness in exact arithmetic
    for (int a = 0; a < 3; ++a) {
      const double s = std::ldexp(1.0, (int)q[a].e - 127);
      for (int c = 0; c < 4; ++c) if (valid[c]) {
        const double ql = (q[a].qlo >> (8 * c)) & 255, qh = (q[a].qhi >> (8 * c)) & 255;
        if (!((double)q[a].p + ql * s <= (double)lo[a][c] - s && (double)q[a].p + qh * s >= (double)hi[a][c] + s)) ++cons_fail;
      }
    }
    // rays: origins near/inside the node, random directions incl. axis-parallel
    for (int rr = 0; rr < 8; ++rr) {
      float o[3], d[3], inv[3];
      for (int a = 0; a < 3; ++a) {
        const double cen = a == 0 ? cx : (a == 1 ? cy : cz);
        o[a] = (float)(cen + (U(rng) - 0.5) * scale * (rr < 4 ? 2.0 : 2000.0));
        d[a] = (float)(U(rng) - 0.5);
        if (U(rng) < 0.1) d[a] = 0.f;
        inv[a] = safe_inv(d[a]);
      }
      const float tmin = 0.001f * (float)scale, tfar = 1e5f;
      for (int c = 0; c < 4; ++c) if (valid[c]) {
        // exact slab (double) with the true box
        double tn = tmin, tf = tfar;
        bool exact_hit = true;
        for (int a = 0; a < 3; ++a) {
          if (d[a] == 0.f) { if (o[a] < lo[a][c] || o[a] > hi[a][c]) exact_hit = false; continue; }
          double t0 = ((double)lo[a][c] - o[a]) / d[a], t1 = ((double)hi[a][c] - o[a]) / d[a];
          tn = std::max(tn, std::min(t0, t1)); tf = std::min(tf, std::max(t0, t1));
        }
        if (!(tn <= tf)) exact_hit = false;
        // traversal emulation
        float tnq = tmin, tfq = tfar;
        for (int a = 0; a < 3; ++a) {
          const float A = bits(q[a].e << 23) * inv[a];
          const float B = (q[a].p - o[a]) * inv[a];
          const bool neg = inv[a] < 0.f;
          const uint32_t nw = neg ? q[a].qhi : q[a].qlo, fw = neg ? q[a].qlo : q[a].qhi;
          const float tnn = std::fma((float)((nw >> (8 * c)) & 255u), A, B);
          const float tff = std::fma((float)((fw >> (8 * c)) & 255u), A, B);
          tnq = std::max(tnq, tnn); tfq = std::min(tfq, tff);
        }
        const bool qhit = tnq <= tfq * 1.00000048f;
        ++tests;
        if (exact_hit && !qhit) { if (fails < 5) std::printf("miss: it %d c %d tn %.9g tf %.9g | q %.9g %.9g\n", it, c, tn, tf, tnq, tfq); ++fails; }
        if (!exact_hit && qhit) ++loose;
        if (exact_hit && qhit && tnq > tn * (1 + 1e-6) + 1e-30) { if (fails < 5) std::printf("entry late: %.9g > %.9g\n", tnq, tn); ++fails; }
      }
    }
  }
  std::printf("tests %ld fails %ld conservative-fails %ld loose %ld\n", tests, fails, cons_fail, loose);
  return fails || cons_fail;
}

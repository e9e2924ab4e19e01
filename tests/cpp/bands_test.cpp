// Property test of the column-band arithmetic (csrc/bands.h) on random frames and budgets:
// bands partition the columns, each band traces every column its filter reads and nothing
// outside the image, band sample ranges start at the band's first sample id, and a band fits
// the room it was planned for (or the planner says that not even one column fits).
#include <cstdio>
#include <random>

#include "bands.h"

int main() {
  std::mt19937_64 rng(7);
  for (int trial = 0; trial < 200000; ++trial) {
    const uint32_t W = 1 + rng() % 700, H = 1 + rng() % 300, spp = 1 + rng() % 70;
    const bool box = rng() & 1;
    const uint64_t per_col = (uint64_t)H * spp;
    const uint64_t room = rng() % (per_col * (W + 3) + 1);
    const uint64_t cols = swr::band_width_for(W, per_col, box, room);
    const uint32_t halo = swr::band_halo(box);
    if (cols == 0) {
      if ((uint64_t)(1 + 2 * halo) * per_col <= room && W * per_col > room) return std::printf("missed room %d\n", trial), 1;
      continue;
    }
    if (cols > W) return std::printf("too wide %d\n", trial), 1;
    const uint64_t max_band = swr::band_max_samples(W, per_col, box, cols);
    if (cols < W && max_band > room) return std::printf("band exceeds room %d\n", trial), 1;
    uint32_t covered = 0;
    for (uint32_t x0 = 0; x0 < W; x0 += (uint32_t)cols) {
      const uint32_t x1 = x0 + (uint32_t)cols < W ? x0 + (uint32_t)cols : W;
      if (x0 != covered) return std::printf("gap %d\n", trial), 1;
      covered = x1;
      const swr::BandLayout b = swr::band_layout(W, H, spp, box, x0, x1);
      if (b.band_x0 + b.band_cols > W) return std::printf("outside image %d\n", trial), 1;
      if (b.sid_base != b.band_x0 * H * spp) return std::printf("sid base %d\n", trial), 1;
      if ((uint64_t)b.band_cols * per_col > max_band) return std::printf("band larger than planned %d\n", trial), 1;
      // every column the filter reads for an output column lies in the band
      for (uint32_t x = x0; x < x1; ++x)
        for (int d = -(int)halo; d <= (int)halo; ++d) {
          const int q = (int)x + d;
          if (q < 0 || q >= (int)W) continue;
          if (q < (int)b.band_x0 || q >= (int)(b.band_x0 + b.band_cols)) return std::printf("missing column %d\n", trial), 1;
        }
      // and nothing beyond those columns
      if (b.band_x0 + halo < x0 || b.band_x0 + b.band_cols > x1 + halo) return std::printf("extra columns %d\n", trial), 1;
    }
    if (covered != W) return std::printf("not covered %d\n", trial), 1;
  }
  std::printf("ok\n");
  return 0;
}

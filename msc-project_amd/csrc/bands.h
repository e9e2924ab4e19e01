// bands.h -- column bands of a frame's samples (the out-of-core form of per-sample state).
//
// Sample ids are column-major (x * height * spp + y * spp + i, Camera.cpp:28), so the samples of
// a run of image columns are one contiguous id range, and the tent filter of a pixel reads the
// samples of its own and its two neighbouring columns (TentFilter.cpp:22-51; the box filter only
// its own). A band that writes columns [x0, x1) therefore traces the columns [x0 - 1, x1 + 1)
// clipped to the image (tent) or [x0, x1) (box). Pure functions, host and device.
#pragma once

#include <cstdint>

#ifndef __HIPCC__
#define SWR_HD
#else
#define SWR_HD __host__ __device__
#endif

namespace swr {

struct BandLayout {
  uint32_t band_x0, band_cols;  // columns whose samples the band traces
  uint32_t sid_base;            // first sample id of the band
};

// Neighbour columns a band needs on each side: 1 for the tent filter, 0 for the box filter.
SWR_HD inline uint32_t band_halo(bool box_filter) { return box_filter ? 0u : 1u; }

// Samples traced by the band that filters columns [x0, x1) of a width x height x spp frame.
SWR_HD inline BandLayout band_layout(uint32_t width, uint32_t height, uint32_t spp, bool box_filter, uint32_t x0,
                                     uint32_t x1) {
  const uint32_t halo = band_halo(box_filter);
  BandLayout b;
  b.band_x0 = x0 >= halo ? x0 - halo : 0u;
  const uint32_t end = x1 + halo < width ? x1 + halo : width;
  b.band_cols = end - b.band_x0;
  b.sid_base = b.band_x0 * height * spp;
  return b;
}

// Columns each band filters so that a band's samples fit `room` samples: the whole width when
// the frame fits, else the most columns whose band (with both neighbour columns) fits; 0 when
// not even one column fits.
SWR_HD inline uint64_t band_width_for(uint64_t width, uint64_t per_col, bool box_filter, uint64_t room) {
  if (width * per_col <= room) return width;
  const uint64_t halo = 2 * band_halo(box_filter);
  const uint64_t cols = per_col ? room / per_col : width;
  if (cols <= halo) return 0;
  return cols - halo < width ? cols - halo : width;
}

// Samples the largest band of that width traces (an interior band: its columns and both
// neighbour columns, clipped to the frame).
SWR_HD inline uint64_t band_max_samples(uint64_t width, uint64_t per_col, bool box_filter, uint64_t band_cols) {
  const uint64_t cols = band_cols + 2 * band_halo(box_filter);
  return (cols < width ? cols : width) * per_col;
}

}  // namespace swr

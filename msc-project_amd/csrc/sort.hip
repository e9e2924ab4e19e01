// sort.hip -- stable LSD radix sort of (u32 key, u32 value) pairs for gfx950.
//
// Replaces the reference's two sorts on the hot path: the recursive ray sort
// (src/core/RaySort.cpp:7-118, driven from Pathtracer.cpp:321-327) and the hit-point sort
// (tbb::parallel_sort with CompareHit, Pathtracer.cpp:335-339). Both become one stable
// radix sort of 32-bit keys; the permutation it returns is bit-exact against a stable sort
// of the same keys.
//
// Design (onesweep, Adinets & Merrill 2022, restated for CDNA4 wave64):
//   1. one read of the keys builds the 256-bin histograms of every 8-bit digit position;
//   2. one kernel per digit: each 512-lane workgroup takes a tile of 512 * kpt keys (ticketed in
//      launch order), ranks keys stably inside each wave, publishes its per-digit counts,
//      resolves its global offsets by decoupled look-back over earlier tiles, stages the tile in
//      LDS in digit order and writes runs of equal digits contiguously.
// A key's peers (the lanes of its wave holding the same digit) come from LDS: each lane ORs its
// lane bit into its wave's 64-bit word for its digit, reads the word back and clears it. That is
// about a quarter of the VALU work of matching the 8 digit bits by ballots, which made the pass
// ALU-bound. Values are loaded after the ranking, so their registers are not live across it.
// Look-back status words are 64-bit: [63:32] epoch tag of the pass, [31:30] flag
// (1 aggregate, 2 inclusive prefix), [29:0] count. Words from older passes carry an older
// epoch and read as "not ready", so the status buffer never needs clearing between passes.
// The epoch counter and the element count can live in device memory, so that a captured
// graph of the wavefront loop replays the sort with fresh tags and the current wave size.
// The reduce-then-scan alternative replaces the look-back by an upsweep of per-tile digit counts
// and a per-digit scan over tiles.
#include "swr_internal.h"

namespace swr {
namespace {

constexpr int kRadix = 256;
constexpr int kWaves = kSortBlock / 64;
constexpr int kHistBlock = 1024;
// The histogram kernel adds into kHistCopies copies of the 4 x 256 counts (workgroup i into copy
// i % kHistCopies, i.e. one copy per XCD under round-robin dispatch): adds from every CU to one
// word serialise, and spreading them over copies cuts that chain eightfold.
constexpr int kHistCopies = 8;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPre = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1;

// counters: [0..3] tile tickets of the passes; [kEpochWord] epoch of the last pass of the last
// sort, on its own cache line: every tile reads it while the tickets take atomics, and a read
// of a line under atomics waits behind them.
constexpr int kEpochWord = 64;
__global__ void sort_prepare_kernel(uint32_t* hist, uint32_t* counters, int passes) {
  const int t = threadIdx.x;
#pragma unroll
  for (int c = 0; c < kHistCopies * 4; ++c) hist[c * kRadix + t] = 0;
  if (t < 4) counters[t] = 0;
  if (t == 0) counters[kEpochWord] += (uint32_t)passes;  // pass p uses that - passes + 1 + p
}

// Adds key digits to an LDS histogram: one LDS atomic per lane, or one per wave when every
// active lane holds the same digit (skewed digits such as a few direction octants or a mostly-
// zero top byte would otherwise serialise 64 atomics on one bank). Called by whole waves.
__device__ inline void hist_add(uint32_t* h, uint32_t d, bool ok) {
  const uint64_t active = __ballot(ok);
  if (!active) return;
  const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);  // lane 0 (the whole wave is active)
  if ((active & 1ull) && __ballot(ok && d == d0) == active) {
    if ((threadIdx.x & 63) == 0) atomicAdd(&h[d0], (uint32_t)__popcll(active));
  } else if (ok) {
    atomicAdd(&h[d], 1u);
  }
}

// Histograms of all digit positions in one read of the keys: one 1024-lane workgroup per CU,
// grid-stride over 16-byte loads, four LDS copies (one per four waves); each workgroup adds its
// counts to its global copy once.
template <bool kVec4>
__global__ __launch_bounds__(kHistBlock) void sort_histogram_kernel(const uint32_t* __restrict__ keys, size_t n_host,
                                                                    const uint32_t* __restrict__ n_dev, int passes,
                                                                    uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_hist[4][4 * kRadix];
  const size_t n = n_dev ? (size_t)*n_dev : n_host;
  for (int i = threadIdx.x; i < 4 * 4 * kRadix; i += kHistBlock) (&s_hist[0][0])[i] = 0;
  __syncthreads();
  uint32_t* h = s_hist[threadIdx.x >> 8];
  const size_t n4 = kVec4 ? n / 4 : 0;  // keys not 16-byte aligned: all go through the scalar loop
  const uint4* k4 = reinterpret_cast<const uint4*>(keys);
  const size_t stride = (size_t)gridDim.x * kHistBlock;
  // Whole waves iterate together (the bound is rounded up per wave) so the ballots see every lane.
  const size_t first = (size_t)blockIdx.x * kHistBlock + threadIdx.x;
  const size_t wave_first = first - (threadIdx.x & 63);
  for (size_t wi = wave_first, i = first; wi < n4; wi += 2 * stride, i += 2 * stride) {
    const bool ok0 = i < n4, ok1 = i + stride < n4;
    const uint4 a = ok0 ? k4[i] : make_uint4(0u, 0u, 0u, 0u);
    const uint4 b = ok1 ? k4[i + stride] : make_uint4(0u, 0u, 0u, 0u);
    for (int p = 0; p < passes; ++p) {
      const int sh = 8 * p;
      uint32_t* hp = h + p * kRadix;
      hist_add(hp, (a.x >> sh) & 255u, ok0);
      hist_add(hp, (a.y >> sh) & 255u, ok0);
      hist_add(hp, (a.z >> sh) & 255u, ok0);
      hist_add(hp, (a.w >> sh) & 255u, ok0);
      hist_add(hp, (b.x >> sh) & 255u, ok1);
      hist_add(hp, (b.y >> sh) & 255u, ok1);
      hist_add(hp, (b.z >> sh) & 255u, ok1);
      hist_add(hp, (b.w >> sh) & 255u, ok1);
    }
  }
  // The keys after the last full 16-byte group (all of them without vector loads).
  for (size_t wi = 4 * n4 + first - (threadIdx.x & 63), i = 4 * n4 + first; wi < n; wi += stride, i += stride) {
    const bool ok = i < n;
    const uint32_t k = ok ? keys[i] : 0u;
    for (int p = 0; p < passes; ++p) hist_add(h + p * kRadix, (k >> (8 * p)) & 255u, ok);
  }
  __syncthreads();
  uint32_t* copy = hist + (size_t)(blockIdx.x % kHistCopies) * 4 * kRadix;
  for (int i = threadIdx.x; i < passes * kRadix; i += kHistBlock) {
    const uint32_t v = s_hist[0][i] + s_hist[1][i] + s_hist[2][i] + s_hist[3][i];
    if (v) atomicAdd(&copy[i], v);
  }
}

// Exclusive scan of one value per thread over a kBlock-thread workgroup.
template <int kBlock>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp) {
  constexpr int kW = kBlock / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  uint32_t pre = 0;
#pragma unroll
  for (int i = 0; i < kW; ++i)
    if (i < w) pre += s_tmp[i];
  __syncthreads();
  return pre + x - v;
}

// Upsweep of a scanned pass: digit counts of kUpTiles consecutive tiles per workgroup (the same
// tiles the pass kernel will take), digit-major (counts[d * stride + tile]): each digit row gets a
// run of kUpTiles consecutive words from a workgroup instead of single scattered words.
// (Two tiles: 2^24 32-bit keys sort in 344 us against 358 with four and 347 with one; loading
// all of a workgroup's keys before counting was slower.)
constexpr int kUpTiles = 2;
template <int kKpt>
__global__ __launch_bounds__(kSortBlock) void upsweep_kernel(const uint32_t* __restrict__ keys, uint32_t n_host,
                                                             const uint32_t* __restrict__ n_dev, int shift,
                                                             uint32_t* __restrict__ counts, uint32_t stride) {
  constexpr int kTile = kSortBlock * kKpt;
  __shared__ uint32_t s_h[kUpTiles][kRadix];
  const uint32_t n = n_dev ? *n_dev : n_host;
  const uint32_t tiles = (n + kTile - 1) / kTile;
  // With the count known at launch, the same contiguous run of tiles per XCD as the scanned pass
  // kernel (see there), so that keys read here may still be in that XCD's L2 when the pass kernel
  // reads them again.
  const uint32_t groups = (tiles + kUpTiles - 1) / kUpTiles, gper = groups / 8u, grem = groups % 8u;
  const uint32_t gx = blockIdx.x % 8u;
  const uint32_t group =
      !n_dev && blockIdx.x < groups ? gx * gper + min(gx, grem) + blockIdx.x / 8u : blockIdx.x;
  const uint32_t tile0 = group * (uint32_t)kUpTiles;
  if (tile0 * (uint32_t)kTile >= n) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < kUpTiles * kRadix; i += kSortBlock) (&s_h[0][0])[i] = 0;
  __syncthreads();
#pragma unroll 1
  for (int g = 0; g < kUpTiles; ++g) {
    if (tile0 + g >= tiles) break;
    const uint32_t wbase = (tile0 + g) * (uint32_t)kTile + (uint32_t)w * 64u * kKpt;
    uint32_t k[kKpt];
#pragma unroll
    for (int i = 0; i < kKpt; ++i) {
      const uint32_t p = wbase + i * 64 + lane;
      k[i] = p < n ? keys[p] : 0u;
    }
#pragma unroll
    for (int i = 0; i < kKpt; ++i) hist_add(s_h[g], (k[i] >> shift) & 255u, wbase + i * 64 + lane < n);
  }
  __syncthreads();
  for (int i = t; i < kUpTiles * kRadix; i += kSortBlock) {
    const int d = i / kUpTiles, g = i % kUpTiles;
    if (tile0 + g < tiles) counts[(size_t)d * stride + tile0 + g] = s_h[g][d];
  }
}

// Exclusive prefix over the tiles of one digit's counts (in place; workgroup = digit, each
// thread a run of consecutive tiles) and the digit's total.
__global__ __launch_bounds__(256) void tile_scan_kernel(uint32_t* __restrict__ counts, uint32_t stride,
                                                        uint32_t n_host, const uint32_t* __restrict__ n_dev,
                                                        uint32_t tile_keys, uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_tmp[4];
  const uint32_t n = n_dev ? *n_dev : n_host;
  const uint32_t tiles = (n + tile_keys - 1) / tile_keys;
  uint32_t* row = counts + (size_t)blockIdx.x * stride;
  const uint32_t per = (tiles + 255) / 256;
  const uint32_t b = threadIdx.x * per, e = min(tiles, b + per);
  uint32_t sum = 0;
  for (uint32_t i = b; i < e; ++i) sum += row[i];
  const uint32_t pre = block_exclusive_scan<256>(sum, s_tmp);
  uint32_t run = pre;
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t c = row[i];
    row[i] = run;
    run += c;
  }
  if (threadIdx.x == 255) totals[blockIdx.x] = pre + sum;
}

// One digit pass over a tile of 512 * kKpt keys: stable rank inside the tile, then scatter
// keys and values to their global positions. The tile's global offset per digit comes either from
// a decoupled look-back over earlier tiles (onesweep: tiles ticketed in launch order, statuses
// tagged by epoch) or, with kScanned, from the per-tile digit prefixes of an upsweep + scan
// (tile = workgroup index, no cross-tile waiting).
template <bool kIota, int kKpt, bool kScanned>
__global__ __launch_bounds__(kSortBlock) void onesweep_kernel(const uint32_t* __restrict__ keys_in,
                                                              const uint32_t* __restrict__ vals_in,
                                                              uint32_t* __restrict__ keys_out,
                                                              uint32_t* __restrict__ vals_out, uint32_t n_host,
                                                              const uint32_t* __restrict__ n_dev, int shift,
                                                              const uint32_t* __restrict__ hist, int hist_copies,
                                                              uint32_t* __restrict__ counters, int pass, int passes,
                                                              uint64_t* __restrict__ status,
                                                              const uint32_t* __restrict__ tile_prefix,
                                                              uint32_t tile_stride) {
  constexpr int kTile = kSortBlock * kKpt;
  __shared__ uint32_t s_buf[kTile];
  __shared__ uint64_t s_peer[kWaves][kRadix];
  __shared__ uint32_t s_wave[kWaves][kRadix];
  __shared__ uint32_t s_start[kRadix];
  __shared__ uint32_t s_global[kRadix];
  __shared__ uint32_t s_tmp[kWaves];
  __shared__ uint32_t s_tile, s_epoch;

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) {
    if (kScanned) {
      // Workgroups go round robin to the 8 XCDs (one L2 each). Tile t's run of a digit ends in
      // the cache line where tile t + 1's run begins; giving each XCD one contiguous run of tiles
      // lets most such shared lines be completed in one L2 instead of leaving two partial lines,
      // and each XCD walks its run from the end, which its upsweep workgroups read last. Only when
      // the count is known at launch: with the count in device memory (sorts sized for a wave's
      // capacity) the extra dependent load before the tile is known cost more than the runs gain.
      const uint32_t g = (uint32_t)(((uint64_t)n_host + kTile - 1) / kTile), per = g / 8u, rem = g % 8u;
      const uint32_t x = blockIdx.x % 8u, cnt = per + (x < rem ? 1u : 0u);
      s_tile = !n_dev && blockIdx.x < g ? x * per + min(x, rem) + (cnt - 1u - blockIdx.x / 8u) : blockIdx.x;
      s_epoch = 0;
    } else {
      s_tile = atomicAdd(counters + pass, 1u);
      s_epoch = counters[kEpochWord] - (uint32_t)passes + 1u + (uint32_t)pass;
    }
  }
  for (int i = t; i < kWaves * kRadix; i += kSortBlock) {
    (&s_wave[0][0])[i] = 0;
    (&s_peer[0][0])[i] = 0;
  }
  const uint32_t n = n_dev ? *n_dev : n_host;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t epoch = s_epoch;
  const uint32_t tile_base = tile * (uint32_t)kTile;
  // Launches are sized for the largest wave; tickets past the current count have no keys, and
  // no earlier tile waits on them.
  if (tile_base >= n) return;
  const uint32_t valid = min((uint32_t)kTile, n - tile_base);
  const uint32_t wbase = (uint32_t)w * 64u * kKpt;
  // Thread t < 256 owns digit t from here to the end.
  uint32_t digit_total = 0;
  if (t < kRadix)
    for (int c = 0; c < hist_copies; ++c) digit_total += hist[(size_t)c * 4 * kRadix + t];
  const uint32_t tile_pre = kScanned && t < kRadix ? tile_prefix[(size_t)t * tile_stride + tile] : 0u;

  uint32_t key[kKpt], rank[kKpt];
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    key[i] = p < valid ? keys_in[tile_base + p] : 0xFFFFFFFFu;
  }

  // Stable rank of every key among equal digits of its wave, in original order: the running
  // count of the digit in this wave (an LDS add by the lowest peer) plus the peers below the lane.
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t my_bit = 1ull << lane;
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const bool ok = wbase + i * 64 + lane < valid;
    const uint32_t d = (key[i] >> shift) & 255u;
    const uint64_t active = __ballot(ok);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    uint64_t peers;
    if ((active & 1ull) && __ballot(ok && d == d0) == active) {
      peers = active;
    } else {
      uint64_t* word = &s_peer[w][d];
      if (ok) __hip_atomic_fetch_or(word, my_bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      peers = ok ? __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0ull;
      if (ok) __hip_atomic_store(word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    const int leader = peers ? (__ffsll((unsigned long long)peers) - 1) : 0;
    uint32_t old = 0;
    if (ok && lane == leader) old = atomicAdd(&s_wave[w][d], (uint32_t)__popcll(peers));
    rank[i] = (uint32_t)__shfl((int)old, leader, 64) + (uint32_t)__popcll(peers & lt_mask);
  }
  uint32_t val[kKpt];
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    if (kIota)
      val[i] = tile_base + p;
    else
      val[i] = p < valid ? vals_in[tile_base + p] : 0u;
  }
  __syncthreads();

  // Per-digit wave prefixes and tile totals.
  uint32_t total = 0;
  if (t < kRadix) {
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      const uint32_t c = s_wave[ww][t];
      s_wave[ww][t] = total;
      total += c;
    }
  }
  const uint64_t tag = (uint64_t)epoch << 32;
  uint64_t* my_status = kScanned ? nullptr : status + (size_t)tile * kRadix + t;
  if (!kScanned && t < kRadix)
    __hip_atomic_store(my_status, tag | (uint64_t)((tile == 0 ? kFlagPre : kFlagAgg) | total), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // Global digit bases of this pass: exclusive scan of the histogram; the tile's digit starts.
  const uint32_t digit_base = block_exclusive_scan<kSortBlock>(digit_total, s_tmp);
  const uint32_t start = block_exclusive_scan<kSortBlock>(total, s_tmp);

  if (t < kRadix) {
    uint32_t excl = tile_pre;
    if (!kScanned && tile > 0) {
      // Decoupled look-back over earlier tiles for digit t.
      int pred = (int)tile - 1;
      for (;;) {
        const uint64_t v = __hip_atomic_load(status + (size_t)pred * kRadix + t, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t lo = (uint32_t)v;
        const uint32_t flag = lo & ~kCountMask;
        if ((uint32_t)(v >> 32) != epoch || flag == 0) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += lo & kCountMask;
        if (flag == kFlagPre) break;
        --pred;
      }
      __hip_atomic_store(my_status, tag | (uint64_t)(kFlagPre | (excl + total)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_start[t] = start;
    s_global[t] = digit_base + excl - start;  // wraps; start <= tile position of the digit's keys
  }
  __syncthreads();

  // Stage keys in digit order.
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    if (p < valid) {
      const uint32_t d = (key[i] >> shift) & 255u;
      const uint32_t pos = s_start[d] + s_wave[w][d] + rank[i];
      s_buf[pos] = key[i];
      rank[i] = pos;
    }
  }
  __syncthreads();
  uint32_t dest[kKpt];
#pragma unroll
  for (int j = 0; j < kKpt; ++j) {
    const uint32_t p = j * kSortBlock + t;
    if (p < valid) {
      const uint32_t k = s_buf[p];
      dest[j] = s_global[(k >> shift) & 255u] + p;
      keys_out[dest[j]] = k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    if (p < valid) s_buf[rank[i]] = val[i];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kKpt; ++j) {
    const uint32_t p = j * kSortBlock + t;
    if (p < valid) vals_out[dest[j]] = s_buf[p];
  }
}

// ---- bucket sort: one partition pass over the key range, then every bucket sorted on chip ----
// Four 8-bit LSD passes move each pair through HBM four times. The bucket sort moves it twice:
//   1. the range [kmin, kmax] of the keys (one read of the keys);
//   2. a stable partition into 2048 buckets by the top 11 bits of key - kmin over that range
//      (upsweep of per-tile bucket counts, a scan over tiles, then the scatter: the scanned pass
//      structure above with an 11-bit digit and 8 K-key tiles);
//   3. each bucket (about n / 2048 pairs on keys spread over their range) loaded into LDS, sorted
//      stably on its remaining bits by LSD passes of at most 8 bits inside the workgroup, and
//      written once into the caller's buffers.
// A bucket larger than the LDS holds is sorted by its workgroup through device memory instead
// (LSD passes over its range of both buffers), which is correct for any key distribution and
// slow only for heavily clustered keys.
constexpr int kBkBits = 11;
constexpr int kBkBins = 1 << kBkBits;
constexpr int kBkBlock = 512;
constexpr int kBkWaves = kBkBlock / 64;
// Keys per lane of a partition tile: 16 (8 K-key tiles, two workgroups per CU) measured faster
// than 32 (16 K-key tiles, one workgroup per CU): 0.373 against 0.386 ms for 2^24 ray keys.
constexpr int kBkKpt = 16;
constexpr int kBkTile = kBkBlock * kBkKpt;
constexpr int kBkRangeBlock = 1024;
constexpr int kBkRangeGroups = 256;  // at most: the consumers reduce one (min, max) pair per group
constexpr int kBkLocalBlock = 1024;  // lanes of the local sort kernels
constexpr int kBkLocalWaves = kBkLocalBlock / 64;
constexpr int kBkLocalKpt = 12;
constexpr int kBkCap = kBkLocalBlock * kBkLocalKpt;  // largest bucket sorted in LDS (12288 pairs)

// Per-workgroup (min, max) of the keys: mm[2 g], mm[2 g + 1]. Also clears the count of buckets
// beyond the LDS that the local kernel appends to.
__global__ __launch_bounds__(kBkRangeBlock) void bk_range_kernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                                 uint32_t* __restrict__ mm, uint32_t* __restrict__ big) {
  __shared__ uint32_t s_lo[kBkRangeBlock / 64], s_hi[kBkRangeBlock / 64];
  const uint32_t t = threadIdx.x, stride = gridDim.x * kBkRangeBlock;
  if (blockIdx.x == 0 && t == 0) big[0] = 0;
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  const uint32_t n4 = (reinterpret_cast<uintptr_t>(keys) & 15u) == 0 ? n / 4 : 0;
  const uint4* k4 = reinterpret_cast<const uint4*>(keys);
  uint32_t i = blockIdx.x * kBkRangeBlock + t;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const uint4 a = k4[i], b = k4[i + stride], c = k4[i + 2 * stride], d = k4[i + 3 * stride];
    lo = min(lo, min(min(min(a.x, a.y), min(a.z, a.w)), min(min(b.x, b.y), min(b.z, b.w))));
    lo = min(lo, min(min(min(c.x, c.y), min(c.z, c.w)), min(min(d.x, d.y), min(d.z, d.w))));
    hi = max(hi, max(max(max(a.x, a.y), max(a.z, a.w)), max(max(b.x, b.y), max(b.z, b.w))));
    hi = max(hi, max(max(max(c.x, c.y), max(c.z, c.w)), max(max(d.x, d.y), max(d.z, d.w))));
  }
  for (; i < n4; i += stride) {
    const uint4 a = k4[i];
    lo = min(lo, min(min(a.x, a.y), min(a.z, a.w)));
    hi = max(hi, max(max(a.x, a.y), max(a.z, a.w)));
  }
  for (uint32_t j = 4 * n4 + blockIdx.x * kBkRangeBlock + t; j < n; j += stride) {
    lo = min(lo, keys[j]);
    hi = max(hi, keys[j]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, off, 64));
  }
  if ((t & 63) == 0) {
    s_lo[t >> 6] = lo;
    s_hi[t >> 6] = hi;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < kBkRangeBlock / 64; ++w) {
      lo = min(lo, s_lo[w]);
      hi = max(hi, s_hi[w]);
    }
    mm[2 * blockIdx.x] = lo;
    mm[2 * blockIdx.x + 1] = hi;
  }
}

// The partition digit of a key is (key - kmin) >> shift, with shift chosen so that the digit of
// kmax is below 2048: the top 11 bits of the range. Buckets are ordered like their keys.
struct BkRange {
  uint32_t kmin, shift;
};
// Reduces the range kernel's per-group pairs; every thread of the kBlock-lane workgroup calls it
// (s_red: 2 kBlock / 64 words).
template <int kBlock = kBkBlock>
__device__ inline BkRange bk_key_range(const uint32_t* __restrict__ mm, int groups, uint32_t* s_red) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  constexpr int kW = kBlock / 64;
  uint32_t lo = t < groups ? mm[2 * t] : 0xFFFFFFFFu;
  uint32_t hi = t < groups ? mm[2 * t + 1] : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, off, 64));
  }
  if (lane == 0) {
    s_red[w] = lo;
    s_red[kW + w] = hi;
  }
  __syncthreads();
  lo = s_red[0];
  hi = s_red[kW];
#pragma unroll
  for (int i = 1; i < kW; ++i) {
    lo = min(lo, s_red[i]);
    hi = max(hi, s_red[kW + i]);
  }
  __syncthreads();
  const uint32_t span = hi - lo;
  const uint32_t bits = span ? 32u - (uint32_t)__clz((int)span) : 0u;
  return BkRange{lo, bits > (uint32_t)kBkBits ? bits - (uint32_t)kBkBits : 0u};
}

// Tile t of a partition launch with a grid of exactly `tiles` workgroups: each XCD takes one
// contiguous run of tiles and walks it from its end (see the scanned pass kernel above).
__device__ inline uint32_t bk_tile_of_block(uint32_t tiles) {
  const uint32_t per = tiles / 8u, rem = tiles % 8u, x = blockIdx.x % 8u, cnt = per + (x < rem ? 1u : 0u);
  return x * per + min(x, rem) + (cnt - 1u - blockIdx.x / 8u);
}

// Bucket counts of one partition tile, tile-major: counts[tile * 2048 + bucket].
template <int kKpt>
__global__ __launch_bounds__(kBkBlock) void bk_upsweep_kernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                              const uint32_t* __restrict__ mm, int groups,
                                                              uint32_t* __restrict__ counts) {
  constexpr int kTile = kBkBlock * kKpt;
  __shared__ uint32_t s_h[kBkBins];
  __shared__ uint32_t s_red[2 * kBkWaves];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint32_t tiles = (n + kTile - 1) / kTile;
  const uint32_t tile = bk_tile_of_block(tiles);
  for (int i = t; i < kBkBins; i += kBkBlock) s_h[i] = 0;
  const BkRange r = bk_key_range(mm, groups, s_red);  // synchronises the workgroup
  const uint32_t wbase = tile * (uint32_t)kTile + (uint32_t)w * 64u * kKpt;
  uint32_t k[kKpt];
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    k[i] = p < n ? keys[p] : 0u;
  }
#pragma unroll
  for (int i = 0; i < kKpt; ++i) hist_add(s_h, (k[i] - r.kmin) >> r.shift, wbase + i * 64 + lane < n);
  __syncthreads();
  uint32_t* row = counts + (size_t)tile * kBkBins;
  for (int i = t; i < kBkBins; i += kBkBlock) row[i] = s_h[i];
}

// Exclusive prefix over tiles of every bucket's counts (in place) and each bucket's total.
// A workgroup takes 16 buckets; lane = (tile group g = lane / 16, bucket = lane % 16), so a wave
// reads four 64-byte row pieces per load.
constexpr int kBkScanBlock = 1024;
constexpr int kBkScanCols = 16;
__global__ __launch_bounds__(kBkScanBlock) void bk_scan_kernel(uint32_t* __restrict__ counts, uint32_t tiles,
                                                               uint32_t* __restrict__ totals) {
  constexpr int kGroups = kBkScanBlock / kBkScanCols;
  __shared__ uint32_t s_part[kGroups][kBkScanCols];
  const int t = threadIdx.x, c = t % kBkScanCols, g = t / kBkScanCols;
  const uint32_t b = blockIdx.x * kBkScanCols + c;
  const uint32_t per = (tiles + kGroups - 1) / kGroups, tb = min(tiles, g * per), te = min(tiles, tb + per);
  uint32_t sum = 0;
#pragma unroll 8
  for (uint32_t i = tb; i < te; ++i) sum += counts[(size_t)i * kBkBins + b];
  s_part[g][c] = sum;
  __syncthreads();
  uint32_t pre = 0;
  for (int gg = 0; gg < g; ++gg) pre += s_part[gg][c];
  if (g == kGroups - 1) totals[b] = pre + sum;
#pragma unroll 8
  for (uint32_t i = tb; i < te; ++i) {
    uint32_t* p = counts + (size_t)i * kBkBins + b;
    const uint32_t v = *p;
    *p = pre;
    pre += v;
  }
}

// The partition pass: a tile of 512 * kKpt keys ranked stably by bucket inside each wave (an LDS
// peer word over the bucket's low 8 bits, refined by three ballots over its top 3 bits), staged in
// LDS in bucket order and written as runs of equal buckets to bucket base + earlier tiles' count
// + position in the tile's run. Per-wave bucket counts are 16-bit, two to a word.
template <bool kIota, int kKpt>
__global__ __launch_bounds__(kBkBlock) void bk_partition_kernel(const uint32_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out, uint32_t n,
                                                                const uint32_t* __restrict__ mm, int groups,
                                                                const uint32_t* __restrict__ tile_prefix,
                                                                const uint32_t* __restrict__ totals) {
  constexpr int kTile = kBkBlock * kKpt;
  constexpr int kWords = kBkBins / 2;
  static_assert(kTile < 65536, "16-bit per-wave bucket counts");
  // Per-wave bucket counts (two to a word, then a word for lanes past the end of the tile) while
  // ranking; once every key's place in the tile is known, the same LDS stages the tile.
  constexpr int kCountWords = kBkWaves * (kWords + 1);
  __shared__ uint32_t s_mem[kTile > kCountWords ? kTile : kCountWords];
  uint32_t(*s_wave)[kWords + 1] = reinterpret_cast<uint32_t(*)[kWords + 1]>(s_mem);
  uint32_t* s_buf = s_mem;
  // Peer words while ranking (by a bucket's low 8 bits; word 256 for lanes past the end);
  // afterwards (all zero again) the tile's global bucket offsets.
  __shared__ uint64_t s_peer[kBkWaves][kRadix + 1];
  __shared__ uint32_t s_red[2 * kBkWaves];
  static_assert(sizeof(uint64_t) * kBkWaves * kRadix >= sizeof(uint32_t) * kBkBins, "offsets fit the peer words");
  uint32_t* s_global = reinterpret_cast<uint32_t*>(&s_peer[0][0]);

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint32_t tiles = (n + kTile - 1) / kTile;
  const uint32_t tile = bk_tile_of_block(tiles);
  for (int i = t; i < kBkWaves * (kWords + 1); i += kBkBlock) (&s_wave[0][0])[i] = 0;
  for (int i = t; i < kBkWaves * (kRadix + 1); i += kBkBlock) (&s_peer[0][0])[i] = 0;
  const BkRange r = bk_key_range(mm, groups, s_red);  // synchronises the workgroup
  const uint32_t tile_base = tile * (uint32_t)kTile;
  const uint32_t valid = min((uint32_t)kTile, n - tile_base);
  const uint32_t wbase = (uint32_t)w * 64u * kKpt;

  uint32_t key[kKpt], rank[kKpt];
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    key[i] = p < valid ? keys_in[tile_base + p] : 0u;
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t my_bit = 1ull << lane;
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    // lanes past the end of the tile rank under bucket 2048, in peer word 256 and counter word
    // 1024 of their own: no per-lane branches
    const bool ok = wbase + i * 64 + lane < valid;
    const uint32_t d = ok ? (key[i] - r.kmin) >> r.shift : (uint32_t)kBkBins;
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    uint64_t peers;
    if (__ballot(d == d0) == ~0ull) {
      peers = ~0ull;
    } else {
      uint64_t* word = &s_peer[w][ok ? (d & 255u) : (uint32_t)kRadix];
      __hip_atomic_fetch_or(word, my_bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      peers = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      __hip_atomic_store(word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      const uint32_t hi = d >> 8;
      const uint64_t b0 = __ballot(hi & 1u), b1 = __ballot(hi & 2u), b2 = __ballot(hi & 4u);
      peers &= ((hi & 1u) ? b0 : ~b0) & ((hi & 2u) ? b1 : ~b1) & ((hi & 4u) ? b2 : ~b2);
    }
    const int leader = __ffsll((unsigned long long)peers) - 1;  // peers hold the lane itself
    uint32_t old = 0;
    if (lane == leader) {
      const uint32_t sh = 16u * (d & 1u);
      old = (atomicAdd(&s_wave[w][d >> 1], (uint32_t)__popcll(peers) << sh) >> sh) & 0xFFFFu;
    }
    rank[i] = (uint32_t)__shfl((int)old, leader, 64) + (uint32_t)__popcll(peers & lt_mask);
  }
  uint32_t val[kKpt];
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    if (kIota)
      val[i] = tile_base + p;
    else
      val[i] = p < valid ? vals_in[tile_base + p] : 0u;
  }
  __syncthreads();

  // Thread t owns buckets 4t .. 4t+3 (words 2t and 2t+1 of every wave's row): per-wave prefixes,
  // the tile's bucket counts, then the bucket starts in the tile and the global run bases.
  uint32_t run0 = 0, run1 = 0, run2 = 0, run3 = 0;
#pragma unroll
  for (int ww = 0; ww < kBkWaves; ++ww) {
    const uint32_t c0 = s_wave[ww][2 * t], c1 = s_wave[ww][2 * t + 1];
    s_wave[ww][2 * t] = run0 | (run1 << 16);
    s_wave[ww][2 * t + 1] = run2 | (run3 << 16);
    run0 += c0 & 0xFFFFu;
    run1 += c0 >> 16;
    run2 += c1 & 0xFFFFu;
    run3 += c1 >> 16;
  }
  const uint4 gt = reinterpret_cast<const uint4*>(totals)[t];
  const uint4 tp = reinterpret_cast<const uint4*>(tile_prefix + (size_t)tile * kBkBins)[t];
  const uint32_t gpre = block_exclusive_scan<kBkBlock>(gt.x + gt.y + gt.z + gt.w, s_red);
  const uint32_t tpre = block_exclusive_scan<kBkBlock>(run0 + run1 + run2 + run3, s_red);
  const uint32_t st0 = tpre, st1 = st0 + run0, st2 = st1 + run1, st3 = st2 + run2;
  const uint32_t gb0 = gpre, gb1 = gb0 + gt.x, gb2 = gb1 + gt.y, gb3 = gb2 + gt.z;
  // wraps; a bucket's start in the tile is at most the position of its keys
  s_global[4 * t + 0] = gb0 + tp.x - st0;
  s_global[4 * t + 1] = gb1 + tp.y - st1;
  s_global[4 * t + 2] = gb2 + tp.z - st2;
  s_global[4 * t + 3] = gb3 + tp.w - st3;
#pragma unroll
  for (int ww = 0; ww < kBkWaves; ++ww) {
    s_wave[ww][2 * t] += st0 | (st1 << 16);
    s_wave[ww][2 * t + 1] += st2 | (st3 << 16);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t d = (key[i] - r.kmin) >> r.shift;
    if (wbase + i * 64 + lane < valid) rank[i] += (s_wave[w][d >> 1] >> (16u * (d & 1u))) & 0xFFFFu;
  }
  __syncthreads();  // the counts are dead: their LDS now stages the tile
#pragma unroll
  for (int i = 0; i < kKpt; ++i)
    if (wbase + i * 64 + lane < valid) s_buf[rank[i]] = key[i];
  __syncthreads();
  uint32_t dest[kKpt];
#pragma unroll
  for (int j = 0; j < kKpt; ++j) {
    const uint32_t p = j * kBkBlock + t;
    if (p < valid) {
      const uint32_t k = s_buf[p];
      dest[j] = s_global[(k - r.kmin) >> r.shift] + p;
      keys_out[dest[j]] = k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    const uint32_t p = wbase + i * 64 + lane;
    if (p < valid) s_buf[rank[i]] = val[i];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kKpt; ++j) {
    const uint32_t p = j * kBkBlock + t;
    if (p < valid) vals_out[dest[j]] = s_buf[p];
  }
}

// Digit of a key inside one bucket: the relative key r = (key - kmin) mod 2^shift, less the
// bucket's smallest r, shifted down to the pass's bits.
struct BkDigit {
  uint32_t kmin, rmask, lo, sh, dmask;
  __device__ uint32_t operator()(uint32_t k) const { return ((((k - kmin) & rmask) - lo) >> sh) & dmask; }
};

// Lanes past the end of a chunk rank under this digit, with a peer word and a counter of their
// own, so that the ranking loop needs no per-lane branches; its counts are never read.
constexpr int kBkSink = kRadix;

// LDS of the local sort kernels.
struct BkLocalShared {
  uint2 buf[kBkCap];  // (key, value) pairs staged in digit order
  uint64_t peer[kBkLocalWaves][kRadix + 1];
  uint32_t wave[kBkLocalWaves][kRadix + 1];
  uint32_t start[kRadix];
  uint32_t run[kRadix];
  uint32_t red[2 * kBkLocalWaves];
  uint32_t bucket[2];
};

// Stable 8-bit ranking of the keys key[0 .. ke) of every lane (positions wbase + i * 64 + lane of
// the chunk, those from `valid` on in the sink) inside each wave, by the peer-word scheme of the
// pass kernels; then per-digit wave prefixes and the chunk's digit starts, so that a key's place in
// the chunk in digit order is sh.wave[w][digit] + rank. Thread t < 256 receives digit t's count in
// the chunk as the result and its start in *start. (Issuing every key's peer-word operations
// before any counter add, one wait per phase instead of per key, was slower: the kernel is bound
// by LDS throughput, not by the wait.)
template <int kKpt>
__device__ inline uint32_t bk_rank8(const uint32_t (&key)[kKpt], const BkDigit& dg, uint32_t (&rank)[kKpt],
                                    uint32_t ke, uint32_t wbase, uint32_t valid, BkLocalShared& sh,
                                    uint32_t* start) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t my_bit = 1ull << lane;
#pragma unroll
  for (int i = 0; i < kKpt; ++i) {
    uint32_t rk = 0;  // assigned to rank[i] once (see the partition kernel)
    if ((uint32_t)i < ke) {
      const uint32_t d = wbase + i * 64 + lane < valid ? dg(key[i]) : (uint32_t)kBkSink;
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
      uint64_t peers;
      if (__ballot(d == d0) == ~0ull) {
        peers = ~0ull;
      } else {
        uint64_t* word = &sh.peer[w][d];
        __hip_atomic_fetch_or(word, my_bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        peers = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_store(word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
      // Every lane reads its digit's running count (before this key's group adds to it); the
      // group's lowest lane alone writes the new count. A wave's LDS operations execute in order,
      // so the next key reads what this one wrote.
      uint32_t* cnt = &sh.wave[w][d];
      const uint32_t before = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      if (lane == __ffsll((unsigned long long)peers) - 1)
        __hip_atomic_store(cnt, before + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      rk = before + (uint32_t)__popcll(peers & lt_mask);
    }
    rank[i] = rk;
  }
  __syncthreads();
  uint32_t total = 0;
  if (t < kRadix) {
#pragma unroll
    for (int ww = 0; ww < kBkLocalWaves; ++ww) {
      const uint32_t c = sh.wave[ww][t];
      sh.wave[ww][t] = total;
      total += c;
    }
  }
  const uint32_t st = block_exclusive_scan<kBkLocalBlock>(total, sh.red);
  if (t < kRadix) {
#pragma unroll
    for (int ww = 0; ww < kBkLocalWaves; ++ww) sh.wave[ww][t] += st;
  }
  *start = st;
  __syncthreads();
  return total;
}

// Clears the per-wave digit counts (sink included) for the next ranking.
__device__ inline void bk_clear_counts(BkLocalShared& sh) {
  for (int i = threadIdx.x; i < kBkLocalWaves * (kRadix + 1); i += kBkLocalBlock) (&sh.wave[0][0])[i] = 0;
}

// Bucket b's place in the partition output (thread b / 4 writes it) and the key range; every
// thread of the workgroup calls it.
__device__ inline BkRange bk_bucket(uint32_t b, const uint32_t* __restrict__ totals, const uint32_t* __restrict__ mm,
                                    int groups, BkLocalShared& sh) {
  const int t = threadIdx.x;
  const uint4 gt = t < kBkBins / 4 ? reinterpret_cast<const uint4*>(totals)[t] : make_uint4(0u, 0u, 0u, 0u);
  const uint32_t pre = block_exclusive_scan<kBkLocalBlock>(gt.x + gt.y + gt.z + gt.w, sh.red);
  if ((uint32_t)t == b / 4u) {
    const uint32_t c[4] = {gt.x, gt.y, gt.z, gt.w};
    uint32_t s = pre;
    for (uint32_t k = 0; k < b % 4u; ++k) s += c[k];
    sh.bucket[0] = s;
    sh.bucket[1] = s + c[b % 4u];
  }
  return bk_key_range<kBkLocalBlock>(mm, groups, sh.red);  // synchronises the workgroup
}

// Smallest and largest of a (local sort) workgroup's lo / hi.
__device__ inline void bk_block_minmax(uint32_t& lo, uint32_t& hi, uint32_t* s_red) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, off, 64));
  }
  if (lane == 0) {
    s_red[w] = lo;
    s_red[kBkLocalWaves + w] = hi;
  }
  __syncthreads();
  lo = s_red[0];
  hi = s_red[kBkLocalWaves];
#pragma unroll
  for (int i = 1; i < kBkLocalWaves; ++i) {
    lo = min(lo, s_red[i]);
    hi = max(hi, s_red[kBkLocalWaves + i]);
  }
  __syncthreads();
}

// Pass count and digit function of a bucket whose relative keys span [lo, hi]: the fewest LSD
// passes of at most 8 bits, all of one width.
__device__ inline BkDigit bk_digit_plan(const BkRange& r, uint32_t rmask, uint32_t lo, uint32_t hi, uint32_t* passes,
                                        uint32_t* width) {
  const uint32_t span = hi - lo;
  const uint32_t bits = span ? 32u - (uint32_t)__clz((int)span) : 0u;
  *passes = (bits + 7u) / 8u;
  *width = *passes ? (bits + *passes - 1u) / *passes : 0u;
  return BkDigit{r.kmin, rmask, lo, 0u, (1u << *width) - 1u};
}

// The buckets beyond the LDS, listed by the local kernel (big[0] buckets in big[1 ..]): each
// sorted by one workgroup through the bucket's range of both buffer pairs, by LSD passes, each a
// histogram of the whole bucket, then chunks of kBkCap pairs in order, each ranked in LDS and
// written at the running digit offsets. An odd pass count ends in (keys_out, vals_out) by itself;
// an even one first copies the bucket there and runs its passes back from it. Correct for any key
// distribution; slow only when the keys crowd into few buckets.
__global__ __launch_bounds__(kBkLocalBlock) void bk_local_big_kernel(uint32_t* keys_in, uint32_t* vals_in,
                                                                uint32_t* keys_out, uint32_t* vals_out,
                                                                const uint32_t* __restrict__ mm, int groups,
                                                                const uint32_t* __restrict__ totals,
                                                                const uint32_t* __restrict__ big) {
  constexpr int kKpt = kBkLocalKpt;
  __shared__ BkLocalShared sh;
  if (blockIdx.x >= big[0]) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint32_t wbase = (uint32_t)w * 64u * kKpt;
  for (int i = t; i < kBkLocalWaves * (kRadix + 1); i += kBkLocalBlock) (&sh.peer[0][0])[i] = 0;
  bk_clear_counts(sh);
  const BkRange r = bk_bucket(big[1 + blockIdx.x], totals, mm, groups, sh);
  const uint32_t bs = sh.bucket[0], m = sh.bucket[1] - sh.bucket[0];
  const uint32_t rmask = r.shift ? (0xFFFFFFFFu >> (32u - r.shift)) : 0u;
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  for (uint32_t p = t; p < m; p += kBkLocalBlock) {
    const uint32_t rk = (keys_in[bs + p] - r.kmin) & rmask;
    lo = min(lo, rk);
    hi = max(hi, rk);
  }
  bk_block_minmax(lo, hi, sh.red);
  uint32_t passes, width;
  BkDigit dg = bk_digit_plan(r, rmask, lo, hi, &passes, &width);
  uint32_t* src_k = keys_in;
  uint32_t* src_v = vals_in;
  uint32_t* dst_k = keys_out;
  uint32_t* dst_v = vals_out;
  if ((passes & 1u) == 0) {
    for (uint32_t p = t; p < m; p += kBkLocalBlock) {
      keys_out[bs + p] = keys_in[bs + p];
      vals_out[bs + p] = vals_in[bs + p];
    }
    if (passes == 0) return;
    __threadfence_block();
    __syncthreads();
    src_k = keys_out;
    src_v = vals_out;
    dst_k = keys_in;
    dst_v = vals_in;
  }
  uint32_t key[kKpt], rank[kKpt];
  for (uint32_t q = 0; q < passes; ++q) {
    dg.sh = q * width;
    if (t < kRadix) sh.run[t] = 0;
    __syncthreads();
    for (uint32_t p0 = 0; p0 < m; p0 += kBkLocalBlock) {
      const uint32_t p = p0 + t;
      const bool ok = p < m;
      hist_add(sh.run, dg(ok ? src_k[bs + p] : 0u), ok);
    }
    __syncthreads();
    {
      const uint32_t c = t < kRadix ? sh.run[t] : 0u;
      const uint32_t pre = block_exclusive_scan<kBkLocalBlock>(c, sh.red);
      if (t < kRadix) sh.run[t] = pre;
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < m; c0 += kBkCap) {
      const uint32_t valid = min((uint32_t)kBkCap, m - c0);
#pragma unroll
      for (int i = 0; i < kKpt; ++i) {
        const uint32_t p = wbase + i * 64 + lane;
        key[i] = p < valid ? src_k[bs + c0 + p] : 0u;
      }
      uint32_t st;
      const uint32_t total = bk_rank8<kKpt>(key, dg, rank, kKpt, wbase, valid, sh, &st);
      if (t < kRadix) sh.start[t] = st;
#pragma unroll
      for (int i = 0; i < kKpt; ++i) {
        const uint32_t p = wbase + i * 64 + lane;
        if (p < valid) sh.buf[rank[i] + sh.wave[w][dg(key[i])]] = make_uint2(key[i], src_v[bs + c0 + p]);
      }
      __syncthreads();
      bk_clear_counts(sh);
#pragma unroll
      for (int j = 0; j < kKpt; ++j) {
        const uint32_t p = j * kBkLocalBlock + t;
        if (p < valid) {
          const uint2 kv = sh.buf[p];
          const uint32_t d = dg(kv.x);
          const uint32_t dest = bs + sh.run[d] + p - sh.start[d];
          dst_k[dest] = kv.x;
          dst_v[dest] = kv.y;
        }
      }
      __syncthreads();
      if (t < kRadix) sh.run[t] += total;
      __syncthreads();
    }
    // this pass's output is the next pass's input, read by other lanes of the workgroup
    __threadfence_block();
    __syncthreads();
    uint32_t* tk = src_k;
    uint32_t* tv = src_v;
    src_k = dst_k;
    src_v = dst_v;
    dst_k = tk;
    dst_v = tv;
  }
}

// Every bucket of the partition sorted on the bits below its digit, stably, from (keys_in,
// vals_in) into (keys_out, vals_out) at the same positions. A persistent workgroup per CU takes
// buckets g, g + G, g + 2G, ... (G workgroups) and loads its next bucket's pairs into registers
// while it sorts the current one: with one 150 KB workgroup per CU, a bucket's loads would
// otherwise wait behind the previous bucket's sort (the load and store of every bucket alone took
// 80 us for 2^24 pairs, twice the copy time). A bucket's pairs are spread evenly over the waves
// (ceil(m / 1024) per lane, in blocks per wave). Inside the bucket only the relative key
// r = (key - kmin) mod 2^shift - min r matters; its ceil(bits / 8) LSD passes of equal width run
// in LDS. Buckets of more than kBkCap pairs are appended to big (count in big[0], cleared by the
// range kernel) for bk_local_big_kernel.
__global__ __launch_bounds__(kBkLocalBlock) void bk_local_kernel(const uint32_t* __restrict__ keys_in,
                                                            const uint32_t* __restrict__ vals_in,
                                                            uint32_t* __restrict__ keys_out,
                                                            uint32_t* __restrict__ vals_out,
                                                            const uint32_t* __restrict__ mm, int groups,
                                                            const uint32_t* __restrict__ totals,
                                                            uint32_t* __restrict__ big) {
  constexpr int kKpt = kBkLocalKpt;
  __shared__ BkLocalShared sh;
  __shared__ uint32_t s_first[kBkBins + 1];  // bucket starts in the partition output
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < kBkLocalWaves * (kRadix + 1); i += kBkLocalBlock) (&sh.peer[0][0])[i] = 0;
  bk_clear_counts(sh);
  {
    const uint4 gt = t < kBkBins / 4 ? reinterpret_cast<const uint4*>(totals)[t] : make_uint4(0u, 0u, 0u, 0u);
    const uint32_t pre = block_exclusive_scan<kBkLocalBlock>(gt.x + gt.y + gt.z + gt.w, sh.red);
    if (t < kBkBins / 4) {
      s_first[4 * t] = pre;
      s_first[4 * t + 1] = pre + gt.x;
      s_first[4 * t + 2] = pre + gt.x + gt.y;
      s_first[4 * t + 3] = pre + gt.x + gt.y + gt.z;
      if (t == kBkBins / 4 - 1) s_first[kBkBins] = pre + gt.x + gt.y + gt.z + gt.w;
    }
  }
  const BkRange r = bk_key_range<kBkLocalBlock>(mm, groups, sh.red);  // synchronises the workgroup
  const uint32_t rmask = r.shift ? (0xFFFFFFFFu >> (32u - r.shift)) : 0u;

  uint32_t key[kKpt], val[kKpt], nkey[kKpt], nval[kKpt];
  // Loads bucket bb's pairs (if it fits the LDS) into (k, v): ceil(m / 1024) per lane.
  auto load = [&](uint32_t bb, uint32_t (&k)[kKpt], uint32_t (&v)[kKpt]) {
    const uint32_t bs = s_first[bb], m = s_first[bb + 1] - bs;
    const uint32_t ke = m <= (uint32_t)kBkCap ? (m + kBkLocalBlock - 1) / kBkLocalBlock : 0u;
    const uint32_t wbase = (uint32_t)w * 64u * ke;
#pragma unroll
    for (int i = 0; i < kKpt; ++i) {
      const uint32_t p = wbase + i * 64 + lane;
      const bool ok = (uint32_t)i < ke && p < m;
      k[i] = ok ? keys_in[bs + p] : 0u;
      v[i] = ok ? vals_in[bs + p] : 0u;
    }
  };
  uint32_t b = blockIdx.x;
  if (b < (uint32_t)kBkBins) load(b, key, val);
  for (; b < (uint32_t)kBkBins; b += gridDim.x) {
    const uint32_t bs = s_first[b], m = s_first[b + 1] - bs;
    // the next bucket's pairs, in flight while this one sorts
    if (b + gridDim.x < (uint32_t)kBkBins) load(b + gridDim.x, nkey, nval);
    if (m > (uint32_t)kBkCap) {
      if (t == 0) big[1 + atomicAdd(big, 1u)] = b;
    } else if (m > 0) {
      const uint32_t ke = (m + kBkLocalBlock - 1) / kBkLocalBlock;  // keys per lane
      const uint32_t wbase = (uint32_t)w * 64u * ke;
      uint32_t lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
      for (int i = 0; i < kKpt; ++i) {
        if ((uint32_t)i < ke && wbase + i * 64 + lane < m) {
          const uint32_t rk = (key[i] - r.kmin) & rmask;
          lo = min(lo, rk);
          hi = max(hi, rk);
        }
      }
      bk_block_minmax(lo, hi, sh.red);
      uint32_t passes, width;
      BkDigit dg = bk_digit_plan(r, rmask, lo, hi, &passes, &width);
      uint32_t rank[kKpt];
      for (uint32_t q = 0; q < passes; ++q) {
        dg.sh = q * width;
        uint32_t st;
        bk_rank8<kKpt>(key, dg, rank, ke, wbase, m, sh, &st);
#pragma unroll
        for (int i = 0; i < kKpt; ++i) {
          if ((uint32_t)i < ke && wbase + i * 64 + lane < m)
            sh.buf[rank[i] + sh.wave[w][dg(key[i])]] = make_uint2(key[i], val[i]);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kKpt; ++i) {
          const uint32_t p = wbase + i * 64 + lane;
          if ((uint32_t)i < ke && p < m) {
            const uint2 kv = sh.buf[p];
            key[i] = kv.x;
            val[i] = kv.y;
          }
        }
        bk_clear_counts(sh);  // every wave is past reading them
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < kKpt; ++i) {
        const uint32_t p = wbase + i * 64 + lane;
        if ((uint32_t)i < ke && p < m) {
          keys_out[bs + p] = key[i];
          vals_out[bs + p] = val[i];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kKpt; ++i) {
      key[i] = nkey[i];
      val[i] = nval[i];
    }
  }
}

}  // namespace

size_t sort_status_bytes(size_t n) {
  const size_t tile = (size_t)kSortBlock * kSortMinKeysPerThread;  // the smallest tile sizes the buffer
  const size_t tiles = (n + tile - 1) / tile;
  return (tiles ? tiles : 1) * kRadix * sizeof(uint64_t);
}

void sort_reserve(swr_context* ctx, size_t n) {
  SortWorkspace& ws = ctx->sort;
  ensure(ctx, ws.hist, kHistCopies * 4 * kRadix * sizeof(uint32_t));
  if (!ws.counters.ptr) {
    ensure(ctx, ws.counters, 2 * kEpochWord * sizeof(uint32_t));
    SWR_HIP(hipMemset(ws.counters.ptr, 0, ws.counters.bytes));
  }
  const size_t need = sort_status_bytes(n);
  if (ws.status.bytes < need) {
    ensure(ctx, ws.status, need);
    sort_reset_epochs(ctx, ctx->stream);
    SWR_HIP(hipStreamSynchronize(ctx->stream));
  }
  // Per-tile digit counts of the scanned passes: 256 rows of the tile count at the smallest tile,
  // rounded up to whole upsweep groups; the bucket sort's 2048 counts per 8 K-key tile share the
  // buffer.
  size_t tiles = (n + (size_t)kSortBlock * kSortMinKeysPerThread - 1) / ((size_t)kSortBlock * kSortMinKeysPerThread);
  tiles = (tiles + kUpTiles - 1) / kUpTiles * kUpTiles;
  const size_t stride = std::max<size_t>(tiles, ws.tile_stride);
  const size_t words = std::max((size_t)kRadix * stride, (size_t)kBkBins * ((n + kBkTile - 1) / kBkTile));
  if (ws.tile_counts.bytes < words * sizeof(uint32_t)) ensure(ctx, ws.tile_counts, words * sizeof(uint32_t));
  ws.tile_stride = (uint32_t)stride;
}

void sort_reset_epochs(swr_context* ctx, hipStream_t stream) {
  // Fresh or recycled memory may hold any bit pattern, including a current epoch tag.
  SortWorkspace& ws = ctx->sort;
  SWR_HIP(hipMemsetAsync(ws.status.ptr, 0, ws.status.bytes, stream));
  SWR_HIP(hipMemsetAsync(ws.counters.ptr, 0, ws.counters.bytes, stream));
  ws.epoch = 0;
}

void sort_epoch_guard(swr_context* ctx, uint64_t passes_ahead, hipStream_t stream) {
  if (ctx->sort.epoch + passes_ahead >= 0xFFFFFF00ull) sort_reset_epochs(ctx, stream);
}

namespace {

// Device buffers of one sort besides the pairs and their ping-pong partners.
struct PassBuffers {
  uint32_t* hist;         // digit totals (onesweep: kHistCopies copies summed by the pass kernel)
  uint32_t* counters;     // onesweep tile tickets and epoch counter
  uint64_t* status;       // onesweep look-back words
  uint32_t* tile_counts;  // reduce-then-scan per-tile digit counts, 256 rows of tile_stride
  uint32_t tile_stride;
  size_t tile_counts_bytes;
};

// Launches the passes of one sort (histogram first for onesweep); shapes checked on the host.
void launch_passes(const PassBuffers& b, bool scanned, int kpt, int num_cus, uint32_t* keys, uint32_t* vals,
                   uint32_t* alt_keys, uint32_t* alt_vals, size_t n, int passes, bool iota, hipStream_t stream,
                   const uint32_t* n_dev) {
  uint32_t* hist = b.hist;
  uint32_t* counters = b.counters;
  uint64_t* status = b.status;
  uint32_t* tile_counts = b.tile_counts;
  const uint32_t tile_stride = b.tile_stride;
  const uint32_t tile_keys = (uint32_t)(kSortBlock * kpt);
  const uint32_t tiles = checked_grid((n + tile_keys - 1) / tile_keys, kSortBlock, "sort pass");
  const uint32_t up_groups = checked_grid((tiles + kUpTiles - 1) / kUpTiles, kSortBlock, "sort upsweep");
  if (scanned && (uint64_t)up_groups * kUpTiles > tile_stride)
    fail(SWR_E_INVALID, "sort: per-tile count rows shorter than the tile count");
  if (!scanned) {
    sort_prepare_kernel<<<1, kRadix, 0, stream>>>(hist, counters, passes);
    const int hist_blocks = (int)std::max<size_t>(
        1, std::min<size_t>((n + 8 * kHistBlock - 1) / (8 * kHistBlock), (size_t)std::max(num_cus, 1)));
    if ((reinterpret_cast<uintptr_t>(keys) & 15u) == 0)
      sort_histogram_kernel<true><<<hist_blocks, kHistBlock, 0, stream>>>(keys, n, n_dev, passes, hist);
    else
      sort_histogram_kernel<false><<<hist_blocks, kHistBlock, 0, stream>>>(keys, n, n_dev, passes, hist);
  }
  uint32_t *sk = keys, *sv = vals, *dk = alt_keys, *dv = alt_vals;
  for (int p = 0; p < passes; ++p) {
    const bool first_iota = iota && p == 0;
    // Onesweep: the pass kernel alone. Reduce-then-scan: upsweep, per-digit scan (the digit
    // totals land in hist), then the pass kernel with the scanned per-tile offsets.
#define SWR_PASS(KPT)                                                                                           \
  if (scanned) {                                                                                                \
    upsweep_kernel<KPT><<<up_groups, kSortBlock, 0, stream>>>(sk, (uint32_t)n, n_dev, 8 * p, tile_counts,        \
                                                              tile_stride);                                     \
    tile_scan_kernel<<<kRadix, 256, 0, stream>>>(tile_counts, tile_stride, (uint32_t)n, n_dev, tile_keys,        \
                                                 hist + p * kRadix);                                            \
    if (first_iota)                                                                                             \
      onesweep_kernel<true, KPT, true><<<tiles, kSortBlock, 0, stream>>>(sk, sv, dk, dv, (uint32_t)n, n_dev,     \
          8 * p, hist + p * kRadix, 1, counters, p, passes, status, tile_counts, tile_stride);                   \
    else                                                                                                        \
      onesweep_kernel<false, KPT, true><<<tiles, kSortBlock, 0, stream>>>(sk, sv, dk, dv, (uint32_t)n, n_dev,    \
          8 * p, hist + p * kRadix, 1, counters, p, passes, status, tile_counts, tile_stride);                   \
  } else if (first_iota) {                                                                                      \
    onesweep_kernel<true, KPT, false><<<tiles, kSortBlock, 0, stream>>>(sk, sv, dk, dv, (uint32_t)n, n_dev,      \
        8 * p, hist + p * kRadix, kHistCopies, counters, p, passes, status, tile_counts, tile_stride);           \
  } else {                                                                                                      \
    onesweep_kernel<false, KPT, false><<<tiles, kSortBlock, 0, stream>>>(sk, sv, dk, dv, (uint32_t)n, n_dev,     \
        8 * p, hist + p * kRadix, kHistCopies, counters, p, passes, status, tile_counts, tile_stride);           \
  }
    switch (kpt) {
      case 8: SWR_PASS(8) break;
      case 12: SWR_PASS(12) break;
      case 16: SWR_PASS(16) break;
      case 20: SWR_PASS(20) break;
      case 24: SWR_PASS(24) break;
      default: fail(SWR_E_INVALID, "sort: keys per thread must be 8, 12, 16, 20 or 24");
    }
#undef SWR_PASS
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  SWR_HIP(hipGetLastError());
}

// hist holds the bucket totals, the range kernel's (min, max) pairs and the oversize list.
static_assert(kBkBins + 2 * kBkRangeGroups + 1 + kBkBins <= kHistCopies * 4 * kRadix, "bucket sort words fit hist");

// The bucket sort of n pairs from (keys, vals) back into them, through (alt_keys, alt_vals).
void launch_bucket_sort(const PassBuffers& b, uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals,
                        size_t n, bool iota, int num_cus, hipStream_t stream) {
  uint32_t* totals = b.hist;                                // 2048 bucket totals
  uint32_t* mm = b.hist + kBkBins;                          // per-group (min, max) of the keys
  uint32_t* big = b.hist + kBkBins + 2 * kBkRangeGroups;    // count, then the buckets beyond the LDS
  const uint32_t groups = checked_grid(
      std::min<size_t>(kBkRangeGroups, (n + 16 * kBkRangeBlock - 1) / (16 * kBkRangeBlock)), kBkRangeBlock,
      "bucket sort range");
  const uint32_t tiles = checked_grid((n + kBkTile - 1) / kBkTile, kBkBlock, "bucket sort partition");
  if ((size_t)tiles * kBkBins * sizeof(uint32_t) > b.tile_counts_bytes)
    fail(SWR_E_INVALID, "bucket sort: per-tile count buffer smaller than the tile count");
  bk_range_kernel<<<groups, kBkRangeBlock, 0, stream>>>(keys, (uint32_t)n, mm, big);
  bk_upsweep_kernel<kBkKpt><<<tiles, kBkBlock, 0, stream>>>(keys, (uint32_t)n, mm, (int)groups, b.tile_counts);
  bk_scan_kernel<<<kBkBins / kBkScanCols, kBkScanBlock, 0, stream>>>(b.tile_counts, tiles, totals);
  if (iota)
    bk_partition_kernel<true, kBkKpt><<<tiles, kBkBlock, 0, stream>>>(keys, vals, alt_keys, alt_vals, (uint32_t)n, mm,
                                                                      (int)groups, b.tile_counts, totals);
  else
    bk_partition_kernel<false, kBkKpt><<<tiles, kBkBlock, 0, stream>>>(keys, vals, alt_keys, alt_vals, (uint32_t)n, mm,
                                                                       (int)groups, b.tile_counts, totals);
  // one persistent local-sort workgroup per CU (its LDS allows no second)
  const uint32_t local_groups = (uint32_t)std::max(1, std::min(kBkBins, num_cus));
  bk_local_kernel<<<local_groups, kBkLocalBlock, 0, stream>>>(alt_keys, alt_vals, keys, vals, mm, (int)groups, totals,
                                                              big);
  // at most n / (kBkCap + 1) buckets exceed the LDS
  const uint32_t big_groups = (uint32_t)std::min<size_t>(kBkBins, n / (kBkCap + 1) + 1);
  bk_local_big_kernel<<<big_groups, kBkLocalBlock, 0, stream>>>(alt_keys, alt_vals, keys, vals, mm, (int)groups,
                                                                totals, big);
  SWR_HIP(hipGetLastError());
}

}  // namespace

void sort_pairs(swr_context* ctx, uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals, size_t n,
                int key_bits, bool iota, uint32_t** out_keys, uint32_t** out_vals, hipStream_t stream,
                const uint32_t* n_dev) {
  if (key_bits < 1 || key_bits > 32) fail(SWR_E_INVALID, "sort: key_bits must be in 1..32");
  if (n > (size_t(1) << 30)) fail(SWR_E_INVALID, "sort: at most 2^30 elements");
  const int passes = (key_bits + 7) / 8;
  // The final buffers depend only on the pass count, so a captured sort returns them too.
  *out_keys = (passes & 1) ? alt_keys : keys;
  *out_vals = (passes & 1) ? alt_vals : vals;
  if (n == 0) {
    *out_keys = keys;
    *out_vals = vals;
    return;
  }
  SortWorkspace& ws = ctx->sort;
  if (!n_dev) {
    sort_reserve(ctx, n);
    sort_epoch_guard(ctx, (uint64_t)passes, stream);
  }
  const PassBuffers b{static_cast<uint32_t*>(ws.hist.ptr), static_cast<uint32_t*>(ws.counters.ptr),
                      static_cast<uint64_t*>(ws.status.ptr), static_cast<uint32_t*>(ws.tile_counts.ptr),
                      ws.tile_stride, ws.tile_counts.bytes};
  // The bucket sort needs the count at launch and pays off from three 8-bit passes up; otherwise
  // its tuning runs the reduce-then-scan passes.
  if (ctx->tuning.sort_algorithm == SWR_SORT_BUCKET && !n_dev && passes >= 3) {
    launch_bucket_sort(b, keys, vals, alt_keys, alt_vals, n, iota, ctx->num_cus, stream);
    *out_keys = keys;
    *out_vals = vals;
    return;
  }
  ws.epoch += (uint64_t)passes;  // host mirror of the device epoch counter
  launch_passes(b, ctx->tuning.sort_algorithm != SWR_SORT_ONESWEEP, ctx->tuning.sort_kpt, ctx->num_cus, keys, vals,
                alt_keys, alt_vals, n, passes, iota, stream, n_dev);
}

// ---- the sort without allocation ----
// Everything the reduce-then-scan passes need besides the caller's pairs lives in one
// file-scope device array: the ping-pong partners of keys and values, the per-tile digit counts
// (tiles of the smallest size, rounded up to whole upsweep groups) and the digit totals. No word
// of it needs clearing: each pass writes every count and total it reads. Each device that loads
// the library holds its own copy until the process ends.
namespace {
constexpr size_t kScratchTiles =
    ((kNoAllocMaxPairs + kSortBlock * kSortMinKeysPerThread - 1) / (kSortBlock * kSortMinKeysPerThread) + kUpTiles -
     1) / kUpTiles * kUpTiles;
constexpr size_t kScratchWords = 2 * kNoAllocMaxPairs + kRadix * kScratchTiles + 4 * kRadix;
static __device__ __align__(256) uint32_t scratch[kScratchWords];
}  // namespace

size_t sort_scratch_bytes() { return sizeof(uint32_t) * kScratchWords; }

void sort_scratch_init(swr_context* ctx) {
  void* p = nullptr;
  SWR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail(SWR_E_HIP, "sort: no device address for the sort scratch array");
  ctx->sort_scratch = static_cast<uint32_t*>(p);
}

void sort_pairs_noalloc(swr_context* ctx, uint32_t* keys, uint32_t* vals, size_t n, int key_bits, bool iota,
                        hipStream_t stream) {
  if (key_bits < 1 || key_bits > 32) fail(SWR_E_INVALID, "sort: key_bits must be in 1..32");
  if (n > kNoAllocMaxPairs)
    fail(SWR_E_INVALID, "sort without allocation: " + std::to_string(n) + " pairs exceed its maximum of " +
                            std::to_string(kNoAllocMaxPairs));
  if (n == 0) return;
  uint32_t* base = ctx->sort_scratch;  // the device address of `scratch` on this context's device
  if (!base) fail(SWR_E_INVALID, "sort without allocation: scratch address not set up");
  uint32_t* alt_keys = base;
  uint32_t* alt_vals = base + kNoAllocMaxPairs;
  uint32_t* tile_counts = base + 2 * kNoAllocMaxPairs;
  uint32_t* hist = tile_counts + kRadix * kScratchTiles;
  const int passes = (key_bits + 7) / 8;
  // Reduce-then-scan only: its passes use no look-back words, tickets or epochs.
  const PassBuffers b{hist, nullptr, nullptr, tile_counts, (uint32_t)kScratchTiles,
                      sizeof(uint32_t) * kRadix * kScratchTiles};
  launch_passes(b, true, ctx->tuning.sort_kpt, ctx->num_cus, keys, vals, alt_keys, alt_vals, n, passes, iota, stream,
                nullptr);
  if (passes & 1) {
    SWR_HIP(hipMemcpyAsync(keys, alt_keys, n * 4, hipMemcpyDeviceToDevice, stream));
    SWR_HIP(hipMemcpyAsync(vals, alt_vals, n * 4, hipMemcpyDeviceToDevice, stream));
  }
}

}  // namespace swr

// Stable LSD radix sort of (key, value) pairs with 4-bit digits and no
// atomics: per-thread digit counts in registers, wave shuffle scans, one
// digit-major exclusive scan over the per-tile histograms, then a scatter in
// which every item's destination is fixed by (digit offset, tile, thread, item)
// alone.  Used to bin atoms into cells: keys are cell ids, values the atoms'
// caller-order indices, so equal keys keep ascending index order (the order
// plugins/linkedcell.cc:244-266 threads atoms into each cell).
#pragma once

#include "common.hpp"

namespace lpmdx {

constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kDigits = 16;

__device__ __forceinline__ void digit_counts(const uint32_t* keys, long n, long base, int shift,
                                             uint32_t (&c)[kDigits]) {
#pragma unroll
  for (int b = 0; b < kDigits; ++b)
    c[b] = 0;
  for (int it = 0; it < kSortItems; ++it) {
    long k = base + it;
    if (k < n) {
      uint32_t d = (keys[k] >> shift) & (kDigits - 1);
#pragma unroll
      for (int b = 0; b < kDigits; ++b)
        c[b] += (d == (uint32_t)b) ? 1u : 0u;
    }
  }
}

// hist[b * ntiles + tile] = number of keys with digit b in the tile.
static __global__ __launch_bounds__(kBlock) void k_sort_count(const uint32_t* __restrict__ keys, long n,
                                                       int shift, uint32_t* __restrict__ hist,
                                                       int ntiles) {
  __shared__ uint32_t sh[kDigits][kBlock / kWave];
  const long base = (long)blockIdx.x * kSortTile + (long)threadIdx.x * kSortItems;
  uint32_t c[kDigits];
  digit_counts(keys, n, base, shift, c);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int b = 0; b < kDigits; ++b) {
    uint32_t s = wave_sum(c[b]);
    if (lane == 0)
      sh[b][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < kDigits) {
    uint32_t s = 0;
    for (int q = 0; q < kBlock / kWave; ++q)
      s += sh[threadIdx.x][q];
    hist[(long)threadIdx.x * ntiles + blockIdx.x] = s;
  }
}

// In-place exclusive scan of m uint32 values by one 1024-thread block.
static __global__ __launch_bounds__(1024) void k_scan_excl(uint32_t* __restrict__ data, long m) {
  __shared__ uint32_t sh[1024];
  const long per = (m + 1023) / 1024;
  const long lo = threadIdx.x * per;
  const long hi = lo + per < m ? lo + per : m;
  uint32_t s = 0;
  for (long k = lo; k < hi; ++k)
    s += data[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 partial sums
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t v = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? sh[threadIdx.x - 1] : 0u;
  for (long k = lo; k < hi; ++k) {
    uint32_t v = data[k];
    data[k] = run;
    run += v;
  }
}

static __global__ __launch_bounds__(kBlock) void k_sort_scatter(const uint32_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals, long n,
                                                         int shift,
                                                         const uint32_t* __restrict__ offs,
                                                         int ntiles, uint32_t* __restrict__ okeys,
                                                         uint32_t* __restrict__ ovals) {
  __shared__ uint32_t sh[kDigits][kBlock / kWave];
  const long base = (long)blockIdx.x * kSortTile + (long)threadIdx.x * kSortItems;
  uint32_t c[kDigits];
  digit_counts(keys, n, base, shift, c);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t pos[kDigits];
#pragma unroll
  for (int b = 0; b < kDigits; ++b) {
    // inclusive wave scan of this digit's per-thread counts
    uint32_t v = c[b];
    for (int o = 1; o < kWave; o <<= 1) {
      uint32_t u = __shfl_up(v, o, kWave);
      if (lane >= o)
        v += u;
    }
    if (lane == kWave - 1)
      sh[b][w] = v;
    pos[b] = v - c[b]; // exclusive within the wave
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < kDigits; ++b) {
    uint32_t before = 0;
    for (int q = 0; q < w; ++q)
      before += sh[b][q];
    pos[b] += before + offs[(long)b * ntiles + blockIdx.x];
  }
  for (int it = 0; it < kSortItems; ++it) {
    long k = base + it;
    if (k < n) {
      uint32_t key = keys[k];
      uint32_t d = (key >> shift) & (kDigits - 1);
      uint32_t p = 0;
#pragma unroll
      for (int b = 0; b < kDigits; ++b) {
        if (d == (uint32_t)b) {
          p = pos[b];
          pos[b] = p + 1;
        }
      }
      okeys[p] = key;
      ovals[p] = vals[k];
    }
  }
}

} // namespace lpmdx

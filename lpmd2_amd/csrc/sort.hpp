// Stable LSD radix sort of (key, value) pairs with 8-bit digits.  A chunk of
// kSortChunk consecutive items belongs to one workgroup, read as kSortItems
// coalesced tiles of kBlock items.  Per pass:
//   k_radix_count    per-chunk digit histogram, digit-major: hist[d * nch + j]
//   k_radix_rows     exclusive scan of each digit's row over the chunks, in
//                    place, and the row total into tot[d]
//   k_radix_scatter  every item to (keys of smaller digits) + (its digit's
//                    row offset) + (its rank among the equal digits before it
//                    in the chunk); the rank inside a wavefront comes from
//                    eight ballots, across the wavefronts of a tile from LDS
//                    prefixes, so equal keys keep their input order.
// Every destination is fixed by the input alone; the histogram counts are
// order-free.  Used to bin atoms into cells: keys are cell ids in caller order
// and equal keys keep ascending caller index, the order
// plugins/linkedcell.cc:244-266 threads atoms into each cell.
#pragma once

#include "common.hpp"

namespace lpmdx {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortItems = 16;                  // tiles (items per thread) of a sort chunk
constexpr int kSortChunk = kBlock * kSortItems; // items per sort chunk (one workgroup)
constexpr int kSortWaves = kBlock / kWave;
static_assert(kBlock == kRadix, "one thread per digit in the sort kernels");

// digit passes for keys in [0, kmax]
inline int radix_passes(uint32_t kmax) {
  int bits = 0;
  while (bits < 32 && (kmax >> bits))
    ++bits;
  return bits > 0 ? (bits + kRadixBits - 1) / kRadixBits : 1;
}

// LDS of one scatter workgroup: run[d] = destination of the next item of
// digit d; wcnt[w][d] = items of digit d in wavefront w of the current tile.
struct RankLds {
  uint32_t run[kRadix];
  uint32_t wcnt[kSortWaves][kRadix];
};

// Destination of this thread's item in the current tile of kBlock items
// (thread t holds item t; invalid items take no slot) and advance of run[].
// Every thread of the workgroup must call it (four barriers; the last keeps
// the next tile from clearing wcnt before every thread has read it).
__device__ __forceinline__ uint32_t rank_tile(bool valid, uint32_t d, RankLds& L) {
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  // lanes of this wavefront holding the same digit
  unsigned long long peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const unsigned long long bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
#pragma unroll
  for (int q = 0; q < kSortWaves; ++q)
    L.wcnt[q][t] = 0u;
  __syncthreads();
  if (valid && __ffsll((unsigned long long)peers) - 1 == lane)
    L.wcnt[w][d] = (uint32_t)__popcll(peers);
  __syncthreads();
  {
    // digit t: start of each wavefront's run, then the next tile's start
    uint32_t pre = L.run[t];
#pragma unroll
    for (int q = 0; q < kSortWaves; ++q) {
      const uint32_t c = L.wcnt[q][t];
      L.wcnt[q][t] = pre;
      pre += c;
    }
    L.run[t] = pre;
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t pos = L.wcnt[w][d] + (uint32_t)__popcll(peers & below);
  __syncthreads();
  return pos;
}

// hist[d * nch + j] = items of chunk j whose digit at `shift` is d.
static __global__ __launch_bounds__(kBlock) void k_radix_count(const uint32_t* __restrict__ keys, long n,
                                                        int shift, uint32_t* __restrict__ hist, int nch) {
  __shared__ uint32_t h[kRadix];
  const int t = threadIdx.x, j = blockIdx.x;
  h[t] = 0u;
  __syncthreads();
  const long base = (long)j * kSortChunk + t;
#pragma unroll 4
  for (int u = 0; u < kSortItems; ++u) {
    const long i = base + (long)u * kBlock;
    if (i < n)
      atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u); // a count: order-free
  }
  __syncthreads();
  hist[(long)t * nch + j] = h[t];
}

// One wavefront per digit row: exclusive scan of hist[d * nch ..][0, nch) in
// place (lane l takes a contiguous segment), the row total into tot[d].
static __global__ __launch_bounds__(kBlock) void k_radix_rows(uint32_t* __restrict__ hist, int nch,
                                                       uint32_t* __restrict__ tot) {
  const int lane = threadIdx.x & (kWave - 1);
  const int d = blockIdx.x * kSortWaves + threadIdx.x / kWave;
  if (d >= kRadix)
    return; // uniform over the wavefront
  uint32_t* row = hist + (long)d * nch;
  const int per = (nch + kWave - 1) / kWave;
  const int lo = min(nch, lane * per), hi = min(nch, lo + per);
  uint32_t s = 0;
  for (int k = lo; k < hi; ++k)
    s += row[k];
  // exclusive wave scan of the segment sums
  uint32_t v = s;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, kWave);
    if (lane >= o)
      v += u;
  }
  if (lane == kWave - 1)
    tot[d] = v;
  uint32_t run = v - s;
  for (int k = lo; k < hi; ++k) {
    const uint32_t c = row[k];
    row[k] = run;
    run += c;
  }
}

static __global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals, long n,
                                                          int shift, const uint32_t* __restrict__ hist,
                                                          const uint32_t* __restrict__ tot, int nch,
                                                          uint32_t* __restrict__ okeys,
                                                          uint32_t* __restrict__ ovals) {
  __shared__ RankLds L;
  __shared__ uint32_t wsum[kSortWaves];
  const int t = threadIdx.x, j = blockIdx.x, lane = t & (kWave - 1), w = t / kWave;
  {
    // digit starts: exclusive scan of the row totals over the digits
    const uint32_t c = tot[t];
    uint32_t v = c;
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t u = __shfl_up(v, o, kWave);
      if (lane >= o)
        v += u;
    }
    if (lane == kWave - 1)
      wsum[w] = v;
    __syncthreads();
    uint32_t before = 0;
    for (int q = 0; q < w; ++q)
      before += wsum[q];
    L.run[t] = before + v - c + hist[(long)t * nch + j];
  }
  // every load of the chunk issued before the first rank
  uint32_t kk[kSortItems], vv[kSortItems];
  const long c0 = (long)j * kSortChunk + t;
#pragma unroll
  for (int u = 0; u < kSortItems; ++u) {
    const long i = c0 + (long)u * kBlock;
    kk[u] = i < n ? keys[i] : 0u;
    vv[u] = i < n ? vals[i] : 0u;
  }
#pragma unroll
  for (int u = 0; u < kSortItems; ++u) {
    if ((long)j * kSortChunk + (long)u * kBlock >= n)
      break; // uniform over the workgroup
    const bool valid = c0 + (long)u * kBlock < n;
    const uint32_t pos = rank_tile(valid, (kk[u] >> shift) & (kRadix - 1), L);
    if (valid && pos < (uint32_t)n) {
      okeys[pos] = kk[u];
      ovals[pos] = vv[u];
    }
  }
}

} // namespace lpmdx

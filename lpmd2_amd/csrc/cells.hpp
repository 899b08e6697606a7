// Cell binning and neighbour-list build.
//
// Two grids:
//  * kLinkedCell reproduces plugins/linkedcell.cc: the reference's subdivision
//    (auto mode or explicit nx ny nz), its binning (Fractional, floor, WrapAround,
//    :249-255) and its (2*side+1)^3 stencil in the same nesting order with the
//    same wrap (:215-241), duplicated cells included.  Lists built on it are the
//    reference's lists, in the reference's order.
//  * kExact bins with cells no smaller than the list radius and visits each
//    distinct cell of the 27-cell neighbourhood once, so that the list holds
//    exactly the minimum-image pairs inside the radius (the set
//    plugins/verletlist.cc and plugins/minimumimage.cc deliver).
// Lists are full lists in a tiled, chunked layout.  A tile holds the T = 64/lpa
// atoms one wavefront serves; each atom's entries are stored in chunks of 8
// consecutive words (32 bytes, one L2 sector):
//   entry k of atom i at  tile(i) * T * cap + (k / 8) * T * 8 + (i mod T) * 8 + (k mod 8).
// The build gathers a chunk in LDS and writes it whole, so no sector is ever
// written partially; the lpa lanes serving an atom read consecutive words
// of its chunk, and a tile's list is one contiguous block.
#pragma once

#include "common.hpp"

namespace lpmdx {

enum GridMode { kLinkedCell = 0, kExact = 1 };

struct Grid {
  int nx, ny, nz;
  int side;  // linkedcell stencil half-width
  int mode;  // GridMode
  int ncell; // cell ids: nx*ny*nz, or nbx*nby*nbz*brick^3 on a bricked grid;
             // key == ncell marks an atom outside every cell (linkedcell)
  int brick; // cells per brick edge on a bricked exact grid, 0 otherwise
  int nbx, nby, nbz; // bricks per axis (bricked grids)
  int div;   // exact grids: cells per list radius along each axis (0 or 1: cells
             // of at least the radius, 27-cell stencil; 2: at least half of it,
             // rows of a 5 x 5 x 5 stencil trimmed per atom; unbricked only)
};

// Cell id of cell (i, j, k) (0 <= i < nx, ...).  The plain order is x-fastest.
// A bricked grid numbers cells brick by brick (bricks x-fastest, the cells of
// a brick x-fastest), so the atoms of one brick are one contiguous range of
// the cell-sorted order; ids of cells past the grid edge in a partial brick
// hold no atoms.
__host__ __device__ __forceinline__ int cell_id(const Grid& g, int i, int j, int k) {
  if (g.brick <= 0)
    return (k * g.ny + j) * g.nx + i;
  const int B = g.brick;
  const int bi = i / B, bj = j / B, bk = k / B;
  const int b = (bk * g.nby + bj) * g.nbx + bi;
  return ((b * B + (k - bk * B)) * B + (j - bj * B)) * B + (i - bi * B);
}

__host__ __device__ __forceinline__ void cell_coords(const Grid& g, int id, int& i, int& j, int& k) {
  if (g.brick <= 0) {
    i = id % g.nx;
    j = (id / g.nx) % g.ny;
    k = id / (g.nx * g.ny);
    return;
  }
  const int B = g.brick, B3 = B * B * B;
  const int b = id / B3, r = id - b * B3;
  i = (b % g.nbx) * B + r % B;
  j = ((b / g.nbx) % g.nby) * B + (r / B) % B;
  k = (b / (g.nbx * g.nby)) * B + r / (B * B);
}

// WrapAround, plugins/linkedcell.cc:9-15 (i == -n maps to n, as there).
__device__ __forceinline__ int wrap_around(int i, int n) {
  if (i < 0)
    i += ((-i / n) + 1) * n;
  else if (i > n - 1)
    i -= (i / n) * n;
  return i;
}

__device__ __forceinline__ int pmod(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

__device__ __forceinline__ int f2i(double v) {
  // saturating conversion (v_cvt_i32_f64 semantics), NaN -> 0
  if (!(v == v))
    return 0;
  if (v >= 2147483647.0)
    return 2147483647;
  if (v <= -2147483648.0)
    return -2147483647 - 1;
  return (int)v;
}

__device__ __forceinline__ uint32_t cell_key(const Grid& g, const Box& b, const double* p) {
  // Fractional (orthogonalcell.h:90-94) divides by the cell edge
  double f0 = p[0] / b.L[0], f1 = p[1] / b.L[1], f2 = p[2] / b.L[2];
  int i = f2i(floor(g.nx * f0)), j = f2i(floor(g.ny * f1)), k = f2i(floor(g.nz * f2));
  if (g.mode == kLinkedCell) {
    i = wrap_around(i, g.nx);
    j = wrap_around(j, g.ny);
    k = wrap_around(k, g.nz);
    long q = (long)k * (g.nx * g.ny) + (long)j * g.nx + i;
    return (q >= 0 && q < g.ncell) ? (uint32_t)q : (uint32_t)g.ncell;
  }
  i = pmod(i, g.nx);
  j = pmod(j, g.ny);
  k = pmod(k, g.nz);
  return (uint32_t)cell_id(g, i, j, k);
}

// keys[perm[i]] = cell of the atom in slot i, vals[perm[i]] = i.  Presenting
// atoms to the stable sort in caller order keeps ascending caller index inside
// each cell; the sorted values are then the atoms' current slots.
static __global__ __launch_bounds__(kBlock) void k_cell_keys(const double* __restrict__ x,
                                                      const int* __restrict__ perm, long n, Grid g,
                                                      Box b, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const int o = perm[i];
  keys[o] = cell_key(g, b, x + 3 * i);
  vals[o] = (uint32_t)i;
}

static __global__ __launch_bounds__(kBlock) void k_fill_int(int* __restrict__ p, long m, int v) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < m)
    p[i] = v;
}

// cb[c] = first sorted position of cell c, for c = 0 .. ncell + 1 (ncell + 2
// entries; cb[ncell + 1] = n): the atoms of cell c are [cb[c], cb[c + 1]),
// empty cells included.  n + 1 threads; thread k writes the cells after the
// previous atom's cell up to its own.
static __global__ __launch_bounds__(kBlock) void k_cell_begin(const uint32_t* __restrict__ skeys, long n,
                                                       int ncell, int* __restrict__ cb) {
  const long k = (long)blockIdx.x * kBlock + threadIdx.x;
  if (k > n)
    return;
  const long lo = k == 0 ? 0 : (long)skeys[k - 1] + 1;
  const long hi = k == n ? (long)ncell + 1 : (long)skeys[k];
  for (long c = lo; c <= hi; ++c)
    cb[c] = (int)k;
}

// Largest number of atoms in one cell (lcbinary wants at most one).
static __global__ __launch_bounds__(kBlock) void k_cell_occupancy(const int* __restrict__ cb, int ncell,
                                                           unsigned int* __restrict__ out) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  unsigned int m = c < ncell ? (unsigned int)(cb[c + 1] - cb[c]) : 0u;
  m = wave_max(m);
  if ((threadIdx.x & (kWave - 1)) == 0 && m > 0)
    atomic_max_sparse(out, m);
}

// ---------------------------------------------------------------------------
// Brick halos.  On a bricked grid a force workgroup serves one brick of
// brick^3 cells.  Its atoms' neighbours lie in the (brick + 2)^3 block of
// cells around it (with periodic wrap; on a small grid one cell may stand at
// two places of the block): the brick's halo.  The workgroup copies the halo's
// atoms into LDS once and the neighbour lists store each neighbour as its slot
// there: halo cells in block order (x fastest), a cell's atoms in sorted order.
constexpr int kMaxBrick = 3;
constexpr int kMaxHaloCells = (kMaxBrick + 2) * (kMaxBrick + 2) * (kMaxBrick + 2);
constexpr int kHaloThreads = 128; // k_brick_halo: one thread per halo cell
static_assert(kMaxHaloCells <= kHaloThreads, "one thread per halo cell");

__host__ __device__ __forceinline__ int halo_cells(int brick) { return (brick + 2) * (brick + 2) * (brick + 2); }

// Cell id of halo cell h of brick b.
__device__ __forceinline__ int halo_cell(const Grid& g, int b, int h) {
  const int E = g.brick + 2;
  const int hx = h % E, hy = (h / E) % E, hz = h / (E * E);
  const int i = pmod((b % g.nbx) * g.brick + hx - 1, g.nx);
  const int j = pmod(((b / g.nbx) % g.nby) * g.brick + hy - 1, g.ny);
  const int k = pmod((b / (g.nbx * g.nby)) * g.brick + hz - 1, g.nz);
  return cell_id(g, i, j, k);
}

// Halo table: per brick, halo_cells + 1 entries of int2: (first slot, first
// sorted atom) of every halo cell, then (halo atoms, atoms of the brick).
// One workgroup of kHaloThreads per brick; the largest halo and the largest
// brick go to halo_max and brick_max.
static __global__ __launch_bounds__(kHaloThreads) void k_brick_halo(const int* __restrict__ cb, Grid g,
                                                              int2* __restrict__ out,
                                                              unsigned int* __restrict__ halo_max,
                                                              unsigned int* __restrict__ brick_max) {
  __shared__ int wsum[kHaloThreads / kWave];
  const int b = blockIdx.x, h = threadIdx.x, H = halo_cells(g.brick);
  int first = 0, count = 0;
  if (h < H) {
    const int p = halo_cell(g, b, h);
    first = cb[p];
    count = cb[p + 1] - first;
  }
  // inclusive scan of count: within each wavefront, then across the two
  const int lane = h & (kWave - 1);
  int incl = count;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int t = __shfl_up(incl, o, kWave);
    if (lane >= o)
      incl += t;
  }
  if (lane == kWave - 1)
    wsum[h / kWave] = incl;
  __syncthreads();
  for (int w = 0; w < h / kWave; ++w)
    incl += wsum[w];
  int2* row = out + (long)b * (H + 1);
  if (h < H)
    row[h] = make_int2(incl - count, first);
  if (h == H - 1) {
    const int B3 = g.brick * g.brick * g.brick;
    const int atoms = cb[(b + 1) * B3] - cb[b * B3];
    row[H] = make_int2(incl, atoms);
    atomic_max_sparse(halo_max, (unsigned int)incl);
    atomic_max_sparse(brick_max, (unsigned int)atoms);
  }
}

// FP32 copy of a position folded into [0, L) (in FP64 first), the frame the
// periodic binning uses.
__device__ __forceinline__ float4 fold_f32(const double* x, const Box& b) {
  float w[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double v = x[q];
    v -= b.L[q] * floor(v / b.L[q]);
    w[q] = (float)v;
  }
  return make_float4(w[0], w[1], w[2], 0.0f);
}

// Reorders the per-atom state into the new (cell-sorted) order.  order[k] is
// the current slot of the atom placed at k; perm maps slots to caller indices.
constexpr int kPermuteExtra = 6;

struct PermuteArgs {
  const uint32_t* order;
  const int* perm_old;
  const double* x;
  const double* v;
  const double* a;
  const int* z;
  const double* ffm;
  const double* mass;
  double* x2;
  double* v2;
  double* a2;
  int* z2;
  double* ffm2;
  double* mass2;
  int* perm_new;
  // up to kPermuteExtra more per-atom 3-vectors moved along (null: unused),
  // e.g. the leapfrog half-step velocities and the Beeman history
  const double* ex[kPermuteExtra];
  double* ex2[kPermuteExtra];
  long n;
  // per-atom fixed-atom flags moved along (null: none)
  const int* fix;
  int* fix2;
  // the list build's inputs written on the way, or null: the build positions
  // (a copy of x2) and the FP32 copy of x2 folded into the box (k_to_f32's)
  double* xb;
  float4* xf;
  Box b;
};

// Every load is issued before the first store (the output arrays are distinct
// from the inputs, which the compiler cannot know).
static __global__ __launch_bounds__(kBlock) void k_permute(PermuteArgs A) {
  const long k = (long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= A.n)
    return;
  const long s = A.order[k];
  double x[3], v[3], a[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    x[q] = A.x[3 * s + q];
    v[q] = A.v[3 * s + q];
    a[q] = A.a[3 * s + q];
  }
  const int z = A.z[s], o = A.perm_old[s];
  const int fx = A.fix ? A.fix[s] : 0;
  const double ffm = A.ffm[s], mass = A.mass[s];
  double h[kPermuteExtra][3];
#pragma unroll
  for (int e = 0; e < kPermuteExtra; ++e)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      h[e][q] = A.ex[e] ? A.ex[e][3 * s + q] : 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    A.x2[3 * k + q] = x[q];
    A.v2[3 * k + q] = v[q];
    A.a2[3 * k + q] = a[q];
  }
  if (A.xb) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
      A.xb[3 * k + q] = x[q];
  }
  if (A.xf)
    A.xf[k] = fold_f32(x, A.b);
  A.z2[k] = z;
  A.ffm2[k] = ffm;
  A.mass2[k] = mass;
  A.perm_new[k] = o;
  if (A.fix)
    A.fix2[k] = fx;
#pragma unroll
  for (int e = 0; e < kPermuteExtra; ++e)
    if (A.ex[e]) {
#pragma unroll
      for (int q = 0; q < 3; ++q)
        A.ex2[e][3 * k + q] = h[e][q];
    }
}

constexpr int kChunk = 8;

__host__ __device__ __forceinline__ long list_index(long i, int k, int lpa, int cap) {
  const int T = kWave / lpa;
  return (i / T) * (long)T * cap + (long)(k / kChunk) * (T * kChunk) + (i % T) * kChunk +
         (k % kChunk);
}

struct BuildArgs {
  const double* x;
  const float4* xf; // FP32 copy of x (w unused) by k_to_f32, then kGroup - 1 entries of padding
  const uint32_t* skeys;
  const int* cb; // cell starts (k_cell_begin)
  const int2* bhalo; // brick halo table (k_brick_halo), or null: global indices
  int* nbr;
  int* cnt;
  unsigned int* max_count;
  unsigned long long* sum_count; // one counter, or kSumSlots slots (sum_spread)
  int sum_spread;                // 0: one counter; 1: spread slots (atomic_add_spread)
  Grid g;
  Box b;
  float Lf[3], halff[3];
  double rl2; // list radius squared
  float rl2f; // admission radius^2 in FP32: the list radius widened by the FP32 rounding bound
  long n;
  long stride; // n rounded up to a whole number of tiles
  int cap;     // entries per atom
  int lpa;     // lanes per atom of the consuming kernels
  int path;    // kPathExact64 / kPathMin32 / kPathShift32
};

// Build paths.  The FP32 paths admit candidates whose FP32 distance is inside a
// radius widened by the rounding bound of FP32 positions, so the list is a
// superset of the pairs inside the list radius; every consumer re-tests
// r2 < rc^2 in FP64 exactly as the reference does.
enum BuildPath {
  kPathExact64 = 0, // FP64 distance with the reference's minimum image (small boxes)
  kPathMin32 = 1,   // FP32 distance with a per-pair minimum image (linkedcell stencils)
  kPathShift32 = 2  // FP32 distance, one periodic shift per stencil cell (>= 3 cells per axis)
};

// FP32 copy of the positions folded into [0, L) (in FP64 first), the frame the
// periodic binning uses: an atom at x == L (FittedInside keeps it) or outside
// the box is binned by floor(n x / L) mod n and compared at its folded place.
static __global__ __launch_bounds__(kBlock) void k_to_f32(const double* __restrict__ x, long n, Box b,
                                                   float4* __restrict__ xf) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const double p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  xf[i] = fold_f32(p, b);
}

// Zeroes a build's counters: max_count, pad0 and sum_count (16 bytes from
// max_count), with halo_max and brick_max behind them when `halo`, and the
// spread sum slots (one launch in place of three memsets).
static __global__ __launch_bounds__(kBlock) void k_build_reset(unsigned int* __restrict__ ctrl_words, int halo,
                                                               unsigned long long* __restrict__ slots, int m) {
  const int t = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (t < (halo ? 6 : 4))
    ctrl_words[t] = 0u;
  if (t < m)
    slots[t] = 0ull;
}

// One atom's list being written.  Accepted entries wait in a ring of kRing
// slots in LDS (the thread's own column of a kRing x kBlock array, so a
// wavefront's writes hit 64 distinct banks) and leave for HBM a whole chunk
// (two 16-byte stores) at a time.  A candidate is written to the ring slot
// after the last entry whether or not it is accepted (acceptance only advances
// the count), so the candidate loop has no branch per candidate.  drain() runs
// after every kGroup candidates: with fewer than kChunk entries pending after
// it and at most kGroup accepted before the next, at most kRing - 1 entries are
// ever pending and the ring never overwrites one.
constexpr int kRing = 16;
constexpr int kGroup = 8;
static_assert(kRing >= kChunk + kGroup - 1 && (kRing & (kRing - 1)) == 0, "ring too small");

struct Pending {
  int* col;    // entry slot u at col[u * kBlock]
  long base;   // list_index(i, 0, lpa, cap): the atom's first chunk
  long cstep;  // distance between the atom's consecutive chunks
  int c;       // entries so far
  int f;       // entries already stored (a multiple of kChunk)
};

__device__ __forceinline__ Pending make_pending(const BuildArgs& A, long i, int* ring) {
  Pending P;
  P.col = ring + threadIdx.x;
  // lpa is a power of two, so T = kWave / lpa and i / T is a shift
  const int sh = __builtin_ctz(kWave / A.lpa);
  const long T = 1L << sh;
  P.base = (i >> sh) * T * A.cap + (i & (T - 1)) * kChunk;
  P.cstep = T * kChunk;
  P.c = 0;
  P.f = 0;
  return P;
}

__device__ __forceinline__ void push(Pending& P, int k, bool take) {
  P.col[(P.c & (kRing - 1)) * kBlock] = k;
  P.c += take ? 1 : 0;
}

// store the oldest pending chunk once it is whole (or, at the end, partly filled)
__device__ __forceinline__ void drain(const BuildArgs& A, Pending& P, bool last) {
  if (P.c - P.f >= kChunk || (last && P.c > P.f)) {
    if (P.f < A.cap) {
      const int* s = P.col + (P.f & (kRing - 1)) * kBlock;
      int4* dst = reinterpret_cast<int4*>(A.nbr + P.base + (long)(P.f / kChunk) * P.cstep);
      dst[0] = make_int4(s[0], s[kBlock], s[2 * kBlock], s[3 * kBlock]);
      dst[1] = make_int4(s[4 * kBlock], s[5 * kBlock], s[6 * kBlock], s[7 * kBlock]);
    }
    P.f += kChunk;
  }
}

__device__ __forceinline__ float wrap1f(float d, float L, float h) {
  const float w = (d >= h) ? -L : ((d < -h) ? L : 0.0f);
  return d + w;
}

// FP32 candidates [lo, hi) against atom i (the FP32 paths).  A group of
// kGroup positions is loaded before any is tested: xf holds kGroup - 1 entries
// of padding past the last atom, so the loads of a group read xf[k0 .. k0 + 7]
// at fixed offsets from one address, and the entries at or past hi (the next
// cell's atoms, or padding) are never accepted.  SELF: the cell holds atom i
// itself, which is skipped.  The distance is formed with fused multiply-adds:
// its rounding stays far inside the admission margin of rl2f.
template <bool SELF>
__device__ __forceinline__ void scan_cell_f32(const BuildArgs& A, int lo, int hi, int ii, float sx, float sy,
                                              float sz, Pending& P, int shift, bool wrap) {
  const int m = hi - lo;
  for (int k0 = 0; k0 < m; k0 += kGroup) {
    const float4* q4 = A.xf + lo + k0;
    float4 q[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u)
      q[u] = q4[u];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      // kPathShift32: (sx, sy, sz) is atom i moved by minus this cell's
      // periodic image shift; kPathMin32: a per-pair minimum image
      float dx = q[u].x - sx, dy = q[u].y - sy, dz = q[u].z - sz;
      if (wrap) {
        dx = wrap1f(dx, A.Lf[0], A.halff[0]);
        dy = wrap1f(dy, A.Lf[1], A.halff[1]);
        dz = wrap1f(dz, A.Lf[2], A.halff[2]);
      }
      const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
      const int k = lo + k0 + u;
      bool take = (k0 + u < m) & (d2 < A.rl2f);
      if (SELF)
        take &= k != ii;
      push(P, k + shift, take);
    }
    drain(A, P, false);
  }
}

// candidates of cell p against atom i; entries are written as k + shift (the
// halo slot of sorted atom k, or k itself for global indices)
__device__ __forceinline__ void scan_cell(const BuildArgs& A, int p, long i, const double* xi,
                                          float sx, float sy, float sz, Pending& P, int shift) {
  const int lo = A.cb[p], hi = A.cb[p + 1];
  const int ii = (int)i;
  if (A.path == kPathExact64) {
    // candidates kGroup at a time; the index is clamped to the cell's last
    // atom so every load stays inside the cell, and only k < hi is accepted
    for (int k0 = lo; k0 < hi; k0 += kGroup) {
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int k = min(k0 + u, hi - 1);
        double dx = A.x[3 * (long)k] - xi[0], dy = A.x[3 * (long)k + 1] - xi[1],
               dz = A.x[3 * (long)k + 2] - xi[2];
        min_image(A.b, dx, dy, dz);
        push(P, k + shift, k0 + u < hi && k != ii && sqmod(dx, dy, dz) < A.rl2);
      }
      drain(A, P, false);
    }
  } else {
    // atom i lies in its own cell only: [lo, hi) holds it iff lo <= i < hi
    const bool wrap = A.path == kPathMin32;
    if (lo <= ii && ii < hi)
      scan_cell_f32<true>(A, lo, hi, ii, sx, sy, sz, P, shift, wrap);
    else
      scan_cell_f32<false>(A, lo, hi, ii, sx, sy, sz, P, shift, wrap);
  }
}

// Distance along one axis from a coordinate u inside its cell [0, h) to the
// cell d cells away (0 for its own cell).
__device__ __forceinline__ float cell_gap(int d, float u, float h) {
  return d > 0 ? (float)(d - 1) * h + (h - u) : (d < 0 ? (float)(-d - 1) * h + u : 0.0f);
}

// The stencil of a grid of half-radius cells (g.div == 2, kPathShift32): the
// 25 rows (dz, dy) of five cells along x around atom i's cell, each trimmed
// to the cells within reach of the atom (its offset inside its cell fixes
// how far the row's y and z gaps leave to go along x), and each scanned as
// one contiguous range of sorted atoms (cells are numbered x-fastest), split
// where the row wraps around the x faces.  The reach is the admission
// radius widened by a relative 1e-5 and 1e-5 of a cell edge, so the FP32
// trimming never drops a cell the admission test could take.  Entries come
// row by row, in sorted order along a row.
__device__ __forceinline__ void scan_rows(const BuildArgs& A, long i, int ci, int cj, int ck, float4 fi, Pending& P) {
  const Grid& g = A.g;
  const int ii = (int)i;
  const float hx = A.Lf[0] / (float)g.nx, hy = A.Lf[1] / (float)g.ny, hz = A.Lf[2] / (float)g.nz;
  const float reach = sqrtf(A.rl2f) * (1.0f + 1e-5f) + 1e-5f * fmaxf(hx, fmaxf(hy, hz));
  const float reach2 = reach * reach;
  // the atom's offset inside its cell, held inside [0, h]
  const float ux = fminf(fmaxf(fi.x - (float)ci * hx, 0.0f), hx);
  const float uy = fminf(fmaxf(fi.y - (float)cj * hy, 0.0f), hy);
  const float uz = fminf(fmaxf(fi.z - (float)ck * hz, 0.0f), hz);
  for (int dz = -2; dz <= 2; ++dz) {
    const float gz = cell_gap(dz, uz, hz);
    const float rz2 = reach2 - gz * gz;
    if (!(rz2 > 0.0f))
      continue;
    const int rz = ck + dz;
    const float shz = rz < 0 ? -A.Lf[2] : (rz >= g.nz ? A.Lf[2] : 0.0f);
    const int z2 = pmod(rz, g.nz);
    for (int dy = -2; dy <= 2; ++dy) {
      const float gy = cell_gap(dy, uy, hy);
      const float ry2 = rz2 - gy * gy;
      if (!(ry2 > 0.0f))
        continue;
      const float rx = sqrtf(ry2);
      const int lo = cell_gap(-2, ux, hx) < rx ? -2 : (cell_gap(-1, ux, hx) < rx ? -1 : 0);
      const int hi = cell_gap(2, ux, hx) < rx ? 2 : (cell_gap(1, ux, hx) < rx ? 1 : 0);
      const int ry = cj + dy;
      const float shy = ry < 0 ? -A.Lf[1] : (ry >= g.ny ? A.Lf[1] : 0.0f);
      const int z1 = pmod(ry, g.ny);
      const int row = (z2 * g.ny + z1) * g.nx;
      // up to three runs: cells past the low face, inside, past the high face
      for (int part = 0; part < 3; ++part) {
        int x0 = ci + lo, x1 = ci + hi, wrap = 0;
        float shx = 0.0f;
        if (part == 0) {
          x1 = min(x1, -1);
          wrap = g.nx;
          shx = -A.Lf[0];
        } else if (part == 1) {
          x0 = max(x0, 0);
          x1 = min(x1, g.nx - 1);
        } else {
          x0 = max(x0, g.nx);
          wrap = -g.nx;
          shx = A.Lf[0];
        }
        if (x0 > x1)
          continue;
        const int b = A.cb[row + x0 + wrap], e = A.cb[row + x1 + wrap + 1];
        const float sx = fi.x - shx, sy = fi.y - shy, sz = fi.z - shz;
        if (b <= ii && ii < e)
          scan_cell_f32<true>(A, b, e, ii, sx, sy, sz, P, 0, false);
        else
          scan_cell_f32<false>(A, b, e, ii, sx, sy, sz, P, 0, false);
      }
    }
  }
}

// Where atom i's neighbours go in the list: on a bricked grid (A.bhalo) the
// halo slot of its brick, slot = sorted index + shift of the neighbour cell,
// else the sorted index itself.
struct HaloRow {
  const int2* row = nullptr; // the brick's halo table row, or null
  int li = 0, lj = 0, lk = 0; // atom i's cell within its brick
  // shift of the halo cell at offset (dx, dy, dz) from atom i's cell
  __device__ __forceinline__ int shift(int E, int dx, int dy, int dz) const {
    if (!row)
      return 0;
    const int2 q = row[((lk + 1 + dz) * E + (lj + 1 + dy)) * E + (li + 1 + dx)];
    return q.x - q.y;
  }
};

// List of atom i (every lane of a wavefront must call it: it ends with
// wavefront reductions into max_count / sum_count).
// ring: the block's kRing x kBlock LDS array for pending entries.
__device__ __forceinline__ void build_atom(const BuildArgs& A, long i, bool active, int* ring) {
  Pending P = make_pending(A, i, ring);
  if (active) {
    const Grid& g = A.g;
    uint32_t key = A.skeys[i];
    if (key < (uint32_t)g.ncell) {
      const double xi[3] = {A.x[3 * i], A.x[3 * i + 1], A.x[3 * i + 2]};
      const float4 fi = A.xf[i];
      int ci, cj, ck;
      cell_coords(g, (int)key, ci, cj, ck);
      HaloRow hr;
      const int E = g.brick + 2;
      if (A.bhalo) {
        const int B = g.brick;
        hr.row = A.bhalo + (long)(key / (B * B * B)) * (halo_cells(B) + 1);
        hr.li = ci % B;
        hr.lj = cj % B;
        hr.lk = ck % B;
      }
      if (g.mode == kLinkedCell) {
        const int s = g.side;
        for (int dz = -s; dz <= s; ++dz) {
          const int z2 = wrap_around(ck + dz, g.nz);
          for (int dy = -s; dy <= s; ++dy) {
            const int z1 = wrap_around(cj + dy, g.ny);
            for (int dx = -s; dx <= s; ++dx) {
              const int z0 = wrap_around(ci + dx, g.nx);
              long p = (long)z2 * (g.nx * g.ny) + (long)z1 * g.nx + z0;
              // p == ncell only when side >= n (WrapAround's -n case): the
              // reference reads past its cell table there; treated as empty
              if (p < g.ncell)
                scan_cell(A, (int)p, i, xi, fi.x, fi.y, fi.z, P, 0);
            }
          }
        }
      } else if (A.path == kPathShift32 && g.div == 2) {
        scan_rows(A, i, ci, cj, ck, fi, P);
      } else if (A.path == kPathShift32) {
        for (int dz = -1; dz <= 1; ++dz) {
          const int rz = ck + dz;
          const float shz = rz < 0 ? -A.Lf[2] : (rz >= g.nz ? A.Lf[2] : 0.0f);
          const int z2 = pmod(rz, g.nz);
          for (int dy = -1; dy <= 1; ++dy) {
            const int ry = cj + dy;
            const float shy = ry < 0 ? -A.Lf[1] : (ry >= g.ny ? A.Lf[1] : 0.0f);
            const int z1 = pmod(ry, g.ny);
            for (int dx = -1; dx <= 1; ++dx) {
              const int rx = ci + dx;
              const float shx = rx < 0 ? -A.Lf[0] : (rx >= g.nx ? A.Lf[0] : 0.0f);
              const int z0 = pmod(rx, g.nx);
              // image of cell (z0,z1,z2) next to atom i sits at x_j + sh: compare
              // x_j against x_i - sh
              scan_cell(A, cell_id(g, z0, z1, z2), i, xi, fi.x - shx, fi.y - shy, fi.z - shz, P,
                        hr.shift(E, dx, dy, dz));
            }
          }
        }
      } else {
        // distinct cells of the 27-neighbourhood: offsets {-1,0,1}, {0,1} or {0}
        const int oz0 = g.nz >= 3 ? -1 : 0, oy0 = g.ny >= 3 ? -1 : 0, ox0 = g.nx >= 3 ? -1 : 0;
        const int oz1 = g.nz >= 2 ? 1 : 0, oy1 = g.ny >= 2 ? 1 : 0, ox1 = g.nx >= 2 ? 1 : 0;
        for (int dz = oz0; dz <= oz1; ++dz) {
          const int z2 = pmod(ck + dz, g.nz);
          for (int dy = oy0; dy <= oy1; ++dy) {
            const int z1 = pmod(cj + dy, g.ny);
            for (int dx = ox0; dx <= ox1; ++dx) {
              const int z0 = pmod(ci + dx, g.nx);
              scan_cell(A, cell_id(g, z0, z1, z2), i, xi, fi.x, fi.y, fi.z, P, hr.shift(E, dx, dy, dz));
            }
          }
        }
      }
    }
    drain(A, P, true); // the last, partly filled chunk
    A.cnt[i] = P.c;
  }
  const int c = active ? P.c : 0;
  unsigned int m = wave_max((unsigned int)c);
  unsigned long long t = wave_sum((unsigned long long)c);
  if ((threadIdx.x & (kWave - 1)) == 0 && m > 0) {
    // integer max and sum: the result does not depend on the order
    atomic_max_sparse(A.max_count, m);
    if (A.sum_spread)
      atomic_add_spread(A.sum_count, t);
    else
      atomicAdd(A.sum_count, t);
  }
}

static __global__ __launch_bounds__(kBlock) void k_build_list(BuildArgs A) {
  __shared__ int ring[kRing * kBlock];
  const long i = xcd_block() * kBlock + threadIdx.x;
  build_atom(A, i, i < A.n, ring);
}

// k_build_list for an exact grid on the FP32 one-shift-per-cell path (the
// path of every large box), with the path fixed at compile time: without the
// other paths' code the kernel needs fewer registers (83 VGPRs against 97:
// six wavefronts a SIMD instead of five).
static __global__ __launch_bounds__(kBlock) void k_build_list_shift32(BuildArgs A) {
  __shared__ int ring[kRing * kBlock];
  const long i = xcd_block() * kBlock + threadIdx.x;
  A.path = kPathShift32;
  A.g.mode = kExact;
  build_atom(A, i, i < A.n, ring);
}

// Stencil cells [t0, t1) of a linkedcell stencil (t = ((dz + s) S + dy + s) S
// + dx + s, the order build_atom visits them) against atom i: the number of
// candidates accepted, written from list position off on when write is set.
// The test is scan_cell's, so the entries are the same.
template <bool EXACT>
__device__ __forceinline__ int lc_scan_cells(const BuildArgs& A, long i, int t0, int t1, int ci, int cj, int ck,
                                             const double* xi, float4 fi, int off, bool write) {
  const Grid& g = A.g;
  const int s = g.side, S = 2 * s + 1;
  int c = 0;
  for (int t = t0; t < t1; ++t) {
    const int dx = t % S - s, dy = (t / S) % S - s, dz = t / (S * S) - s;
    const int z2 = wrap_around(ck + dz, g.nz), z1 = wrap_around(cj + dy, g.ny), z0 = wrap_around(ci + dx, g.nx);
    const long p = (long)z2 * (g.nx * g.ny) + (long)z1 * g.nx + z0;
    if (p >= g.ncell)
      continue; // WrapAround's -n case: an empty cell, as in build_atom
    const int lo = A.cb[p], hi = A.cb[p + 1];
    for (int k = lo; k < hi; ++k) {
      bool take;
      if (EXACT) {
        double ex = A.x[3 * (long)k] - xi[0], ey = A.x[3 * (long)k + 1] - xi[1], ez = A.x[3 * (long)k + 2] - xi[2];
        min_image(A.b, ex, ey, ez);
        take = sqmod(ex, ey, ez) < A.rl2;
      } else {
        const float4 q = A.xf[k];
        const float fx = wrap1f(q.x - fi.x, A.Lf[0], A.halff[0]);
        const float fy = wrap1f(q.y - fi.y, A.Lf[1], A.halff[1]);
        const float fz = wrap1f(q.z - fi.z, A.Lf[2], A.halff[2]);
        take = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)) < A.rl2f;
      }
      if (take && k != (int)i) {
        if (write && off + c < A.cap)
          A.nbr[list_index(i, off + c, A.lpa, A.cap)] = k;
        ++c;
      }
    }
  }
  return c;
}

// Small systems on a linkedcell stencil: one wavefront per atom.  With few
// atoms the thread-per-atom build leaves most of the GPU idle while each
// thread walks its stencil (125 cells at two cells a side) one dependent
// cell-range load after another.  Here lane l takes stencil cells
// [l c, (l + 1) c) (c = cells / 64, rounded up) in the stencil's order and
// counts what it accepts; a wavefront prefix sum gives each lane its first
// list position, and the lanes scan again and write there.  The list is the
// thread-per-atom build's, entry for entry (stencil order, sorted order
// within a cell).  Grid: one wavefront per atom.
template <bool EXACT>
static __global__ __launch_bounds__(kBlock) void k_build_list_lc_wave(BuildArgs A) {
  const long i = (xcd_block() * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  int total = 0;
  if (i < A.n) { // the whole wavefront, or none of it
    const Grid& g = A.g;
    const uint32_t key = A.skeys[i];
    if (key < (uint32_t)g.ncell) {
      const double xi[3] = {A.x[3 * i], A.x[3 * i + 1], A.x[3 * i + 2]};
      const float4 fi = A.xf[i];
      int ci, cj, ck;
      cell_coords(g, (int)key, ci, cj, ck);
      const int S = 2 * g.side + 1, nst = S * S * S;
      const int per = (nst + kWave - 1) / kWave;
      const int t0 = min(lane * per, nst), t1 = min(t0 + per, nst);
      const int c = lc_scan_cells<EXACT>(A, i, t0, t1, ci, cj, ck, xi, fi, 0, false);
      int inc = c; // inclusive prefix over the lanes
      for (int o = 1; o < kWave; o <<= 1) {
        const int u = __shfl_up(inc, o, kWave);
        if (lane >= o)
          inc += u;
      }
      total = __shfl(inc, kWave - 1, kWave);
      lc_scan_cells<EXACT>(A, i, t0, t1, ci, cj, ck, xi, fi, inc - c, true);
    }
    if (lane == 0) {
      A.cnt[i] = total;
      if (total > 0) {
        atomic_max_sparse(A.max_count, (unsigned int)total);
        atomic_add_spread(A.sum_count, (unsigned long long)total);
      }
    }
  }
}

} // namespace lpmdx

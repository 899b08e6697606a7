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
  int ncell; // nx*ny*nz; key == ncell marks an atom outside every cell
};

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
  return (uint32_t)(((long)k * g.ny + j) * g.nx + i);
}

// keys[perm[i]] = cell of atom i; vals[o] = o.  Presenting atoms to the
// stable sort in caller order keeps ascending caller index inside each cell.
static __global__ __launch_bounds__(kBlock) void k_cell_keys(const double* __restrict__ x,
                                                      const int* __restrict__ perm, long n, Grid g,
                                                      Box b, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  int o = perm[i];
  keys[o] = cell_key(g, b, x + 3 * i);
  vals[i] = (uint32_t)i;
}

static __global__ __launch_bounds__(kBlock) void k_fill_int(int* __restrict__ p, long m, int v) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < m)
    p[i] = v;
}

// [cstart[c], cend[c]) = the sorted positions of cell c (c == ncell: outside).
static __global__ __launch_bounds__(kBlock) void k_cell_ranges(const uint32_t* __restrict__ skeys, long n,
                                                        int* __restrict__ cstart,
                                                        int* __restrict__ cend) {
  long k = (long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n)
    return;
  uint32_t key = skeys[k];
  if (k == 0 || skeys[k - 1] != key)
    cstart[key] = (int)k;
  if (k == n - 1 || skeys[k + 1] != key)
    cend[key] = (int)(k + 1);
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

// ---------------------------------------------------------------------------
// Tile halos.  The atoms of a 64-atom tile (one wavefront of the build) have
// their neighbours in the 27-cell neighbourhoods of the tile's cells; the
// union of those cells, merged into runs of consecutive cell ids, is a few
// contiguous ranges of the cell-sorted atoms: the tile's halo.  The force
// kernels copy a tile's halo into LDS with coalesced loads and read neighbour
// positions there; the list stores each neighbour as its slot in the halo
// (slots run through the ranges in order).  Tiles whose halo is too large,
// spans too many cell rows, or that use the reference linkedcell stencil keep
// global atom indices (total = -1).
constexpr int kHaloTile = 64;    // atoms per tile
constexpr int kHaloRanges = 32;  // most ranges of one halo
constexpr int kHaloCap = 2000;   // most atoms of one staged halo: 32 B each in LDS, with the
                                 // kernels' static reduction scratch, within 64 KiB
constexpr int kHaloRows = 4;     // most cell rows one staged tile may span

// Tile staging is off unless LPMDX_STAGE=1 is set in the environment: on the
// measured workloads the LDS a 64-atom tile's FP64 halo needs (1000-2000
// atoms, 24-32 B each) costs more occupancy than the L1 gathers it saves.
inline bool stage_enabled() {
  static const bool on = [] {
    const char* e = getenv("LPMDX_STAGE");
    return e && e[0] == '1';
  }();
  return on;
}

struct HaloTile {
  int total; // halo atoms, or -1: global indices
  int nr;    // ranges
  int pad[2];
  int4 r[kHaloRanges]; // (first cell, last cell, first sorted atom, first slot)
};

// Range of cell p in D (the first cell run holding p), or -1.
__device__ __forceinline__ int halo_range(const HaloTile* D, int nr, int p) {
  for (int r = 0; r < nr; ++r) {
    const int4 q = D->r[r];
    if (p >= q.x && p <= q.y)
      return r;
  }
  return -1;
}

// Halo of tile t (atoms [64 t, 64 t + 64) of a system with n atoms sorted by
// cell, sorted keys skeys, cell starts cb) on an exact grid.
__device__ inline void make_halo(const uint32_t* __restrict__ skeys, const int* __restrict__ cb, const Grid& g,
                                 long n, long t, int cap, HaloTile* __restrict__ D) {
  const long a0 = t * kHaloTile, a1 = min(n, a0 + kHaloTile);
  int total = -1, nr = 0;
  int2 iv[kHaloRows * 9 * 2];
  int m = 0;
  bool ok = g.mode == kExact && a0 < a1;
  int cf = 0, cl = 0;
  if (ok) {
    cf = (int)skeys[a0];
    cl = (int)skeys[a1 - 1];
    ok = cl < g.ncell && cl / g.nx - cf / g.nx + 1 <= kHaloRows;
  }
  if (ok) {
    // neighbour offsets per axis that name distinct cells: {-1,0,1}, {0,1} or {0}
    const int zy0 = g.nz >= 3 ? -1 : 0, zy1 = g.nz >= 2 ? 1 : 0;
    const int yy0 = g.ny >= 3 ? -1 : 0, yy1 = g.ny >= 2 ? 1 : 0;
    for (int row = cf / g.nx; row <= cl / g.nx; ++row) {
      const int y = row % g.ny, z = row / g.ny;
      const int xa = row == cf / g.nx ? cf % g.nx : 0;
      const int xb = row == cl / g.nx ? cl % g.nx : g.nx - 1;
      // x pieces of the neighbour rows: [xa-1, xb+1] around the ring, or the whole row
      int p0 = 0, p1 = g.nx - 1, q0 = 1, q1 = 0; // second piece empty unless wrapping
      if (g.nx >= 3 && xb - xa + 3 < g.nx) {
        p0 = xa - 1;
        p1 = xb + 1;
        if (p0 < 0) {
          q0 = 0;
          q1 = p1;
          p0 += g.nx;
          p1 = g.nx - 1;
        } else if (p1 >= g.nx) {
          q0 = 0;
          q1 = p1 - g.nx;
          p1 = g.nx - 1;
        }
      }
      for (int dz = zy0; dz <= zy1; ++dz) {
        const int z2 = pmod(z + dz, g.nz);
        for (int dy = yy0; dy <= yy1; ++dy) {
          const int base = (z2 * g.ny + pmod(y + dy, g.ny)) * g.nx;
          iv[m++] = make_int2(base + p0, base + p1);
          if (q0 <= q1)
            iv[m++] = make_int2(base + q0, base + q1);
        }
      }
    }
    // sort the cell intervals by first cell, then merge overlapping / adjacent ones
    for (int a = 1; a < m; ++a) {
      const int2 v = iv[a];
      int b = a - 1;
      while (b >= 0 && iv[b].x > v.x) {
        iv[b + 1] = iv[b];
        --b;
      }
      iv[b + 1] = v;
    }
    total = 0;
    int k = 0;
    while (k < m && ok) {
      int lo = iv[k].x, hi = iv[k].y;
      ++k;
      while (k < m && iv[k].x <= hi + 1) {
        hi = max(hi, iv[k].y);
        ++k;
      }
      const int s0 = cb[lo], s1 = cb[hi + 1];
      if (s1 > s0) {
        if (nr == kHaloRanges) {
          ok = false;
          break;
        }
        D->r[nr++] = make_int4(lo, hi, s0, total);
        total += s1 - s0;
      }
    }
    if (!ok || total > cap)
      total = -1;
  }
  D->total = ok ? total : -1;
  D->nr = ok ? nr : 0;
}

// Halos of tiles [0, ntile) of one system; halo_max = the largest staged halo.
static __global__ __launch_bounds__(kBlock) void k_halo(const uint32_t* __restrict__ skeys,
                                                 const int* __restrict__ cb, Grid g, long n, long ntile,
                                                 int cap, HaloTile* __restrict__ out,
                                                 unsigned int* __restrict__ halo_max) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= ntile)
    return;
  make_halo(skeys, cb, g, n, t, cap, out + t);
  const int total = out[t].total;
  if (total > 0)
    atomicMax(halo_max, (unsigned int)total);
}

// Reorders the per-atom state into the new (cell-sorted) order.  order[k] is
// the caller index of the atom placed at k; iperm maps caller -> old position.
struct PermuteArgs {
  const uint32_t* order;
  const int* iperm_old;
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
  int* iperm_new;
  long n;
};

static __global__ __launch_bounds__(kBlock) void k_permute(PermuteArgs A) {
  long k = (long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= A.n)
    return;
  int o = (int)A.order[k];
  long s = A.iperm_old[o];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    A.x2[3 * k + q] = A.x[3 * s + q];
    A.v2[3 * k + q] = A.v[3 * s + q];
    A.a2[3 * k + q] = A.a[3 * s + q];
  }
  A.z2[k] = A.z[s];
  A.ffm2[k] = A.ffm[s];
  A.mass2[k] = A.mass[s];
  A.perm_new[k] = o;
  A.iperm_new[o] = (int)k;
}

constexpr int kChunk = 8;

__host__ __device__ __forceinline__ long list_index(long i, int k, int lpa, int cap) {
  const int T = kWave / lpa;
  return (i / T) * (long)T * cap + (long)(k / kChunk) * (T * kChunk) + (i % T) * kChunk +
         (k % kChunk);
}

struct BuildArgs {
  const double* x;
  const float4* xf; // FP32 copy of x (w unused), written by k_to_f32 before the build
  const uint32_t* skeys;
  const int* cb; // cell starts (k_cell_begin)
  const HaloTile* halo; // per 64-atom tile, or null: global indices everywhere
  unsigned int* halo_miss; // set when a neighbour cell lies outside its tile's halo
  int* nbr;
  int* cnt;
  unsigned int* max_count;
  unsigned long long* sum_count;
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
  float w[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double v = x[3 * i + q];
    v -= b.L[q] * floor(v / b.L[q]);
    w[q] = (float)v;
  }
  xf[i] = make_float4(w[0], w[1], w[2], 0.0f);
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

// candidates of cell p against atom i; entries are written as k + shift (the
// halo slot of sorted atom k, or k itself for global indices)
// Candidates are taken kGroup at a time; the index is clamped to the cell's
// last atom so every load stays inside the cell, and only k < hi is accepted.
__device__ __forceinline__ void scan_cell(const BuildArgs& A, int p, long i, const double* xi,
                                          float sx, float sy, float sz, Pending& P, int shift) {
  const int lo = A.cb[p], hi = A.cb[p + 1];
  const int ii = (int)i;
  if (A.path == kPathExact64) {
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
    // the group's positions are loaded before any is tested, and the tests
    // combine with & (no short circuit), so the loads are in flight together
    const bool wrap = A.path == kPathMin32;
    for (int k0 = lo; k0 < hi; k0 += kGroup) {
      float4 q[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u)
        q[u] = A.xf[min(k0 + u, hi - 1)];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int k = min(k0 + u, hi - 1);
        // kPathShift32: (sx, sy, sz) is atom i moved by minus this cell's
        // periodic image shift; kPathMin32: a per-pair minimum image
        float dx = q[u].x - sx, dy = q[u].y - sy, dz = q[u].z - sz;
        if (wrap) {
          dx = wrap1f(dx, A.Lf[0], A.halff[0]);
          dy = wrap1f(dy, A.Lf[1], A.halff[1]);
          dz = wrap1f(dz, A.Lf[2], A.halff[2]);
        }
        const bool in = dx * dx + dy * dy + dz * dz < A.rl2f;
        push(P, k + shift, (k0 + u < hi) & (k != ii) & in);
      }
      drain(A, P, false);
    }
  }
}

// Slot shift of cell p for tile halo D (nr ranges): slot = sorted index + shift.
// A non-empty cell outside the halo cannot happen for an exact grid; it is
// reported (halo_miss) and the cell skipped rather than read wrongly.
__device__ __forceinline__ bool cell_shift(const BuildArgs& A, const HaloTile* D, int nr, int p, int& shift) {
  shift = 0;
  if (!D)
    return true;
  if (A.cb[p + 1] == A.cb[p])
    return true;
  const int r = halo_range(D, nr, p);
  if (r < 0) {
    atomicOr(A.halo_miss, 1u);
    return false;
  }
  const int4 q = D->r[r];
  shift = q.w - q.z;
  return true;
}

// List of atom i (every lane of a wavefront must call it: it ends with
// wavefront reductions into max_count / sum_count).
// ring: the block's kRing x kBlock LDS array for pending entries.
__device__ __forceinline__ void build_atom(const BuildArgs& A, long i, bool active, int* ring) {
  Pending P = make_pending(A, i, ring);
  // the tile's halo (the wavefront serves one tile: a uniform descriptor)
  const HaloTile* D = nullptr;
  int nr = 0;
  if (A.halo) {
    const long t = __builtin_amdgcn_readfirstlane((int)(i / kHaloTile));
    // (a wavefront past the last tile has no descriptor)
    if (t < A.stride / kHaloTile && A.halo[t].total >= 0) {
      D = A.halo + t;
      nr = D->nr;
    }
  }
  if (active) {
    const Grid& g = A.g;
    uint32_t key = A.skeys[i];
    if (key < (uint32_t)g.ncell) {
      const double xi[3] = {A.x[3 * i], A.x[3 * i + 1], A.x[3 * i + 2]};
      const float4 fi = A.xf[i];
      const int ci = key % g.nx, cj = (key / g.nx) % g.ny, ck = key / (g.nx * g.ny);
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
              const int p = (z2 * g.ny + z1) * g.nx + z0;
              int shift;
              if (cell_shift(A, D, nr, p, shift))
                scan_cell(A, p, i, xi, fi.x - shx, fi.y - shy, fi.z - shz, P, shift);
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
              const int p = (z2 * g.ny + z1) * g.nx + z0;
              int shift;
              if (cell_shift(A, D, nr, p, shift))
                scan_cell(A, p, i, xi, fi.x, fi.y, fi.z, P, shift);
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
    atomicMax(A.max_count, m);
    atomicAdd(A.sum_count, t);
  }
}

static __global__ __launch_bounds__(kBlock) void k_build_list(BuildArgs A) {
  __shared__ int ring[kRing * kBlock];
  const long i = xcd_block() * kBlock + threadIdx.x;
  build_atom(A, i, i < A.n, ring);
}

} // namespace lpmdx

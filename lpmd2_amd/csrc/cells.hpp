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
// Lists are full lists in a lane-interleaved transposed layout:
//   entry k of atom i at  (k / lpa) * (stride * lpa) + i * lpa + (k % lpa)
// so that the lanes of a wavefront read consecutive words.
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
__global__ __launch_bounds__(kBlock) void k_cell_keys(const double* __restrict__ x,
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

__global__ __launch_bounds__(kBlock) void k_fill_int(int* __restrict__ p, long m, int v) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < m)
    p[i] = v;
}

// [cstart[c], cend[c]) = the sorted positions of cell c (c == ncell: outside).
__global__ __launch_bounds__(kBlock) void k_cell_ranges(const uint32_t* __restrict__ skeys, long n,
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

__global__ __launch_bounds__(kBlock) void k_permute(PermuteArgs A) {
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

struct BuildArgs {
  const double* x;
  const uint32_t* skeys;
  const int* cstart;
  const int* cend;
  int* nbr;
  int* cnt;
  unsigned int* max_count;
  unsigned long long* sum_count;
  Grid g;
  Box b;
  double rl2; // list radius squared (FP64 r2 < rl2, as the managers test)
  long n;
  long stride; // >= n, row length of the transposed list
  int cap;     // entries per atom
  int lpa;     // lanes per atom of the consuming kernels
};

__device__ __forceinline__ void scan_cell(const BuildArgs& A, int p, long i, const double* xi,
                                          int& c) {
  const int lo = A.cstart[p], hi = A.cend[p];
  const long rowlen = A.stride * A.lpa;
  for (int k = lo; k < hi; ++k) {
    if (k == i)
      continue;
    double dx = A.x[3 * k] - xi[0];
    double dy = A.x[3 * k + 1] - xi[1];
    double dz = A.x[3 * k + 2] - xi[2];
    min_image(A.b, dx, dy, dz);
    double r2 = sqmod(dx, dy, dz);
    if (r2 < A.rl2) {
      if (c < A.cap)
        A.nbr[(long)(c / A.lpa) * rowlen + i * A.lpa + (c % A.lpa)] = k;
      ++c;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_build_list(BuildArgs A) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  int c = 0;
  if (i < A.n) {
    const Grid& g = A.g;
    uint32_t key = A.skeys[i];
    if (key < (uint32_t)g.ncell) {
      const double xi[3] = {A.x[3 * i], A.x[3 * i + 1], A.x[3 * i + 2]};
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
                scan_cell(A, (int)p, i, xi, c);
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
              scan_cell(A, (z2 * g.ny + z1) * g.nx + z0, i, xi, c);
            }
          }
        }
      }
    }
    A.cnt[i] = c;
  }
  unsigned int m = wave_max((unsigned int)c);
  unsigned long long t = wave_sum((unsigned long long)c);
  if ((threadIdx.x & (kWave - 1)) == 0 && m > 0) {
    // integer max and sum: the result does not depend on the order
    atomicMax(A.max_count, m);
    atomicAdd(A.sum_count, t);
  }
}

} // namespace lpmdx

// Device side of the batched engine: many independent periodic systems of
// uneven size advanced together.  Every step is the same fixed sequence of
// launches (captured once in a HIP graph and replayed):
//   k_b_control  one thread per system: step bookkeeping, skin check
//   k_b_sort     one workgroup per system that needs a new list: stable
//                in-workgroup radix sort of its atoms into cells, reorder
//   k_b_build    every tile of those systems: neighbour lists
//   k_b_rho      EAM systems: rho pass            (one launch per kind)
//   k_b_force    all active systems: forces + velocity Verlet (per kind)
// The force and build launches sweep a static list of 64-atom tiles covering
// all systems, so a launch is load balanced across the whole GPU whatever the
// system sizes.  Each system's atoms, cells and lists live in its own slices
// of shared arenas; per-system decisions (rebuild, step counts) are made on
// the device, so a round of steps needs no host round trip.  Nothing a system
// computes depends on the other systems in the batch.
#pragma once

#include "cells.hpp"
#include "common.hpp"
#include "forces.hpp"
#include "sort.hpp"

namespace lpmdx {

constexpr int kMaxTerms = 4;
constexpr int kTile = kWave; // atoms per tile
constexpr int kSortThreads = 512;

// Read-only description of one system for a round.
struct BSys {
  Box b;
  double ll[3];
  Grid g; // exact grid: cells no smaller than the list radius
  float Lf[3], halff[3];
  int path;  // BuildPath
  int live;  // slot in use
  long a0;   // first atom slot (multiple of kTile)
  int n;     // atoms
  int ntile; // tiles
  long c0;   // first slot of its cell ranges (ncell + 1)
  long l0;   // first word of its list block
  int cap;   // list entries per atom (multiple of kChunk)
  int lpa;   // lanes per atom of its kind's force launch
  double rl2;
  float rl2f;
  double budget; // half the skin: rebuild once the displacement bound passes it
  double dt;
  int kind, nterm, metal;
  DevPot P[kMaxTerms];
};

// Mutable per-system state on the device.
struct BState {
  double disp;        // displacement bound since the list build
  long left;          // steps still to run
  long done;          // steps run
  unsigned int maxd2; // largest |dx|^2 of the last step (float bits)
  int act;            // this step advances the system
  int flag;           // this step rebuilds its list first
  int par;            // positions of the next force evaluation are in x[par]
  int upar;           // positions of the last completed step are in x[upar]
  int built;          // a list exists
  int init;           // 1: the first force (and drift) is pending, 2: running
  int overflow;       // a list outgrew cap (sticky)
  unsigned int maxcnt;        // longest list of the last build
  unsigned int pad;
  unsigned long long sumcnt;  // list entries of the last build
  long builds;                // list builds so far
};

// Arena pointers (one entry per atom slot unless noted).
struct BArena {
  double* x[2]; // [3A] position buffers
  double* v;    // [3A]
  double* a;    // [3A]
  int* z;
  double* ffm;
  double* mass;
  int* perm;    // caller's atom index within its system
  float4* xf;
  float* w;     // EAM rho^{-1/2}
  double* rho;
  int* cnt;
  int* nbr;     // list arena (words)
  int* cstart;  // cell arena
  int* cend;
  uint32_t* keys[2];
  uint32_t* vals[2];
  double* sv;   // [3A] reorder scratch
  double* sa;   // [3A]
  int* sz;
  double* sffm;
  double* smass;
  int* sperm;
};

struct BArgs {
  const BSys* sys;
  BState* st;
  BArena R;
  const int* tiles;   // tile list of this launch (global tile index = atom slot / kTile)
  const int* tsys;    // system of every global tile
  long ntiles;        // entries of `tiles`
  int nsys;
};

// ---------------------------------------------------------------- control
static __global__ __launch_bounds__(kBlock) void k_b_control(BArgs B) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= B.nsys)
    return;
  const BSys& S = B.sys[s];
  BState& st = B.st[s];
  // the last step's drift moved the positions into the other buffer
  if (st.act)
    st.par ^= 1;
  if (!S.live) {
    st.act = 0;
    return;
  }
  if (st.init == 2)
    st.init = 0;
  float m;
  {
    unsigned int bits = st.maxd2;
    m = __uint_as_float(bits);
  }
  st.maxd2 = 0u;
  if (st.init == 1) {
    // first force evaluation of the system: a = F(x0), then the drift to x1
    st.init = 2;
    st.act = 1;
    st.flag = 1;
    st.disp = 0.0;
    st.built = 1;
    st.upar = st.par;
    return;
  }
  if (st.left <= 0) {
    st.act = 0;
    return;
  }
  st.act = 1;
  st.left -= 1;
  st.done += 1;
  st.disp += sqrt((double)m);
  st.flag = (!st.built) || (st.disp > S.budget);
  if (st.flag) {
    st.disp = 0.0;
    st.built = 1;
  }
  st.upar = st.par;
}

// ---------------------------------------------------------------- sort
// Stable LSD radix sort (4-bit digits) of one system's atoms by cell, inside
// one workgroup; thread t owns the contiguous items [t*ipt, (t+1)*ipt), so
// equal keys keep their order.  Then cell ranges, the reorder of the per-atom
// state, and the FP32 folded positions for the list build.
static __global__ __launch_bounds__(kSortThreads) void k_b_sort(BArgs B) {
  const int s = blockIdx.x;
  if (s >= B.nsys)
    return;
  const BSys& S = B.sys[s];
  BState& st = B.st[s];
  if (!(st.act && st.flag))
    return;
  __shared__ uint32_t table[kDigits * kSortThreads];
  __shared__ uint32_t part[kSortThreads];
  const int t = threadIdx.x;
  const long a0 = S.a0;
  const int n = S.n;
  const int ipt = (n + kSortThreads - 1) / kSortThreads;
  const int lo = min(n, t * ipt), hi = min(n, lo + ipt);
  const Grid& g = S.g;
  const double* xc = B.R.x[st.par] + 3 * a0;
  uint32_t* ka = B.R.keys[0] + a0;
  uint32_t* va = B.R.vals[0] + a0;
  uint32_t* kb = B.R.keys[1] + a0;
  uint32_t* vb = B.R.vals[1] + a0;
  for (int k = lo; k < hi; ++k) {
    ka[k] = cell_key(g, S.b, xc + 3 * k);
    va[k] = (uint32_t)k;
  }
  int bits = 0;
  while (bits < 32 && ((uint32_t)g.ncell >> bits))
    ++bits;
  const int passes = max(1, (bits + 3) / 4);
  for (int ps = 0; ps < passes; ++ps) {
    const int shift = 4 * ps;
    __syncthreads();
    uint32_t c[kDigits];
#pragma unroll
    for (int d = 0; d < kDigits; ++d)
      c[d] = 0;
    for (int k = lo; k < hi; ++k) {
      const uint32_t d = (ka[k] >> shift) & (kDigits - 1);
#pragma unroll
      for (int e = 0; e < kDigits; ++e)
        c[e] += (d == (uint32_t)e) ? 1u : 0u;
    }
#pragma unroll
    for (int d = 0; d < kDigits; ++d)
      table[d * kSortThreads + t] = c[d];
    __syncthreads();
    // exclusive scan of the digit-major table: thread t scans 16 entries
    {
      uint32_t sum = 0;
      for (int q = 0; q < kDigits; ++q)
        sum += table[t * kDigits + q];
      part[t] = sum;
      __syncthreads();
      for (int o = 1; o < kSortThreads; o <<= 1) {
        uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
      }
      uint32_t run = t ? part[t - 1] : 0u;
      for (int q = 0; q < kDigits; ++q) {
        const uint32_t v = table[t * kDigits + q];
        table[t * kDigits + q] = run;
        run += v;
      }
    }
    __syncthreads();
    uint32_t pos[kDigits];
#pragma unroll
    for (int d = 0; d < kDigits; ++d)
      pos[d] = table[d * kSortThreads + t];
    for (int k = lo; k < hi; ++k) {
      const uint32_t key = ka[k];
      const uint32_t d = (key >> shift) & (kDigits - 1);
      uint32_t p = 0;
#pragma unroll
      for (int e = 0; e < kDigits; ++e)
        if (d == (uint32_t)e) {
          p = pos[e];
          pos[e] = p + 1;
        }
      kb[p] = key;
      vb[p] = va[k];
    }
    uint32_t* tk = ka;
    ka = kb;
    kb = tk;
    uint32_t* tv = va;
    va = vb;
    vb = tv;
  }
  __syncthreads();
  // sorted keys where the list build reads them (keys[0])
  uint32_t* k0 = B.R.keys[0] + a0;
  if (ka != k0)
    for (int k = lo; k < hi; ++k)
      k0[k] = ka[k];
  // cell ranges
  int* cs = B.R.cstart + S.c0;
  int* ce = B.R.cend + S.c0;
  for (int q = t; q <= g.ncell; q += kSortThreads) {
    cs[q] = 0;
    ce[q] = 0;
  }
  __syncthreads();
  for (int k = lo; k < hi; ++k) {
    const uint32_t key = ka[k];
    if (k == 0 || ka[k - 1] != key)
      cs[key] = k;
    if (k == n - 1 || ka[k + 1] != key)
      ce[key] = k + 1;
  }
  // reorder: positions into the other buffer, the rest through scratch
  const int np = st.par ^ 1;
  double* xo = B.R.x[np] + 3 * a0;
  for (int k = lo; k < hi; ++k) {
    const long o = va[k];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      xo[3 * k + q] = xc[3 * o + q];
      B.R.sv[3 * (a0 + k) + q] = B.R.v[3 * (a0 + o) + q];
      B.R.sa[3 * (a0 + k) + q] = B.R.a[3 * (a0 + o) + q];
    }
    B.R.sz[a0 + k] = B.R.z[a0 + o];
    B.R.sffm[a0 + k] = B.R.ffm[a0 + o];
    B.R.smass[a0 + k] = B.R.mass[a0 + o];
    B.R.sperm[a0 + k] = B.R.perm[a0 + o];
  }
  __syncthreads();
  for (int k = lo; k < hi; ++k) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      B.R.v[3 * (a0 + k) + q] = B.R.sv[3 * (a0 + k) + q];
      B.R.a[3 * (a0 + k) + q] = B.R.sa[3 * (a0 + k) + q];
    }
    B.R.z[a0 + k] = B.R.sz[a0 + k];
    B.R.ffm[a0 + k] = B.R.sffm[a0 + k];
    B.R.mass[a0 + k] = B.R.smass[a0 + k];
    B.R.perm[a0 + k] = B.R.sperm[a0 + k];
    // FP32 positions folded into [0, L) for the list build
    float w[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double v = xo[3 * k + q];
      v -= S.b.L[q] * floor(v / S.b.L[q]);
      w[q] = (float)v;
    }
    B.R.xf[a0 + k] = make_float4(w[0], w[1], w[2], 0.0f);
  }
  __syncthreads();
  if (t == 0) {
    st.par = np;
    st.upar = np;
    st.maxcnt = 0u;
    st.sumcnt = 0ull;
    st.builds += 1;
  }
}

// ---------------------------------------------------------------- per-tile views
// Atom `i` (within its system) served by this thread, and lane `l` of it.
struct TileSlot {
  int s;
  long i;
  int l;
  bool ok; // the tile exists
};

template <int LPA> __device__ __forceinline__ TileSlot tile_slot(const BArgs& B) {
  const long t = xcd_block() * kBlock + threadIdx.x;
  const long per = (long)kTile * LPA; // threads per tile
  const long tl = t / per;
  TileSlot r{0, 0, 0, false};
  if (tl >= B.ntiles)
    return r;
  const int tile = B.tiles[tl];
  const int s = B.tsys[tile];
  const long r0 = t % per;
  r.s = s;
  r.i = (long)tile * kTile - B.sys[s].a0 + r0 / LPA;
  r.l = (int)(r0 % LPA);
  r.ok = true;
  return r;
}

__device__ __forceinline__ ForceArgs force_view(const BArgs& B, int s, double* xn_override) {
  const BSys& S = B.sys[s];
  const BState& st = B.st[s];
  const long a0 = S.a0;
  ForceArgs A{};
  A.x = B.R.x[st.par] + 3 * a0;
  A.xn = xn_override ? xn_override : B.R.x[st.par ^ 1] + 3 * a0;
  A.v = B.R.v + 3 * a0;
  A.a = B.R.a + 3 * a0;
  A.z = B.R.z + a0;
  A.ffm = B.R.ffm + a0;
  A.nbr = B.R.nbr + S.l0;
  A.cnt = B.R.cnt + a0;
  A.w_rho = B.R.w + a0;
  A.maxd2 = &B.st[s].maxd2;
  A.P = S.P[0];
  A.terms = S.nterm > 1 ? S.P : nullptr;
  A.nterm = S.nterm;
  A.b = S.b;
  for (int q = 0; q < 3; ++q)
    A.ll[q] = S.ll[q];
  A.dt = S.dt;
  A.n = S.n;
  A.stride = (long)S.ntile * kTile;
  A.cap = S.cap;
  A.mode = st.init == 2 ? kInitDrift : kStep;
  return A;
}

// ---------------------------------------------------------------- build
static __global__ __launch_bounds__(kBlock) void k_b_build(BArgs B) {
  const TileSlot ts = tile_slot<1>(B);
  if (!ts.ok)
    return;
  const BSys& S = B.sys[ts.s];
  BState& st = B.st[ts.s];
  if (!(st.act && st.flag))
    return; // uniform over the wavefront: one tile, one system
  const long a0 = S.a0;
  BuildArgs A{};
  A.x = B.R.x[st.par] + 3 * a0;
  A.xf = B.R.xf + a0;
  A.skeys = B.R.keys[0] + a0;
  A.cstart = B.R.cstart + S.c0;
  A.cend = B.R.cend + S.c0;
  A.nbr = B.R.nbr + S.l0;
  A.cnt = B.R.cnt + a0;
  A.max_count = &st.maxcnt;
  A.sum_count = &st.sumcnt;
  A.g = S.g;
  A.b = S.b;
  for (int q = 0; q < 3; ++q) {
    A.Lf[q] = S.Lf[q];
    A.halff[q] = S.halff[q];
  }
  A.rl2 = S.rl2;
  A.rl2f = S.rl2f;
  A.n = S.n;
  A.stride = (long)S.ntile * kTile;
  A.cap = S.cap;
  A.lpa = S.lpa;
  A.path = S.path;
  build_atom(A, ts.i, ts.i < S.n);
  // each lane checks the count it has just written itself
  if (ts.i < S.n && A.cnt[ts.i] > S.cap)
    st.overflow = 1;
}

// ---------------------------------------------------------------- forces
template <int KIND, int LPA>
static __global__ __launch_bounds__(kBlock) void k_b_rho(BArgs B) {
  const TileSlot ts = tile_slot<LPA>(B);
  if (!ts.ok || !B.st[ts.s].act)
    return; // uniform over the wavefront
  const ForceArgs A = force_view(B, ts.s, nullptr);
  eam_rho_atom<KIND, LPA, false>(A, ts.i, ts.l, ts.i < A.n, B.R.w + B.sys[ts.s].a0, nullptr);
}

template <int KIND, int LPA>
static __global__ __launch_bounds__(kBlock) void k_b_force(BArgs B) {
  const TileSlot ts = tile_slot<LPA>(B);
  if (!ts.ok || !B.st[ts.s].act)
    return; // uniform over the wavefront
  const ForceArgs A = force_view(B, ts.s, nullptr);
  if (B.sys[ts.s].metal)
    eam_force_atom<KIND, LPA, false>(A, ts.i, ts.l, ts.i < A.n);
  else
    pair_atom<KIND, LPA, false>(A, ts.i, ts.l, ts.i < A.n);
}

// Energy of one system at its user positions x[upar] (no state change): the
// launch covers exactly that system's tiles, so every block belongs to it and
// its block partials sum, in a fixed order, to the system's totals.
template <int KIND, int LPA>
static __global__ __launch_bounds__(kBlock) void k_b_eval(BArgs B, int s, double* scratch_a, double* partials,
                                                   int rho_pass) {
  const long t = xcd_block() * kBlock + threadIdx.x;
  const BSys& S = B.sys[s];
  const long i = t / LPA;
  const int l = (int)(t % LPA);
  ForceArgs A = force_view(B, s, nullptr);
  A.x = B.R.x[B.st[s].upar] + 3 * S.a0;
  A.a = scratch_a;
  A.mode = kForcesOnly;
  A.partials = partials;
  A.maxd2 = nullptr;
  if (S.metal) {
    if (rho_pass)
      eam_rho_atom<KIND, LPA, true>(A, i, l, i < A.n, B.R.w + S.a0, nullptr);
    else
      eam_force_atom<KIND, LPA, true>(A, i, l, i < A.n);
  } else {
    pair_atom<KIND, LPA, true>(A, i, l, i < A.n);
  }
}

} // namespace lpmdx

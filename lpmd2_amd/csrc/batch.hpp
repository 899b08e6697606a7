// Device side of the batched engine: many independent periodic systems of
// uneven size advanced together.  Every step is the same fixed sequence of
// launches (captured once in a HIP graph and replayed):
//   k_b_control  one thread per system: step bookkeeping, skin check
//   k_b_count / k_b_scan / k_b_scatter (one triple per 8-bit digit pass)
//                stable radix sort of the atoms of every system that needs a
//                new list by cell: 1024-atom chunks, one workgroup each, so a
//                large system is sorted by many workgroups at once
//   k_b_gather / k_b_commit / k_b_sorted
//                cell starts and the reorder of the per-atom state
//   k_b_build    every tile of those systems: neighbour lists
//   k_b_rho      EAM systems: rho pass            (one launch per kind)
//   k_b_force    all active systems: forces + velocity Verlet (per kind);
//                one workgroup per tile
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
constexpr int kMaxSortCells = 1 << (2 * kRadixBits); // cells per system: at most two digit passes
constexpr int kScanThreads = 1024;
// Sort chunks of the batch: 1024 atoms (4 per thread), a quarter of the plain
// path's, so that the few systems rebuilding in one step still give the
// count and scatter passes many workgroups
constexpr int kBSortItems = 4;
constexpr int kBSortChunk = kBlock * kBSortItems;

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
  double margin; // largest separation of a listed pair (0: always take the minimum image)
  double dt;
  int kind, nterm, metal;
  int one_term; // every species pair present meets at most one of the nterm potentials
  int passes;   // digit passes its cell ids need (1 or 2)
  int ch0, nch; // its first sort chunk and number of chunks
  DevPot P[kMaxTerms];
};

// Sutton-Chen systems with the common integer exponent pairs, (9, 6) and
// (12, 6), run force and rho kernels with the exponents fixed at compile time,
// as the single-system engine does (bit-identical results, about 20 fewer
// instructions per neighbour): the systems of a kind are grouped by this
// key, sc_exp(n, m) for those and 0 for every other system.
__host__ __device__ inline int batch_exp(const BSys& S) {
  if (S.kind != kSuttonChen || S.P[0].ni <= 0 || S.P[0].mi <= 0)
    return 0;
  const int ex = sc_exp(S.P[0].ni, S.P[0].mi);
  return (ex == sc_exp(9, 6) || ex == sc_exp(12, 6)) ? ex : 0;
}

// Mutable per-system state on the device.
struct BState {
  double disp;        // largest displacement from the build positions
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
  long quota;                 // steps the current run may still advance it
};

// Arena pointers (one entry per atom slot unless noted).
struct BArena {
  double* x[2]; // [3A] position buffers
  double* xb;   // [3A] positions at the last list build
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
  int* cb;       // cell arena: cell starts of each system (ncell + 2)
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
  const int2* chunks; // sort chunks: (system, chunk index within it)
  int nchunks;
  uint32_t* hist;     // per system: kRadix x nch digit counts, digit-major
};

// ---------------------------------------------------------------- control
// lite: a step launched without the rebuild chain, in which a system whose
// list has run out (or that still waits for its first list) holds: it does
// not advance and keeps its displacement maximum, so the next full step
// reaches the same rebuild decision.  kind >= 0: only the systems of that
// potential kind and exponent key (batch_exp; the lite steps run on one
// branch per group).  A system advances only while its quota of steps for the
// current run lasts.
static __global__ __launch_bounds__(kBlock) void k_b_control(BArgs B, int lite, int kind, int exp) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= B.nsys)
    return;
  const BSys& S = B.sys[s];
  if (kind >= 0 && (S.kind != kind || batch_exp(S) != exp))
    return;
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
  if ((st.init != 1 && st.left <= 0) || st.quota <= 0) {
    st.act = 0;
    return;
  }
  if (st.init == 1) {
    if (lite) {
      st.act = 0;
      return;
    }
    // first force evaluation of the system: a = F(x0), then the drift to x1
    st.maxd2 = 0u;
    st.quota -= 1;
    st.init = 2;
    st.act = 1;
    st.flag = 1;
    st.disp = 0.0;
    st.built = 1;
    st.upar = st.par;
    return;
  }
  float m;
  {
    unsigned int bits = st.maxd2;
    m = __uint_as_float(bits);
  }
  const double disp = sqrt((double)m); // largest |x - xb| of the positions about to be used
  const int need = (!st.built) || (disp > S.budget);
  if (lite && need) {
    st.act = 0;
    return;
  }
  st.maxd2 = 0u;
  st.quota -= 1;
  st.act = 1;
  st.left -= 1;
  st.done += 1;
  st.disp = disp;
  st.flag = need;
  if (st.flag) {
    st.disp = 0.0;
    st.built = 1;
  }
  st.upar = st.par;
}

// ---------------------------------------------------------------- sort
// Stable LSD radix sort of each rebuilding system's atoms by cell id, 8-bit
// digits.  Pass p: k_b_count histograms digit p per chunk (the first pass
// also computes the cell ids from the positions), k_b_scan turns one
// system's digit-major (digit, chunk) counts into scatter offsets, and
// k_b_scatter writes every key to offset + its rank among the equal digits
// before it in the chunk (wave ballots give the rank inside a wavefront,
// LDS prefixes across the wavefronts), so equal cell ids keep ascending
// atom order.  Keys and values ping-pong between keys[0]/vals[0] and
// keys[1]/vals[1]; a system's sorted order ends in buffer (passes & 1).

__device__ __forceinline__ bool sort_chunk(const BArgs& B, int pass, int& s, int& j) {
  const int c = blockIdx.x;
  if (c >= B.nchunks)
    return false;
  const int2 cj = B.chunks[c];
  s = cj.x;
  j = cj.y;
  const BState& st = B.st[s];
  return B.sys[s].live && st.act && st.flag && pass < B.sys[s].passes;
}

static __global__ __launch_bounds__(kBlock) void k_b_count(BArgs B, int pass) {
  int s, j;
  if (!sort_chunk(B, pass, s, j))
    return; // uniform over the workgroup
  const BSys& S = B.sys[s];
  const BState& st = B.st[s];
  __shared__ uint32_t h[kRadix];
  const int t = threadIdx.x;
  h[t] = 0u;
  __syncthreads();
  const long a0 = S.a0;
  const int n = S.n;
  const int shift = kRadixBits * pass;
  uint32_t* kin = B.R.keys[pass & 1] + a0;
  uint32_t* vin = B.R.vals[pass & 1] + a0;
  const double* xc = B.R.x[st.par] + 3 * a0;
  for (int u = 0; u < kBSortItems; ++u) {
    const int i = j * kBSortChunk + u * kBlock + t;
    if (i < n) {
      uint32_t key;
      if (pass == 0) {
        key = cell_key(S.g, S.b, xc + 3 * (long)i);
        kin[i] = key;
        vin[i] = (uint32_t)i;
      } else {
        key = kin[i];
      }
      atomicAdd(&h[(key >> shift) & (kRadix - 1)], 1u); // a count: order-free
    }
  }
  __syncthreads();
  B.hist[(long)S.ch0 * kRadix + (long)t * S.nch + j] = h[t];
}

static __global__ __launch_bounds__(kScanThreads) void k_b_scan(BArgs B, int pass) {
  const int s = blockIdx.x;
  if (s >= B.nsys)
    return;
  const BSys& S = B.sys[s];
  const BState& st = B.st[s];
  if (!S.live || !(st.act && st.flag) || pass >= S.passes)
    return;
  __shared__ uint32_t part[kScanThreads];
  const int t = threadIdx.x;
  uint32_t* h = B.hist + (long)S.ch0 * kRadix;
  const long m = (long)kRadix * S.nch;
  const long per = (m + kScanThreads - 1) / kScanThreads;
  const long lo = min(m, (long)t * per), hi = min(m, lo + per);
  uint32_t sum = 0;
  for (long k = lo; k < hi; ++k)
    sum += h[k];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (long k = lo; k < hi; ++k) {
    const uint32_t v = h[k];
    h[k] = run;
    run += v;
  }
}

static __global__ __launch_bounds__(kBlock) void k_b_scatter(BArgs B, int pass) {
  int s, j;
  if (!sort_chunk(B, pass, s, j))
    return; // uniform over the workgroup
  const BSys& S = B.sys[s];
  __shared__ RankLds L;
  const int t = threadIdx.x;
  L.run[t] = B.hist[(long)S.ch0 * kRadix + (long)t * S.nch + j];
  const long a0 = S.a0;
  const int n = S.n;
  const int shift = kRadixBits * pass;
  const uint32_t* kin = B.R.keys[pass & 1] + a0;
  const uint32_t* vin = B.R.vals[pass & 1] + a0;
  uint32_t* kout = B.R.keys[(pass + 1) & 1] + a0;
  uint32_t* vout = B.R.vals[(pass + 1) & 1] + a0;
  for (int u = 0; u < kBSortItems; ++u) {
    const int base = j * kBSortChunk + u * kBlock;
    if (base >= n)
      break; // uniform over the workgroup
    const int i = base + t;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t pos = rank_tile(valid, (key >> shift) & (kRadix - 1), L);
    if (valid && pos < (uint32_t)n) {
      kout[pos] = key;
      vout[pos] = val;
    }
  }
}

// Cell ranges from the sorted ids and the reorder of the per-atom state:
// positions into the other position buffer, the rest through scratch.
static __global__ __launch_bounds__(kBlock) void k_b_gather(BArgs B) {
  const long t = xcd_block() * kBlock + threadIdx.x;
  const long tl = t / kTile;
  if (tl >= B.ntiles)
    return;
  const int tile = wave_uniform(B.tiles[tl]); // one tile per wavefront
  const int s = wave_uniform(B.tsys[tile]);
  const BSys& S = B.sys[s];
  const BState& st = B.st[s];
  if (!(st.act && st.flag))
    return;
  const long a0 = S.a0;
  const long k = (long)tile * kTile - a0 + (t % kTile);
  const int n = S.n;
  if (k >= n)
    return;
  const uint32_t* K = B.R.keys[S.passes & 1] + a0;
  const uint32_t* V = B.R.vals[S.passes & 1] + a0;
  // cell starts: atom k writes the cells after the previous atom's cell up to
  // its own; the last atom also the cells after its own
  {
    int* cb = B.R.cb + S.c0;
    const long lo = k == 0 ? 0 : (long)K[k - 1] + 1;
    for (long c = lo; c <= (long)K[k]; ++c)
      cb[c] = (int)k;
    if (k == n - 1)
      for (long c = (long)K[k] + 1; c <= (long)S.g.ncell + 1; ++c)
        cb[c] = n;
  }
  long o = V[k];
  if (o >= n)
    o = k; // unreachable for a complete permutation; keeps every read inside the system
  const double* xc = B.R.x[st.par] + 3 * a0;
  double* xo = B.R.x[st.par ^ 1] + 3 * a0;
  double* xb = B.R.xb + 3 * a0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    xo[3 * k + q] = xc[3 * o + q];
    xb[3 * k + q] = xc[3 * o + q];
    B.R.sv[3 * (a0 + k) + q] = B.R.v[3 * (a0 + o) + q];
    B.R.sa[3 * (a0 + k) + q] = B.R.a[3 * (a0 + o) + q];
  }
  B.R.sz[a0 + k] = B.R.z[a0 + o];
  B.R.sffm[a0 + k] = B.R.ffm[a0 + o];
  B.R.smass[a0 + k] = B.R.mass[a0 + o];
  B.R.sperm[a0 + k] = B.R.perm[a0 + o];
}

static __global__ __launch_bounds__(kBlock) void k_b_commit(BArgs B) {
  const long t = xcd_block() * kBlock + threadIdx.x;
  const long tl = t / kTile;
  if (tl >= B.ntiles)
    return;
  const int tile = wave_uniform(B.tiles[tl]); // one tile per wavefront
  const int s = wave_uniform(B.tsys[tile]);
  const BSys& S = B.sys[s];
  const BState& st = B.st[s];
  if (!(st.act && st.flag))
    return;
  const long a0 = S.a0;
  const long k = (long)tile * kTile - a0 + (t % kTile);
  if (k >= S.n)
    return;
  const long g = a0 + k;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    B.R.v[3 * g + q] = B.R.sv[3 * g + q];
    B.R.a[3 * g + q] = B.R.sa[3 * g + q];
  }
  B.R.z[g] = B.R.sz[g];
  B.R.ffm[g] = B.R.sffm[g];
  B.R.mass[g] = B.R.smass[g];
  B.R.perm[g] = B.R.sperm[g];
  // FP32 positions folded into [0, L) for the list build
  const double* xo = B.R.x[st.par ^ 1] + 3 * a0;
  float w[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double v = xo[3 * k + q];
    v -= S.b.L[q] * floor(v / S.b.L[q]);
    w[q] = (float)v;
  }
  B.R.xf[g] = make_float4(w[0], w[1], w[2], 0.0f);
}

// The sorted positions become the current ones.
static __global__ __launch_bounds__(kBlock) void k_b_sorted(BArgs B) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= B.nsys)
    return;
  BState& st = B.st[s];
  if (!B.sys[s].live || !(st.act && st.flag))
    return;
  const int np = st.par ^ 1;
  st.par = np;
  st.upar = np;
  st.maxcnt = 0u;
  st.sumcnt = 0ull;
  st.builds += 1;
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
  const int tile = wave_uniform(B.tiles[tl]); // one tile per wavefront
  const int s = wave_uniform(B.tsys[tile]);
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
  A.xb = B.R.xb + 3 * a0;
  A.margin = S.margin;
  A.P = S.P[0];
  A.terms = S.nterm > 1 ? S.P : nullptr;
  A.nterm = S.nterm;
  A.one_term = S.one_term;
  A.b = S.b;
  for (int q = 0; q < 3; ++q) {
    A.ll[q] = S.ll[q];
    A.margin_hi[q] = S.b.L[q] - S.margin;
  }
  A.dt = S.dt;
  A.n = S.n;
  A.stride = (long)S.ntile * kTile;
  A.cap = S.cap;
  A.mode = st.init == 2 ? kInitDrift : kStep;
  return A;
}

// ---------------------------------------------------------------- build
static __global__ __launch_bounds__(kBlock) void k_b_build(BArgs B) {
  __shared__ int ring[kRing * kBlock];
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
  A.skeys = B.R.keys[S.passes & 1] + a0;
  A.cb = B.R.cb + S.c0;
  A.bhalo = nullptr;
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
  A = wave_uniform_record(A); // one system per wavefront: its arguments in scalar registers
  build_atom(A, ts.i, ts.i < S.n, ring);
  // each lane checks the count it has just written itself
  if (ts.i < S.n && A.cnt[ts.i] > S.cap)
    st.overflow = 1;
}

// ---------------------------------------------------------------- forces
// Workgroups of T tiles (T = blockDim.x / (64 * LPA): the tile shape the
// autotuner picks, 64, 128 or 256 atoms a workgroup), 64 * LPA threads a
// tile of the launch's tile list; every wavefront lies inside one tile, so
// per-system branches are uniform over a wavefront.  The EAM passes load
// each list batch when it starts (no prefetch): measured faster for these
// small workgroups.
constexpr int kMaxTilesPerGroup = 4;

struct TileLane {
  long tl; // entry of the launch's tile list
  int r;   // thread within the tile
};

template <int LPA> __device__ __forceinline__ TileLane tile_lane() {
  const int per = kWave * LPA;
  return TileLane{xcd_block() * (long)(blockDim.x / per) + threadIdx.x / per, (int)(threadIdx.x % per)};
}

template <int KIND, int LPA, int EXP = 0>
static __global__ __launch_bounds__(kWave * LPA * kMaxTilesPerGroup) void k_b_rho(BArgs B) {
  const TileLane tw = tile_lane<LPA>();
  if (tw.tl >= B.ntiles)
    return; // uniform over the wavefront
  const int tile = wave_uniform(B.tiles[tw.tl]);
  const int s = wave_uniform(B.tsys[tile]);
  if (!B.st[s].act)
    return; // uniform over the wavefront
  const ForceArgs A = force_view(B, s, nullptr);
  const long i = (long)tile * kTile - B.sys[s].a0 + tw.r / LPA;
  const int l = tw.r % LPA;
  double emb = 0.0;
  eam_rho_atom<KIND, LPA, false, false, false, EXP>(A, Stage{}, i, l, i < A.n, B.R.w + B.sys[s].a0, nullptr, emb);
}

template <int KIND, int LPA, int EXP = 0>
static __global__ __launch_bounds__(kWave * LPA * kMaxTilesPerGroup) void k_b_force(BArgs B) {
  const TileLane tw = tile_lane<LPA>();
  if (tw.tl >= B.ntiles)
    return; // uniform over the wavefront
  const int tile = wave_uniform(B.tiles[tw.tl]);
  const int s = wave_uniform(B.tsys[tile]);
  if (!B.st[s].act)
    return; // uniform over the wavefront
  const ForceArgs A = force_view(B, s, nullptr);
  const long i = (long)tile * kTile - B.sys[s].a0 + tw.r / LPA;
  const int l = tw.r % LPA;
  Tally tal;
  // species are read only by the systems whose potentials need them (the
  // branch is uniform: one workgroup, one system)
  if constexpr (metal_kind(KIND)) {
    if (A.P.spc_sum >= 0)
      eam_force_atom<KIND, LPA, false, false, true, false, EXP>(A, Stage{}, i, l, i < A.n, tal);
    else
      eam_force_atom<KIND, LPA, false, false, false, false, EXP>(A, Stage{}, i, l, i < A.n, tal);
  } else {
    if (A.nterm > 1 || A.P.spc_sum >= 0)
      pair_atom<KIND, LPA, false, false, true>(A, Stage{}, i, l, i < A.n, tal);
    else
      pair_atom<KIND, LPA, false, false, false>(A, Stage{}, i, l, i < A.n, tal);
  }
}

// Energy of one system at its user positions x[upar] (no state change): one
// workgroup per tile of that system, block partials summed in a fixed order.
template <int KIND, int LPA>
static __global__ __launch_bounds__(kWave * LPA) void k_b_eval(BArgs B, int s, double* scratch_a,
                                                               double* partials, int rho_pass) {
  const BSys& S = B.sys[s];
  const long i = (long)xcd_block() * kTile + threadIdx.x / LPA;
  const int l = (int)(threadIdx.x % LPA);
  ForceArgs A = force_view(B, s, nullptr);
  A.x = B.R.x[B.st[s].upar] + 3 * S.a0;
  A.a = scratch_a;
  A.mode = kForcesOnly;
  A.partials = partials;
  A.maxd2 = nullptr;
  Tally tal;
  double emb = 0.0;
  if constexpr (metal_kind(KIND)) {
    if (rho_pass)
      eam_rho_atom<KIND, LPA, true, false>(A, Stage{}, i, l, i < A.n, B.R.w + S.a0, nullptr, emb);
    else
      eam_force_atom<KIND, LPA, true, false, true>(A, Stage{}, i, l, i < A.n, tal);
  } else {
    pair_atom<KIND, LPA, true, false, true>(A, Stage{}, i, l, i < A.n, tal);
  }
  write_partials(A, tal, emb);
}

} // namespace lpmdx

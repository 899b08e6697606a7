// Host side of the batched engine (lpmdx_batch_* in include/lpmdx.h): many
// independent periodic systems in shared HBM arenas, advanced together by a
// HIP graph that replays one step of fixed launches (see batch.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <set>

#include "batch.hpp"
#include "host_util.hpp"

static_assert(lpmdx::kNumKinds == LPMDX_NKINDS, "kind count of the C ABI");

using namespace lpmdx;

namespace {

// First-fit allocator over [0, cap) in whole units.
struct Arena {
  long cap = 0;
  std::map<long, long> free; // start -> length
  void init(long c) {
    cap = c;
    free.clear();
    if (c > 0)
      free[0] = c;
  }
  long alloc(long len) {
    if (len <= 0)
      return 0;
    for (auto it = free.begin(); it != free.end(); ++it) {
      if (it->second >= len) {
        const long s = it->first, l = it->second;
        free.erase(it);
        if (l > len)
          free[s + len] = l - len;
        return s;
      }
    }
    return -1;
  }
  void release(long s, long len) {
    if (len <= 0)
      return;
    auto it = free.emplace(s, len).first;
    auto nx = std::next(it);
    if (nx != free.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free.erase(nx);
    }
    if (it != free.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free.erase(it);
      }
    }
  }
  long largest() const {
    long m = 0;
    for (auto& f : free)
      m = std::max(m, f.second);
    return m;
  }
};

struct HostSys {
  bool live = false;
  long n = 0, tiles = 0, atom0 = 0, list0 = 0, listw = 0, cell0 = 0, cells = 0;
  double rl = 0.0; // list radius (cutoff + skin)
  std::vector<std::unique_ptr<PotParams>> pots;
  int kind = 0;
};

constexpr long kBytesPerAtom = 264; // per atom slot over all per-atom arenas
constexpr int kMaxCycle = 16;       // longest cycle of one full and cyc - 1 lite steps

// Cells per list radius of a system's exact grid for div (1 or 2): 2 only
// where five half-radius cells fit along every axis within the sort's cell
// limit and the FP32 one-shift-per-cell path applies.
int grid_div(const double box[3], double rl, int div) {
  if (div != 2)
    return 1;
  const double lmax = std::max(box[0], std::max(box[1], box[2]));
  const double pos_err = 2.0 * std::ldexp(lmax, -24) + 1e-30;
  const double hmin = 0.5 * std::min(box[0], std::min(box[1], box[2]));
  if (!(hmin > rl + 4.0 * pos_err + 1e-9 * hmin))
    return 1;
  long cells = 1;
  for (int q = 0; q < 3; ++q) {
    const double m = std::floor(box[q] / (0.5 * rl * (1.0 + 1e-9)));
    if (m < 5.0 || m > 512.0)
      return 1;
    cells *= (long)m;
  }
  return cells <= kMaxSortCells ? 2 : 1;
}

// The exact grid of system S (box S.b, list radius rl) with cells of at
// least rl / div per axis (div as grid_div resolves it), its sort passes and
// build path.
void set_grid(BSys& S, double rl, int div) {
  const double box[3] = {S.b.L[0], S.b.L[1], S.b.L[2]};
  div = grid_div(box, rl, div);
  int nq[3];
  for (int q = 0; q < 3; ++q)
    nq[q] = (int)std::max(1.0, std::min(std::floor(box[q] / (rl / div * (1.0 + 1e-9))), 512.0));
  // at most kMaxSortCells cells (two digit passes of the sort): coarser
  // cells only add candidates, the list is the same
  while ((long)nq[0] * nq[1] * nq[2] > kMaxSortCells) {
    const int q = (nq[0] >= nq[1] && nq[0] >= nq[2]) ? 0 : (nq[1] >= nq[2] ? 1 : 2);
    nq[q] -= 1;
  }
  S.g.nx = nq[0];
  S.g.ny = nq[1];
  S.g.nz = nq[2];
  S.g.side = 1;
  S.g.mode = kExact;
  S.g.div = div;
  S.g.ncell = nq[0] * nq[1] * nq[2];
  {
    int bits = 0;
    while (bits < 31 && ((uint32_t)(S.g.ncell - 1) >> bits))
      ++bits;
    S.passes = std::max(1, (bits + kRadixBits - 1) / kRadixBits);
  }
  const double lmax = std::max(box[0], std::max(box[1], box[2]));
  const double pos_err = 2.0 * std::ldexp(lmax, -24) + 1e-30;
  const double hmin = std::min(S.b.half[0], std::min(S.b.half[1], S.b.half[2]));
  S.path = kPathExact64;
  if (hmin > rl + 4.0 * pos_err + 1e-9 * hmin)
    S.path = (nq[0] >= 3 && nq[1] >= 3 && nq[2] >= 3) ? kPathShift32 : kPathMin32;
}

} // namespace

struct lpmdx_batch {
  hipStream_t st = nullptr;
  std::vector<hipStream_t> bst; // per-kind branches of a step
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  long acap = 0, lcap = 0, ccap = 0;
  Arena atoms, lists, cells; // atoms counted in tiles
  // per-atom arenas
  DBuf<double> x0, x1, xb, v, a, sv, sa, ffm, mass, sffm, smass, rho;
  DBuf<int> z, sz, perm, sperm, cnt, nbr, cb, tsys;
  DBuf<float4> xf;
  DBuf<float> w;
  DBuf<uint32_t> k0, k1, v0, v1;
  // tables
  std::vector<BSys> hsys;
  std::vector<HostSys> meta;
  DBuf<BSys> dsys;
  DBuf<BState> dst;
  BState* hst = nullptr; // pinned mirror
  // tile lists
  DBuf<int> tiles_all;
  long n_all = 0;
  // sort chunks and their digit counts
  DBuf<int2> chunks;
  DBuf<uint32_t> hist;
  int nchunks = 0, passes = 1;
  struct Group {
    int kind, exp, lpa, tpw; // exp: batch_exp of its systems
    bool metal;
    DBuf<int> tiles;
    long n = 0;
  };
  std::vector<std::unique_ptr<Group>> groups;
  // lanes per atom of each kind's force launch until autotuned or set: 4
  // (served Gupta / Buckingham / Finnis-Sinclair jobs measured 18 / 31 / 3 %
  // faster than at one lane; the tile kernels keep few atoms per workgroup)
  int lpa_of_kind[kNumKinds + 1] = {0, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};
  // cells per list radius of each kind's grids until autotuned or set: 1
  // (measured on the served Gupta / Finnis-Sinclair / Buckingham mix at
  // full load: 2 was 2 % faster for Gupta, 11 % and 3 % slower for the others)
  int div_of_kind[kNumKinds + 1] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  // tiles (of 64 atoms) per workgroup of each kind's force and rho launches
  // until autotuned or set: 1
  int tpw_of_kind[kNumKinds + 1] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  // graphs of one full step and of one cycle (a full step and cyc - 1 lite steps)
  hipGraph_t graph = nullptr, cgraph = nullptr;
  hipGraphExec_t gexec = nullptr, cexec = nullptr;
  int cyc = 4;
  bool dirty = true;
  long dead_tiles = 0; // tiles of systems removed since the last plan
  // evaluation scratch
  DBuf<double> evala, partials, red;
  double* hred = nullptr;
  long steps_run = 0;

  ~lpmdx_batch() {
    if (st)
      (void)hipStreamSynchronize(st);
    release_graph();
    if (hst)
      (void)hipHostFree(hst);
    if (hred)
      (void)hipHostFree(hred);
    meta.clear();
    for (hipStream_t b : bst)
      (void)hipStreamDestroy(b);
    for (hipEvent_t e : ev_join)
      (void)hipEventDestroy(e);
    if (ev_fork)
      (void)hipEventDestroy(ev_fork);
    if (st)
      (void)hipStreamDestroy(st);
  }

  void sync() { hipchk(hipStreamSynchronize(st), "hipStreamSynchronize"); }

  BArena arena() {
    BArena R{};
    R.x[0] = x0.p;
    R.x[1] = x1.p;
    R.xb = xb.p;
    R.v = v.p;
    R.a = a.p;
    R.z = z.p;
    R.ffm = ffm.p;
    R.mass = mass.p;
    R.perm = perm.p;
    R.xf = xf.p;
    R.w = w.p;
    R.rho = rho.p;
    R.cnt = cnt.p;
    R.nbr = nbr.p;
    R.cb = cb.p;
    R.keys[0] = k0.p;
    R.keys[1] = k1.p;
    R.vals[0] = v0.p;
    R.vals[1] = v1.p;
    R.sv = sv.p;
    R.sa = sa.p;
    R.sz = sz.p;
    R.sffm = sffm.p;
    R.smass = smass.p;
    R.sperm = sperm.p;
    return R;
  }

  BArgs args(const int* tiles, long ntiles) {
    BArgs B{};
    B.sys = dsys.p;
    B.st = dst.p;
    B.R = arena();
    B.tiles = tiles;
    B.tsys = tsys.p;
    B.ntiles = ntiles;
    B.nsys = (int)hsys.size();
    B.chunks = chunks.p;
    B.nchunks = nchunks;
    B.hist = hist.p;
    return B;
  }

  void create(long atom_cap, long list_words, long cell_slots) {
    acap = ((atom_cap + kTile - 1) / kTile) * kTile;
    lcap = list_words;
    ccap = cell_slots;
    atoms.init(acap / kTile);
    lists.init(lcap / kTile);
    cells.init(ccap);
    const size_t A = std::max<long>(acap, kTile);
    for (DBuf<double>* b : {&x0, &x1, &xb, &v, &a, &sv, &sa})
      b->alloc(3 * A);
    for (DBuf<double>* b : {&ffm, &mass, &sffm, &smass, &rho})
      b->alloc(A);
    for (DBuf<int>* b : {&z, &sz, &perm, &sperm, &cnt})
      b->alloc(A);
    for (DBuf<uint32_t>* b : {&k0, &k1, &v0, &v1})
      b->alloc(A);
    xf.alloc(A + kGroup); // padding read by the candidate groups of the last system
    w.alloc(A);
    tsys.alloc(A / kTile);
    nbr.alloc(std::max<long>(lcap, kTile));
    cb.alloc(std::max<long>(ccap, 2));
    hipchk(hipMemsetAsync(x0.p, 0, 3 * A * sizeof(double), st), "memset");
    hipchk(hipMemsetAsync(x1.p, 0, 3 * A * sizeof(double), st), "memset");
    hipchk(hipMemsetAsync(nbr.p, 0, std::max<long>(lcap, kTile) * sizeof(int), st), "memset");
    hipchk(hipHostMalloc(&hred, 16 * sizeof(double)), "hipHostMalloc");
    red.alloc(16);
    sync();
  }

  // ---------------------------------------------------------------- systems
  long add(long n, const double box[3], const int* zz, const double* mm, const double* x,
           const double* vel, const char* cm_args, int npot, const char** names, const char** pargs,
           const int* s1, const int* s2, double dt, long steps) {
    if (n <= 0)
      fail(LPMDX_EINVAL, "a batched system needs at least one atom");
    if (npot < 1 || npot > kMaxTerms)
      fail(LPMDX_EINVAL, "a batched system takes 1 to 4 potentials");
    need(box, "box");
    need(zz, "z");
    need(x, "x");
    finite(box, 3, "box");
    finite(x, 3 * n, "positions");
    if (vel)
      finite(vel, 3 * n, "velocities");
    if (!std::isfinite(dt) || steps < 0)
      fail(LPMDX_EINVAL, "bad dt or step count");
    HostSys H;
    BSys S{};
    for (int q = 0; q < npot; ++q) {
      need(names[q], "potential name");
      H.pots.push_back(make_potential(names[q], pargs ? pargs[q] : "", s1 ? s1[q] : -1, s2 ? s2[q] : -1));
      if (H.pots.back()->kind != H.pots[0]->kind)
        fail(LPMDX_EINVAL, "the potentials of one batched system must be of one kind");
    }
    H.kind = H.pots[0]->kind;
    const bool metal = H.pots[0]->metal();
    if (metal && npot > 1)
      fail(LPMDX_EINVAL, "a batched EAM system takes one potential");
    // the cell manager keywords of the plain path's verletlist with a skin
    Keywords k(cm_args, {{"cutoff", "0.0"}, {"skin", "0.3"}, {"maxneighbors", "0"}});
    double cutoff = k.d("cutoff");
    if (!(cutoff > 0.0)) {
      cutoff = 0.0;
      for (auto& p : H.pots)
        cutoff = std::max(cutoff, std::isfinite(p->cutoff) ? p->cutoff : 0.0);
    }
    if (!(cutoff > 0.0))
      fail(LPMDX_EINVAL, "batched system needs a finite cutoff");
    const double skin = k.d("skin");
    if (!(skin >= 0.0))
      fail(LPMDX_EINVAL, "skin must be non-negative");
    const double rl = cutoff + skin;
    for (int q = 0; q < 3; ++q) {
      if (!(box[q] > 0.0))
        fail(LPMDX_EINVAL, "box edges must be positive");
      S.b.L[q] = box[q];
      S.b.half[q] = 0.5 * box[q];
      S.b.inv[q] = 1.0 / box[q];
      S.ll[q] = edge_module(box[q]);
      S.Lf[q] = (float)box[q];
      S.halff[q] = (float)S.b.half[q];
    }
    // the cell slice holds either grid (the finer one is the larger) when the
    // arena has room, else the whole-radius grid only
    set_grid(S, rl, 1);
    const long ncell1 = S.g.ncell;
    set_grid(S, rl, 2);
    long ncell_max = std::max(ncell1, (long)S.g.ncell);
    if (cells.largest() < ncell_max + 2)
      ncell_max = ncell1;
    set_grid(S, rl, ncell_max >= S.g.ncell ? div_of_kind[H.kind] : 1);
    const double lmax = std::max(box[0], std::max(box[1], box[2]));
    const double pos_err = 2.0 * std::ldexp(lmax, -24) + 1e-30;
    const double hmin = std::min(S.b.half[0], std::min(S.b.half[1], S.b.half[2]));
    S.rl2 = rl * rl;
    S.rl2f = (float)((rl + 2.0 * pos_err) * (rl + 2.0 * pos_err) * (1.0 + 1e-6));
    S.budget = 0.5 * skin;
    {
      const double m = std::sqrt((double)S.rl2f) * (1.0 + 1e-6) + 2.0 * S.budget + 1e-9 * lmax;
      S.margin = m < hmin * (1.0 - 1e-6) ? m : 0.0; // room for min_image_near
    }
    S.dt = dt;
    S.kind = H.kind;
    S.metal = metal;
    S.nterm = npot;
    for (int q = 0; q < npot; ++q) {
      const double own = H.pots[q]->cutoff * H.pots[q]->cutoff;
      const double f = std::min(own, rl * rl);
      S.P[q] = device_view(*H.pots[q], f, f);
    }
    {
      // species pairs of this system that more than one potential applies to
      // (Potential::AppliesTo, potential.cc:23-27): none lets the force kernel
      // evaluate one selected term per pair
      std::set<int> zs(zz, zz + n);
      int most = 0;
      for (int za : zs)
        for (int zb : zs) {
          int c = 0;
          for (int q = 0; q < npot; ++q)
            c += (S.P[q].spc_sum < 0 || S.P[q].spc_sum == za + zb) ? 1 : 0;
          most = std::max(most, c);
        }
      S.one_term = npot > 1 && most <= 1;
    }
    const double vol = box[0] * box[1] * box[2];
    const double est = (double)n / vol * 4.18879 * rl * rl * rl;
    long cap = (long)std::min<double>((double)n, est * 1.5 + 32);
    const long maxn = k.i("maxneighbors");
    if (maxn > 0)
      cap = std::max(cap, maxn);
    cap = ((cap + kChunk - 1) / kChunk) * kChunk;
    const long tiles = (n + kTile - 1) / kTile;
    const long t0 = atoms.alloc(tiles);
    if (t0 < 0)
      fail(LPMDX_ENOMEM, "atom arena full");
    const long lw = tiles * kTile * cap;
    const long l0 = lists.alloc((lw + kTile - 1) / kTile);
    if (l0 < 0) {
      atoms.release(t0, tiles);
      fail(LPMDX_ENOMEM, "list arena full");
    }
    const long c0 = cells.alloc(ncell_max + 2);
    if (c0 < 0) {
      atoms.release(t0, tiles);
      lists.release(l0, (lw + kTile - 1) / kTile);
      fail(LPMDX_ENOMEM, "cell arena full");
    }
    S.live = 1;
    S.a0 = t0 * kTile;
    S.n = (int)n;
    S.ntile = (int)tiles;
    S.c0 = c0;
    S.l0 = l0 * kTile;
    S.cap = (int)cap;
    S.lpa = lpa_of_kind[H.kind];
    H.live = true;
    H.n = n;
    H.tiles = tiles;
    H.atom0 = S.a0;
    H.list0 = l0;
    H.listw = (lw + kTile - 1) / kTile;
    H.cell0 = c0;
    H.cells = ncell_max + 2;
    H.rl = rl;
    // upload the atoms
    std::vector<double> ff(n), ms(n);
    std::vector<int> pm(n), zz2(n);
    for (long i = 0; i < n; ++i) {
      const double mi = mm ? mm[i] : ElemMass(zz[i]);
      if (!(mi > 0.0) || !std::isfinite(mi))
        fail(LPMDX_EINVAL, "no mass for atomic number " + std::to_string(zz[i]));
      ff[i] = kForceFactor / mi;
      ms[i] = mi;
      pm[i] = (int)i;
      zz2[i] = zz[i];
    }
    std::vector<double> zero(3 * n, 0.0);
    const long a0 = S.a0;
    hipchk(hipMemcpyAsync(x0.p + 3 * a0, x, 3 * n * sizeof(double), hipMemcpyHostToDevice, st), "memcpy");
    hipchk(hipMemcpyAsync(v.p + 3 * a0, vel ? vel : zero.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, st),
           "memcpy");
    hipchk(hipMemcpyAsync(a.p + 3 * a0, zero.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, st), "memcpy");
    hipchk(hipMemcpyAsync(ffm.p + a0, ff.data(), n * sizeof(double), hipMemcpyHostToDevice, st), "memcpy");
    hipchk(hipMemcpyAsync(mass.p + a0, ms.data(), n * sizeof(double), hipMemcpyHostToDevice, st), "memcpy");
    hipchk(hipMemcpyAsync(z.p + a0, zz2.data(), n * sizeof(int), hipMemcpyHostToDevice, st), "memcpy");
    hipchk(hipMemcpyAsync(perm.p + a0, pm.data(), n * sizeof(int), hipMemcpyHostToDevice, st), "memcpy");
    // slot
    long sid = -1;
    for (size_t q = 0; q < hsys.size(); ++q)
      if (!meta[q].live) {
        sid = (long)q;
        break;
      }
    if (sid < 0) {
      sid = (long)hsys.size();
      hsys.push_back(S);
      meta.push_back(std::move(H));
    } else {
      hsys[sid] = S;
      meta[sid] = std::move(H);
    }
    BState s0{};
    s0.left = steps;
    s0.init = 1;
    grow_state();
    hst[sid] = s0;
    hipchk(hipMemcpyAsync(dst.p + sid, &hst[sid], sizeof(BState), hipMemcpyHostToDevice, st), "memcpy");
    sync();
    dirty = true;
    return sid;
  }

  void grow_state() {
    const size_t m = std::max<size_t>(hsys.size(), 1);
    if (dst.n >= m && hst)
      return;
    const size_t cap2 = std::max<size_t>(m, 2 * dst.n);
    DBuf<BState> nd;
    nd.alloc(cap2);
    BState* nh = nullptr;
    hipchk(hipHostMalloc(&nh, cap2 * sizeof(BState)), "hipHostMalloc");
    memset(nh, 0, cap2 * sizeof(BState));
    if (dst.p && hst) {
      hipchk(hipMemcpyAsync(nd.p, dst.p, dst.n * sizeof(BState), hipMemcpyDeviceToDevice, st), "memcpy");
      memcpy(nh, hst, dst.n * sizeof(BState));
    } else {
      hipchk(hipMemsetAsync(nd.p, 0, cap2 * sizeof(BState), st), "memset");
    }
    sync();
    std::swap(dst.p, nd.p);
    std::swap(dst.n, nd.n);
    if (hst)
      (void)hipHostFree(hst);
    hst = nh;
  }

  void remove(long sid) {
    check_sid(sid);
    HostSys& H = meta[sid];
    const long gone_tiles = H.tiles;
    atoms.release(H.atom0 / kTile, H.tiles);
    lists.release(H.list0, H.listw);
    cells.release(H.cell0, H.cells);
    H = HostSys();
    hsys[sid].live = 0;
    hst[sid] = BState{};
    hipchk(hipMemcpyAsync(dst.p + sid, &hst[sid], sizeof(BState), hipMemcpyHostToDevice, st), "memcpy");
    // A system of the current plan may leave it without a new plan: its
    // device table entry says it is gone, so the control kernel keeps it
    // inactive and every kernel of the captured steps skips its tiles, sort
    // chunks and cells (each checks live, or act, first).  Its slices are free
    // for the next add, which plans again before any step runs.  That saves
    // the re-plan and the two graph instantiations (about 1.5 ms), but every
    // step still launches the gone systems' empty workgroups, so the batch
    // plans again once they pass an eighth of the live tiles (at once when
    // few systems are left).
    long live_tiles = 0;
    for (const HostSys& m : meta)
      if (m.live)
        live_tiles += m.tiles;
    dead_tiles += gone_tiles;
    if (!dirty && (size_t)sid < dsys.n && dead_tiles * 8 <= live_tiles)
      hipchk(hipMemcpyAsync(dsys.p + sid, &hsys[sid], sizeof(BSys), hipMemcpyHostToDevice, st), "memcpy");
    else
      dirty = true;
    sync();
  }

  void check_sid(long sid) const {
    if (sid < 0 || sid >= (long)meta.size() || !meta[sid].live)
      fail(LPMDX_EINVAL, "no such system in the batch");
  }

  // ---------------------------------------------------------------- launch plan
  void plan() {
    // The graph of the last plan goes first: its kernels read the tables and
    // buffers reallocated below.  The stream is idle (every run ends with a
    // synchronisation); it is waited on again so that no launch of the old
    // graph can still be in flight when it is destroyed.
    sync();
    release_graph();
    dead_tiles = 0;
    // system table and sort chunks
    // systems whose kind changed its cells per list radius get the new grid
    // and build their lists again at their next step
    std::vector<size_t> regrid;
    for (size_t s = 0; s < hsys.size(); ++s) {
      if (!meta[s].live)
        continue;
      BSys& S = hsys[s];
      const double box[3] = {S.b.L[0], S.b.L[1], S.b.L[2]};
      if (grid_div(box, meta[s].rl, div_of_kind[S.kind]) != S.g.div) {
        BSys T = S;
        set_grid(T, meta[s].rl, div_of_kind[S.kind]);
        if (T.g.ncell + 2 > meta[s].cells)
          continue; // its cell slice holds the whole-radius grid only
        S = T;
        regrid.push_back(s);
      }
    }
    if (!regrid.empty()) {
      pull_state();
      for (size_t s : regrid)
        if (hst[s].init != 1)
          hst[s].built = 0;
      hipchk(hipMemcpyAsync(dst.p, hst, hsys.size() * sizeof(BState), hipMemcpyHostToDevice, st), "memcpy");
    }
    std::vector<int2> ch;
    passes = 1;
    for (size_t s = 0; s < hsys.size(); ++s) {
      if (!meta[s].live)
        continue;
      BSys& S = hsys[s];
      S.lpa = lpa_of_kind[S.kind];
      S.ch0 = (int)ch.size();
      S.nch = (S.n + kBSortChunk - 1) / kBSortChunk;
      for (int j = 0; j < S.nch; ++j)
        ch.push_back(make_int2((int)s, j));
      passes = std::max(passes, S.passes);
    }
    nchunks = (int)ch.size();
    chunks.alloc(std::max<size_t>(ch.size(), 1));
    hist.alloc(std::max<size_t>(ch.size(), 1) * kRadix);
    if (!ch.empty())
      hipchk(hipMemcpyAsync(chunks.p, ch.data(), ch.size() * sizeof(int2), hipMemcpyHostToDevice, st), "memcpy");
    dsys.alloc(std::max<size_t>(hsys.size(), 1));
    if (!hsys.empty())
      hipchk(hipMemcpyAsync(dsys.p, hsys.data(), hsys.size() * sizeof(BSys), hipMemcpyHostToDevice, st), "memcpy");
    // tiles
    std::vector<int> ts(acap / kTile, 0), all;
    std::map<std::pair<int, int>, std::vector<int>> by_kind; // (kind, batch_exp)
    for (size_t s = 0; s < hsys.size(); ++s) {
      if (!meta[s].live)
        continue;
      const long t0 = hsys[s].a0 / kTile;
      for (long t = 0; t < meta[s].tiles; ++t) {
        ts[t0 + t] = (int)s;
        all.push_back((int)(t0 + t));
        by_kind[{hsys[s].kind, batch_exp(hsys[s])}].push_back((int)(t0 + t));
      }
    }
    hipchk(hipMemcpyAsync(tsys.p, ts.data(), ts.size() * sizeof(int), hipMemcpyHostToDevice, st), "memcpy");
    tiles_all.alloc(std::max<size_t>(all.size(), 1));
    n_all = (long)all.size();
    if (!all.empty())
      hipchk(hipMemcpyAsync(tiles_all.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, st), "memcpy");
    groups.clear();
    for (auto& kv : by_kind) {
      auto G = std::make_unique<Group>();
      G->kind = kv.first.first;
      G->exp = kv.first.second;
      G->lpa = lpa_of_kind[G->kind];
      G->tpw = tpw_of_kind[G->kind];
      G->metal = metal_kind(G->kind);
      G->n = (long)kv.second.size();
      G->tiles.alloc(kv.second.size());
      hipchk(hipMemcpyAsync(G->tiles.p, kv.second.data(), kv.second.size() * sizeof(int), hipMemcpyHostToDevice, st),
             "memcpy");
      groups.push_back(std::move(G));
    }
    // the branch streams and events of the step exist before the capture
    // starts: nothing is created while the stream is being captured
    if (groups.size() > 1)
      ensure_branches(groups.size());
    sync();
    capture();
    dirty = false;
  }

  template <int KIND, int EXP = 0> void launch_group(Group& G, hipStream_t s) {
    BArgs B = args(G.tiles.p, G.n);
    const int T = G.tpw;
    const long nb = (G.n + T - 1) / T;
    switch (G.lpa) {
    case 1:
      if (G.metal)
        k_b_rho<KIND, 1, EXP><<<nb, T * kWave, 0, s>>>(B);
      k_b_force<KIND, 1, EXP><<<nb, T * kWave, 0, s>>>(B);
      break;
    case 2:
      if (G.metal)
        k_b_rho<KIND, 2, EXP><<<nb, T * 2 * kWave, 0, s>>>(B);
      k_b_force<KIND, 2, EXP><<<nb, T * 2 * kWave, 0, s>>>(B);
      break;
    default:
      if (G.metal)
        k_b_rho<KIND, 4, EXP><<<nb, T * 4 * kWave, 0, s>>>(B);
      k_b_force<KIND, 4, EXP><<<nb, T * 4 * kWave, 0, s>>>(B);
      break;
    }
  }

  // Streams and events of the per-kind branches of a step (created once,
  // default priority).
  void ensure_branches(size_t k) {
    if (!ev_fork)
      hipchk(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
    while (bst.size() < k) {
      hipStream_t b;
      hipEvent_t e;
      hipchk(hipStreamCreateWithFlags(&b, hipStreamNonBlocking), "hipStreamCreate");
      hipchk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      bst.push_back(b);
      ev_join.push_back(e);
    }
  }

  void launch_kind(Group& G, hipStream_t s) {
    switch (G.kind) {
    case kLJ: launch_group<kLJ>(G, s); break;
    case kFastLJ: launch_group<kFastLJ>(G, s); break;
    case kMorse: launch_group<kMorse>(G, s); break;
    case kBuckingham: launch_group<kBuckingham>(G, s); break;
    case kSuttonChen:
      if (G.exp == sc_exp(9, 6))
        launch_group<kSuttonChen, sc_exp(9, 6)>(G, s);
      else if (G.exp == sc_exp(12, 6))
        launch_group<kSuttonChen, sc_exp(12, 6)>(G, s);
      else
        launch_group<kSuttonChen>(G, s);
      break;
    case kFinnisSinclair: launch_group<kFinnisSinclair>(G, s); break;
    case kGupta: launch_group<kGupta>(G, s); break;
    case kHarmonic: launch_group<kHarmonic>(G, s); break;
    case kLJMod: launch_group<kLJMod>(G, s); break;
    case kFinnisSinclairExt: launch_group<kFinnisSinclairExt>(G, s); break;
    }
  }

  void enqueue_step() {
    const int ns = (int)hsys.size();
    BArgs B = args(tiles_all.p, n_all);
    k_b_control<<<blocks_for(ns), kBlock, 0, st>>>(B, 0, -1, 0);
    if (nchunks) {
      for (int p = 0; p < passes; ++p) {
        k_b_count<<<nchunks, kBlock, 0, st>>>(B, p);
        k_b_scan<<<ns, kScanThreads, 0, st>>>(B, p);
        k_b_scatter<<<nchunks, kBlock, 0, st>>>(B, p);
      }
    }
    if (n_all) {
      k_b_gather<<<blocks_for(n_all * kTile), kBlock, 0, st>>>(B);
      k_b_commit<<<blocks_for(n_all * kTile), kBlock, 0, st>>>(B);
    }
    k_b_sorted<<<blocks_for(ns), kBlock, 0, st>>>(B);
    if (n_all) {
      k_b_build<<<blocks_for(n_all * kTile), kBlock, 0, st>>>(B);
    }
    if (groups.size() == 1) {
      launch_kind(*groups[0], st);
    } else if (groups.size() > 1) {
      // kinds touch disjoint systems: one branch each, joined before the step ends
      if (bst.size() < groups.size() || !ev_fork)
        fail(LPMDX_EDEVICE, "batch: branch streams missing (internal error)");
      hipchk(hipEventRecord(ev_fork, st), "hipEventRecord");
      for (size_t g = 0; g < groups.size(); ++g) {
        hipchk(hipStreamWaitEvent(bst[g], ev_fork, 0), "hipStreamWaitEvent");
        launch_kind(*groups[g], bst[g]);
        hipchk(hipEventRecord(ev_join[g], bst[g]), "hipEventRecord");
        hipchk(hipStreamWaitEvent(st, ev_join[g], 0), "hipStreamWaitEvent");
      }
    }
    hipchk(hipGetLastError(), "kernel launch");
  }

  // `nlite` steps without the rebuild chain: per potential kind a chain of
  // (control, rho, force) launches over that kind's systems, the kinds'
  // chains on branches of their own, joined once at the end.
  void enqueue_lite(int nlite) {
    const int ns = (int)hsys.size();
    BArgs B = args(tiles_all.p, n_all);
    if (groups.size() == 1) {
      for (int k = 0; k < nlite; ++k) {
        k_b_control<<<blocks_for(ns), kBlock, 0, st>>>(B, 1, groups[0]->kind, groups[0]->exp);
        launch_kind(*groups[0], st);
      }
    } else if (groups.size() > 1) {
      if (bst.size() < groups.size() || !ev_fork)
        fail(LPMDX_EDEVICE, "batch: branch streams missing (internal error)");
      hipchk(hipEventRecord(ev_fork, st), "hipEventRecord");
      for (size_t g = 0; g < groups.size(); ++g) {
        hipchk(hipStreamWaitEvent(bst[g], ev_fork, 0), "hipStreamWaitEvent");
        for (int k = 0; k < nlite; ++k) {
          k_b_control<<<blocks_for(ns), kBlock, 0, bst[g]>>>(B, 1, groups[g]->kind, groups[g]->exp);
          launch_kind(*groups[g], bst[g]);
        }
        hipchk(hipEventRecord(ev_join[g], bst[g]), "hipEventRecord");
        hipchk(hipStreamWaitEvent(st, ev_join[g], 0), "hipStreamWaitEvent");
      }
    }
    hipchk(hipGetLastError(), "kernel launch");
  }

  void release_graph() {
    for (hipGraphExec_t* e : {&gexec, &cexec})
      if (*e) {
        (void)hipGraphExecDestroy(*e);
        *e = nullptr;
      }
    for (hipGraph_t* g : {&graph, &cgraph})
      if (*g) {
        (void)hipGraphDestroy(*g);
        *g = nullptr;
      }
  }

  // Two graphs: one full step (control, the rebuild chain, forces), and a
  // cycle of one full step followed by cyc - 1 lite steps (no rebuild chain:
  // a system that needs a new list holds until the next full step).
  void capture() {
    release_graph();
    hipchk(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    enqueue_step();
    hipchk(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
    hipchk(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    hipchk(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    enqueue_step();
    enqueue_lite(cyc - 1);
    hipchk(hipStreamEndCapture(st, &cgraph), "hipStreamEndCapture");
    hipchk(hipGraphInstantiate(&cexec, cgraph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }

  // ---------------------------------------------------------------- running
  void pull_state() {
    if (!hsys.empty())
      hipchk(hipMemcpyAsync(hst, dst.p, hsys.size() * sizeof(BState), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
  }

  long pending_steps() {
    long m = 0;
    for (size_t s = 0; s < hsys.size(); ++s)
      if (meta[s].live)
        m = std::max(m, hst[s].left + (hst[s].init == 1 ? 1 : 0));
    return m;
  }

  // Steps between list builds of the batch's fastest-rebuilding system (its
  // mean so far), the cycle length that wastes the fewest lite steps on
  // holds; 4 before any system has rebuilt twice.
  int cycle_length() const {
    long best = 0;
    for (size_t s = 0; s < hsys.size(); ++s)
      if (meta[s].live && hst[s].builds >= 2) {
        const long iv = hst[s].done / hst[s].builds;
        best = best ? std::min(best, iv) : iv;
      }
    if (!best)
      return 4;
    return (int)std::max<long>(2, std::min<long>(best, kMaxCycle));
  }

  // graph steps still owed to the live systems by the current run
  long owed() const {
    long m = 0;
    for (size_t s = 0; s < hsys.size(); ++s)
      if (meta[s].live && !hst[s].overflow)
        m = std::max(m, std::min(hst[s].quota, hst[s].left + (hst[s].init == 1 ? 1 : 0)));
    return m;
  }

  // advances every live system by up to `steps` steps (the first step of a new
  // system is its initial force evaluation): each system gets a quota of
  // min(steps, its pending steps), spent in cycles of one full step and
  // cyc - 1 lite steps; systems that held in a lite step are owed steps, and
  // more cycles run until every quota is spent.
  void run(long steps) {
    if (dirty) {
      // the cycle length first, so that the plan's capture is the only one
      // of this run
      if (!hsys.empty()) {
        pull_state();
        const int want = cycle_length();
        if (want < cyc || want >= cyc + 2)
          cyc = want;
      }
      plan();
    }
    if (hsys.empty())
      return;
    pull_state();
    const long todo = std::min(steps, pending_steps());
    if (todo <= 0)
      return;
    const int want = cycle_length();
    if (want < cyc || want >= cyc + 2) {
      cyc = want;
      sync();
      capture();
    }
    for (size_t s = 0; s < hsys.size(); ++s)
      hst[s].quota = meta[s].live ? todo : 0;
    hipchk(hipMemcpyAsync(dst.p, hst, hsys.size() * sizeof(BState), hipMemcpyHostToDevice, st), "memcpy");
    long rem = todo;
    while (rem > 0) {
      for (long k = 0; k < rem / cyc; ++k)
        hipchk(hipGraphLaunch(cexec, st), "hipGraphLaunch");
      for (long k = 0; k < rem % cyc; ++k)
        hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
      steps_run += (rem / cyc) * cyc + rem % cyc;
      pull_state();
      rem = owed();
    }
    std::string bad;
    for (size_t s = 0; s < hsys.size(); ++s) {
      if (meta[s].live && hst[s].overflow)
        bad += " " + std::to_string(s);
    }
    if (!bad.empty())
      fail(LPMDX_ECAPACITY, "neighbour list capacity exceeded in system(s):" + bad);
  }

  // ---------------------------------------------------------------- results
  void get(long sid, double* xo, double* vo, double* ao, long* info) {
    check_sid(sid);
    pull_state();
    const long n = meta[sid].n, a0 = hsys[sid].a0;
    const BState& sst = hst[sid];
    std::vector<int> pm(n);
    std::vector<double> tx(3 * n), tv(3 * n), ta(3 * n);
    const double* xs = (sst.upar ? x1.p : x0.p) + 3 * a0;
    hipchk(hipMemcpyAsync(pm.data(), perm.p + a0, n * sizeof(int), hipMemcpyDeviceToHost, st), "memcpy");
    hipchk(hipMemcpyAsync(tx.data(), xs, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    hipchk(hipMemcpyAsync(tv.data(), v.p + 3 * a0, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    hipchk(hipMemcpyAsync(ta.data(), a.p + 3 * a0, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
    for (long i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q) {
        if (xo)
          xo[3 * pm[i] + q] = tx[3 * i + q];
        if (vo)
          vo[3 * pm[i] + q] = tv[3 * i + q];
        if (ao)
          ao[3 * pm[i] + q] = ta[3 * i + q];
      }
    if (info) {
      info[0] = sst.done;
      info[1] = sst.left;
      info[2] = sst.builds;
      info[3] = sst.init == 1 ? 0 : 1; // 2: the first force evaluation has run
      info[4] = (long)sst.maxcnt;
      info[5] = hsys[sid].cap;
      info[6] = hsys[sid].lpa;
      info[7] = sst.overflow;
    }
  }

  template <int KIND> void eval_k(long sid, int lpa, long nb, double* pa, int rho_pass) {
    BArgs B = args(nullptr, 0);
    switch (lpa) {
    case 1: k_b_eval<KIND, 1><<<nb, kWave, 0, st>>>(B, (int)sid, evala.p, pa, rho_pass); break;
    case 2: k_b_eval<KIND, 2><<<nb, 2 * kWave, 0, st>>>(B, (int)sid, evala.p, pa, rho_pass); break;
    default: k_b_eval<KIND, 4><<<nb, 4 * kWave, 0, st>>>(B, (int)sid, evala.p, pa, rho_pass); break;
    }
  }

  // potential energy, virial, kinetic energy of one system at its last step
  void energies(long sid, double out[3]) {
    check_sid(sid);
    if (dirty)
      plan();
    pull_state();
    if (hst[sid].init == 1)
      fail(LPMDX_ESTATE, "the system has not run its first force evaluation yet");
    const BSys& S = hsys[sid];
    const long n = meta[sid].n;
    const int lpa = S.lpa;
    const long nb = S.ntile; // one workgroup per tile
    evala.alloc(3 * std::max<long>(n, 1));
    partials.alloc(nb * 12 * 2);
    double* p1 = partials.p;
    double* p2 = partials.p + nb * 12;
    auto launch = [&](double* pa, int rho_pass) {
      switch (S.kind) {
      case kLJ: eval_k<kLJ>(sid, lpa, nb, pa, rho_pass); break;
      case kFastLJ: eval_k<kFastLJ>(sid, lpa, nb, pa, rho_pass); break;
      case kMorse: eval_k<kMorse>(sid, lpa, nb, pa, rho_pass); break;
      case kBuckingham: eval_k<kBuckingham>(sid, lpa, nb, pa, rho_pass); break;
      case kSuttonChen: eval_k<kSuttonChen>(sid, lpa, nb, pa, rho_pass); break;
      case kFinnisSinclair: eval_k<kFinnisSinclair>(sid, lpa, nb, pa, rho_pass); break;
      case kGupta: eval_k<kGupta>(sid, lpa, nb, pa, rho_pass); break;
      case kHarmonic: eval_k<kHarmonic>(sid, lpa, nb, pa, rho_pass); break;
      case kLJMod: eval_k<kLJMod>(sid, lpa, nb, pa, rho_pass); break;
      case kFinnisSinclairExt: eval_k<kFinnisSinclairExt>(sid, lpa, nb, pa, rho_pass); break;
      }
    };
    double e_emb = 0.0;
    if (S.metal) {
      launch(p2, 1);
      k_reduce_partials<<<1, kBlock, 0, st>>>(p2, nb, 12, red.p);
      hipchk(hipMemcpyAsync(hred, red.p, 12 * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
      sync();
      e_emb = hred[1];
    }
    launch(p1, 0);
    k_reduce_partials<<<1, kBlock, 0, st>>>(p1, nb, 12, red.p);
    // kinetic energy of the system's slice
    k_kinetic<<<blocks_for(n), kBlock, 0, st>>>(v.p + 3 * S.a0, mass.p + S.a0, n, p2);
    k_reduce_partials<<<1, kBlock, 0, st>>>(p2, blocks_for(n), 4, red.p + 12);
    hipchk(hipGetLastError(), "kernel launch");
    hipchk(hipMemcpyAsync(hred, red.p, 16 * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
    out[0] = hred[0] + e_emb;
    out[1] = hred[2];
    out[2] = hred[12] * kKin2eV;
  }

  // ---------------------------------------------------------------- autotuner
  // Per potential kind, the lanes per atom of its force launches (1, 2, 4)
  // and the cells per list radius of its grids (1, 2: how many candidates a
  // list build visits against how many cell ranges it reads), chosen by
  // timing a few steps of this batch per candidate pair, then the tiles per
  // workgroup of its force and rho launches (1, 2, 4) with that pair; on a
  // snapshot of the batch's state, restored afterwards, so tuning leaves the
  // trajectories alone.
  void autotune(long trial_steps, int out_lpa[kNumKinds + 1], int out_div[kNumKinds + 1],
                int out_tpw[kNumKinds + 1]) {
    if (dirty)
      plan();
    pull_state();
    const size_t A = acap;
    DBuf<double> bx0, bx1, bv, ba, bffm, bmass;
    DBuf<int> bz, bperm;
    DBuf<BState> bst;
    for (DBuf<double>* b : {&bx0, &bx1, &bv, &ba})
      b->alloc(3 * A);
    for (DBuf<double>* b : {&bffm, &bmass})
      b->alloc(A);
    bz.alloc(A);
    bperm.alloc(A);
    bst.alloc(std::max<size_t>(hsys.size(), 1));
    auto copy = [&](bool save) {
      auto cp = [&](void* d, void* s, size_t bytes) {
        hipchk(hipMemcpyAsync(save ? d : s, save ? s : d, bytes, hipMemcpyDeviceToDevice, st), "memcpy");
      };
      cp(bx0.p, x0.p, 3 * A * sizeof(double));
      cp(bx1.p, x1.p, 3 * A * sizeof(double));
      cp(bv.p, v.p, 3 * A * sizeof(double));
      cp(ba.p, a.p, 3 * A * sizeof(double));
      cp(bffm.p, ffm.p, A * sizeof(double));
      cp(bmass.p, mass.p, A * sizeof(double));
      cp(bz.p, z.p, A * sizeof(int));
      cp(bperm.p, perm.p, A * sizeof(int));
      cp(bst.p, dst.p, hsys.size() * sizeof(BState));
      sync();
    };
    copy(true);
    std::vector<BState> saved(hst, hst + hsys.size());
    // every trial starts from the saved state with its lists invalidated,
    // so each candidate pays for one build, as a steady run would
    auto reset = [&]() {
      copy(false);
      for (size_t s = 0; s < hsys.size(); ++s) {
        hst[s] = saved[s];
        hst[s].built = 0;
        hst[s].left = std::max<long>(hst[s].left, trial_steps + 1);
        hst[s].quota = trial_steps + 8;
      }
      hipchk(hipMemcpyAsync(dst.p, hst, hsys.size() * sizeof(BState), hipMemcpyHostToDevice, st), "memcpy");
      sync();
    };
    std::vector<int> kinds; // each kind once (its exponent groups share its shapes)
    for (auto& G : groups)
      if (std::find(kinds.begin(), kinds.end(), G->kind) == kinds.end())
        kinds.push_back(G->kind);
    for (int kind : kinds) {
      double best = 1e300;
      int best_lpa = lpa_of_kind[kind], best_div = div_of_kind[kind];
      for (int dv : {1, 2})
        for (int cand : {1, 2, 4}) {
          lpa_of_kind[kind] = cand;
          div_of_kind[kind] = dv;
          plan();
          reset();
          // warm up (includes the build), then time full steps, each with
          // the rebuild chain a build would run
          hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
          hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
          sync();
          auto t0 = std::chrono::steady_clock::now();
          for (long k = 0; k < trial_steps; ++k)
            hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
          sync();
          double dt_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
          if (dt_s < best) {
            best = dt_s;
            best_lpa = cand;
            best_div = dv;
          }
        }
      lpa_of_kind[kind] = best_lpa;
      div_of_kind[kind] = best_div;
      // then the tile shape with those lanes and cells
      int best_tpw = tpw_of_kind[kind];
      for (int T : {1, 2, 4}) {
        if (T == best_tpw)
          continue; // timed above
        tpw_of_kind[kind] = T;
        plan();
        reset();
        hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
        hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
        sync();
        auto t0 = std::chrono::steady_clock::now();
        for (long k = 0; k < trial_steps; ++k)
          hipchk(hipGraphLaunch(gexec, st), "hipGraphLaunch");
        sync();
        double dt_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (dt_s < best) {
          best = dt_s;
          best_tpw = T;
        }
      }
      tpw_of_kind[kind] = best_tpw;
    }
    plan();
    copy(false);
    for (size_t s = 0; s < hsys.size(); ++s) {
      hst[s] = saved[s];
      hst[s].built = 0; // lists are rebuilt for the chosen layout
    }
    hipchk(hipMemcpyAsync(dst.p, hst, hsys.size() * sizeof(BState), hipMemcpyHostToDevice, st), "memcpy");
    sync();
    for (int q = 0; q <= kNumKinds; ++q) {
      if (out_lpa)
        out_lpa[q] = lpa_of_kind[q];
      if (out_div)
        out_div[q] = div_of_kind[q];
      if (out_tpw)
        out_tpw[q] = tpw_of_kind[q];
    }
  }

  static void need(const void* p, const char* what) {
    if (!p)
      fail(LPMDX_EINVAL, std::string(what) + " is NULL");
  }
  static void finite(const double* p, long m, const char* what) {
    for (long q = 0; q < m; ++q)
      if (!std::isfinite(p[q]))
        fail(LPMDX_ENONFINITE, std::string(what) + " holds a non-finite value");
  }
};

extern "C" {

int lpmdx_batch_create(lpmdx_batch** out, long atom_capacity, long list_words, long cell_slots) {
  return guard([&] {
    if (!out)
      fail(LPMDX_EINVAL, "out is NULL");
    *out = nullptr;
    if (atom_capacity <= 0 || list_words <= 0 || cell_slots <= 0)
      fail(LPMDX_EINVAL, "arena capacities must be positive");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0)
      fail(LPMDX_EDEVICE, "no HIP device available");
    auto B = std::make_unique<lpmdx_batch>();
    hipchk(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking), "hipStreamCreate");
    B->create(atom_capacity, list_words, cell_slots);
    *out = B.release();
  });
}

int lpmdx_batch_destroy(lpmdx_batch* b) {
  return guard([&] { delete b; });
}

int lpmdx_batch_add(lpmdx_batch* b, long n, const double box[3], const int* z, const double* mass,
                    const double* x, const double* v, const char* cm_args, int npot,
                    const char** names, const char** args, const int* s1, const int* s2, double dt,
                    long steps, long* sid) {
  return guard([&] {
    if (!b || !sid)
      fail(LPMDX_EINVAL, "batch or sid is NULL");
    if (!names)
      fail(LPMDX_EINVAL, "names is NULL");
    *sid = b->add(n, box, z, mass, x, v, cm_args, npot, names, args, s1, s2, dt, steps);
  });
}

int lpmdx_batch_remove(lpmdx_batch* b, long sid) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    b->remove(sid);
  });
}

int lpmdx_batch_run(lpmdx_batch* b, long steps) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    if (steps < 0)
      fail(LPMDX_EINVAL, "negative step count");
    b->run(steps);
  });
}

int lpmdx_batch_get(lpmdx_batch* b, long sid, double* x, double* v, double* a, long info[8]) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    b->get(sid, x, v, a, info);
  });
}

int lpmdx_batch_energies(lpmdx_batch* b, long sid, double out[3]) {
  return guard([&] {
    if (!b || !out)
      fail(LPMDX_EINVAL, "batch or out is NULL");
    b->energies(sid, out);
  });
}

int lpmdx_batch_autotune(lpmdx_batch* b, long trial_steps, int lanes_out[LPMDX_NKINDS + 1],
                         int cells_out[LPMDX_NKINDS + 1], int tiles_out[LPMDX_NKINDS + 1]) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    if (trial_steps <= 0)
      fail(LPMDX_EINVAL, "trial steps must be positive");
    b->autotune(trial_steps, lanes_out, cells_out, tiles_out);
  });
}

int lpmdx_batch_set_tiles(lpmdx_batch* b, int kind, int tiles_per_group) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    if (kind < 1 || kind > kNumKinds || !(tiles_per_group == 1 || tiles_per_group == 2 || tiles_per_group == 4))
      fail(LPMDX_EINVAL, "kind 1.." + std::to_string(kNumKinds) + " and 1, 2 or 4 tiles per workgroup");
    b->tpw_of_kind[kind] = tiles_per_group;
    b->dirty = true; // launch shapes only: the lists stay
  });
}

int lpmdx_batch_set_cells(lpmdx_batch* b, int kind, int per_radius) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    if (kind < 1 || kind > kNumKinds || !(per_radius == 1 || per_radius == 2))
      fail(LPMDX_EINVAL, "kind 1.." + std::to_string(kNumKinds) + " and 1 or 2 cells per list radius");
    b->div_of_kind[kind] = per_radius;
    b->dirty = true; // the next plan gives the kind's systems their grids and new lists
  });
}

int lpmdx_batch_set_lanes(lpmdx_batch* b, int kind, int lanes) {
  return guard([&] {
    if (!b)
      fail(LPMDX_EINVAL, "batch is NULL");
    if (kind < 1 || kind > kNumKinds || !(lanes == 1 || lanes == 2 || lanes == 4))
      fail(LPMDX_EINVAL, "kind 1.." + std::to_string(kNumKinds) + " and lanes 1, 2 or 4");
    b->lpa_of_kind[kind] = lanes;
    b->pull_state();
    for (size_t s = 0; s < b->hsys.size(); ++s)
      if (b->meta[s].live && b->hsys[s].kind == kind)
        b->hst[s].built = 0;
    hipchk(hipMemcpyAsync(b->dst.p, b->hst, b->hsys.size() * sizeof(BState), hipMemcpyHostToDevice, b->st),
           "memcpy");
    b->sync();
    b->dirty = true;
  });
}

int lpmdx_batch_info(lpmdx_batch* b, long out[8]) {
  return guard([&] {
    if (!b || !out)
      fail(LPMDX_EINVAL, "batch or out is NULL");
    long live = 0, atoms_used = 0;
    for (auto& m : b->meta)
      if (m.live) {
        ++live;
        atoms_used += m.n;
      }
    out[0] = live;
    out[1] = atoms_used;
    out[2] = b->acap;
    out[3] = b->atoms.largest() * kTile;
    out[4] = b->lists.largest() * kTile;
    out[5] = b->cells.largest();
    out[6] = b->steps_run;
    out[7] = (long)b->hsys.size();
  });
}

long lpmdx_batch_bytes_per_atom(void) { return kBytesPerAtom; }

int lpmdx_batch_progress(lpmdx_batch* b, long nslots, long* left, int* flags) {
  return guard([&] {
    if (!b || (nslots > 0 && (!left || !flags)))
      fail(LPMDX_EINVAL, "batch, left or flags is NULL");
    b->pull_state();
    for (long s = 0; s < nslots; ++s) {
      left[s] = 0;
      flags[s] = 0;
      if (s >= (long)b->hsys.size() || !b->meta[s].live)
        continue;
      const BState& q = b->hst[s];
      const int pending = q.init == 1 ? 1 : 0;
      left[s] = q.left + pending;
      flags[s] = 1 | (q.overflow ? 2 : 0) | (pending ? 4 : 0);
    }
  });
}

} // extern "C"

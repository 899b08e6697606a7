// Host engine and C ABI of the MI355X LPMD force-and-integration path.
//
// One lpmdx_system is one Configuration resident in HBM (positions, velocities,
// accelerations in AoS FP64, internally kept in cell-sorted order), with its
// cell manager, its potentials and its own HIP stream.  The entry points are
// declared in include/lpmdx.h, each next to the reference interface it serves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lpmdx.h"
#include "cells.hpp"
#include "common.hpp"
#include "forces.hpp"
#include "host_util.hpp"
#include "sort.hpp"

using namespace lpmdx;

namespace lpmdx {
thread_local std::string g_err;
}

namespace {


// timing slots of lpmdx_system_profile_read
enum ProfSlot { kProfForce = 0, kProfRho = 1, kProfBuild = 2, kProfOther = 3, kProfSlots = 4 };

// max |x - xb|^2 over atoms through the minimum image (xb: positions at the build)
__global__ __launch_bounds__(kBlock) void k_disp(ForceArgs A) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float d2 = 0.0f;
  if (i < A.n) {
    double dx = A.x[3 * i] - A.xn[3 * i];
    double dy = A.x[3 * i + 1] - A.xn[3 * i + 1];
    double dz = A.x[3 * i + 2] - A.xn[3 * i + 2];
    min_image(A.b, dx, dy, dz);
    d2 = (float)sqmod(dx, dy, dz);
  }
  float m = wave_max(d2);
  if ((threadIdx.x & (kWave - 1)) == 0 && m > 0.0f)
    atomic_max_sparse(A.maxd2, __float_as_uint(m));
}

// Start of a chunk of steps: the per-step displacement slots md[0, m) cleared.
__global__ __launch_bounds__(kBlock) void k_chunk_begin(unsigned int* __restrict__ md, long m) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t < m)
    md[t] = 0u;
}

// Integrator::GoBack (liblpmd/simulation/integrator.cc:31-42) for the leapfrog
// and Beeman set-ups: x and a saved, vh = v - a dt (when vh is given),
// x = FittedInside(x - v dt).
__global__ __launch_bounds__(kBlock) void k_lf_goback(double* __restrict__ x, const double* __restrict__ v,
                                                      const double* __restrict__ a, double* __restrict__ keep_x,
                                                      double* __restrict__ keep_a, double* __restrict__ vh, long n,
                                                      double dt, double l0, double l1, double l2) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const double ll[3] = {l0, l1, l2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double xq = x[3 * i + q], vq = v[3 * i + q], aq = a[3 * i + q];
    keep_x[3 * i + q] = xq;
    keep_a[3 * i + q] = aq;
    if (vh)
      vh[3 * i + q] = vq - aq * dt;
    x[3 * i + q] = fitted_inside(xq - vq * dt, ll[q]);
  }
}

// leapfrog.cc:63-66: vh += (a_old * 0.5) * dt
__global__ __launch_bounds__(kBlock) void k_lf_halfkick(double* __restrict__ vh, const double* __restrict__ a,
                                                        long n, double dt) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    vh[3 * i + q] += (a[3 * i + q] * 0.5) * dt;
}

// SimulationEngine::SetTemperature (liblpmd/simulation/simulationbuilder.cc:53-71)
// with the kinetic sum already reduced on the device (red[0] = sum of
// 0.5 m v^2): T_now = (2/3) K / (kB N), K = KIN2EV * red[0]
// (properties.h:22-53), then v *= sqrt(T / T_now).  A system at rest is left
// as it is (the reference would divide by zero).
// gate / gate_b: as for the force kernels of a speculative step.
__global__ __launch_bounds__(kBlock) void k_set_temperature(double* __restrict__ v, const double* __restrict__ red,
                                                            long n, double temp, double q,
                                                            const unsigned int* gate, double gate_b) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  if (gate && (*gate == kGateClosed || sqrt((double)__uint_as_float(*gate)) > gate_b))
    return;
  const double ti = (2.0 / 3.0) * (red[0] * kKin2eV) / (kBoltzmann * double(n));
  if (!(ti > 0.0))
    return;
  // q < 0: SetTemperature; q = dt / tau: the berendsen modifier's factor
  // (plugins/berendsen.cc:62-70)
  const double f = q < 0.0 ? sqrt(temp / ti) : sqrt(1.0 + q * (temp / ti - 1.0));
#pragma unroll
  for (int q = 0; q < 3; ++q)
    v[3 * i + q] *= f;
}

__global__ void k_fill_double(double* p, long m, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m)
    p[i] = v;
}

} // namespace

// ---------------------------------------------------------------------------
// Launches of the brick-staged kernels: one workgroup per brick.  Dynamic LDS
// above 64 KiB has to be allowed per kernel; each instantiation does so once.
struct BrickLaunch {
  long bricks;
  int threads;
  size_t lds;
  hipStream_t st;
  const int2* bhalo;
  int brick;
  int cap;
};

inline bool allow_stage_lds(const void* kernel) {
  hipchk(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, kStageLdsLimit),
         "hipFuncSetAttribute");
  return true;
}

template <int KIND, int LPA, bool EVAL> struct BrickKernels {
  template <bool SPC> static void pair(const BrickLaunch& L, const ForceArgs& A, int parts) {
    static const bool ok = allow_stage_lds((const void*)k_pair_force_bk<KIND, LPA, EVAL, SPC>);
    (void)ok;
    k_pair_force_bk<KIND, LPA, EVAL, SPC><<<L.bricks, L.threads, L.lds, L.st>>>(A, L.bhalo, L.brick, L.cap, parts);
  }
  template <int EXP = 0> static void rho(const BrickLaunch& L, const ForceArgs& A, float* w, double* rho_out) {
    static const bool ok = allow_stage_lds((const void*)k_eam_rho_bk<KIND, LPA, EVAL, EXP>);
    (void)ok;
    k_eam_rho_bk<KIND, LPA, EVAL, EXP><<<L.bricks, L.threads, L.lds, L.st>>>(A, L.bhalo, L.brick, L.cap, w, rho_out);
  }
  template <bool SPC, int EXP = 0> static void eam(const BrickLaunch& L, const ForceArgs& A, int parts) {
    static const bool ok = allow_stage_lds((const void*)k_eam_force_bk<KIND, LPA, EVAL, SPC, EXP>);
    (void)ok;
    k_eam_force_bk<KIND, LPA, EVAL, SPC, EXP><<<L.bricks, L.threads, L.lds, L.st>>>(A, L.bhalo, L.brick, L.cap,
                                                                                      parts);
  }
};

// ---------------------------------------------------------------------------
// the system

struct lpmdx_system {
  long n = 0;
  Box box{};
  double ll[3] = {0, 0, 0};
  hipStream_t st = nullptr;
  // state, internal order
  DBuf<double> x, xn, v, a, ffm, mass, asum, xb;
  DBuf<int> z, perm; // perm[k]: caller index of the atom in slot k
  // scratch for re-ordering
  DBuf<double> x2, v2, a2, ffm2, mass2;
  DBuf<int> z2, perm2;
  // sort
  DBuf<uint32_t> keys, vals, keys2, vals2, hist, htot;
  DBuf<unsigned long long> sumslots; // spread list-length sum of a build
  // cells
  DBuf<int> cb;           // cell starts of the sorted atoms (ncell + 2)
  DBuf<int2> bhalo;       // brick halo table (bricked grids)
  bool staged = false;    // lists hold brick halo slots (brick-staged force kernels)
  int halo_cap = 0;       // LDS slots of the staged kernels: >= the largest halo
  int brick_threads = 0;  // workgroup size of the staged kernels
  long nbrick = 0;
  Grid grid{};
  // lists
  DBuf<int> nbr, cnt;
  DBuf<float4> xf;
  long stride = 0;
  int cap = 0;
  int lpa = 1, lpa_req = 0;
  double rl = 0.0; // list radius
  double rl_min = 0.0; // a list radius an export asked for (minimumimage, lcbinary)
  // EAM
  DBuf<float> w_rho;
  DBuf<double4> xw; // EAM: packed (x, y, z, rho^{-1/2}) for the unstaged force pass
  DBuf<double> rho;
  // reductions
  DBuf<double> partials, partials2, red;
  DBuf<Ctrl> ctrl;
  Ctrl* hctrl = nullptr; // pinned
  double* hred = nullptr;
  // host-side model
  CellMgr cm;
  std::vector<std::unique_ptr<PotParams>> pots;
  bool list_built = false;
  bool forces_ok = false; // a = F(x) for the user-visible positions
  bool drifted = false;   // user positions are in xn; x = drift of them
  // leapfrog (plugins/leapfrog.cc): half-step velocities v(t - dt/2), in
  // internal order, once lpmdx_leapfrog_init has set them
  DBuf<double> vh, vh2;
  bool lf_ready = false;
  // Beeman: the acceleration of the step before a's (ab) and its reorder
  // copy; drift_beeman says which scheme made the look-ahead x
  DBuf<double> ab, ab2;
  bool bm_ready = false, drift_beeman = false;
  double bm_dt = 0.0;
  double lf_dt = 0.0;
  // more arrays that reorders carry along (with their scratch twins)
  std::vector<std::pair<DBuf<double>*, DBuf<double>*>> carry;
  DBuf<double> keep_x, keep_x2, keep_a, keep_a2;
  long pos_version = 0;   // bumped whenever the user positions change
  long cell_version = -1; // pos_version at the last UpdateCell
  double disp_cur = 0.0; // largest |x - xb| of the positions of the next force evaluation
  long chunk = 8;        // steps of the next velocity-Verlet chunk
  long n_cut_chunks = 0; // chunks cut short by the skin test
  double last_e[12] = {0};
  bool last_e_valid = false;
  long n_builds = 0, n_forces = 0, n_steps = 0, max_list = 0, sum_list = 0;
  // the tempscaling modifier run after every step it is active in
  // (lpmd/application.cc:404-459, plugins/tempscaling.cc:58-63); steps are
  // counted by n_steps, the simulation's CurrentStep
  struct TempSchedule {
    bool on = false;
    double from = 0.0, to = 0.0;
    long start = 0, end = -1, each = 1;
    // Stepper::IsActiveInStep (liblpmd/lpmd/stepper.h:17-25)
    bool active(long k) const {
      if (!on || k < start)
        return false;
      if (end != -1 && k > end)
        return false;
      return ((k - start) % each) == 0;
    }
    // Stepper::ValueAtStep (stepper.h:27-31)
    double value(long k) const {
      if (start == end)
        return from;
      return from + double(k - start) / double(end - start) * (to - from);
    }
  } tsched;

  // the berendsen modifier (plugins/berendsen.cc:14-80): its first active
  // step only arms it (the stop step = that step + tau / dt, each = 1), then
  // every step up to the stop step scales v by
  // sqrt(1 + (dt / tau) (T_set / T - 1)), and the step after it disarms it
  // (each back to its value).  The state moves with the step numbers only,
  // so the host advances it as it launches steps.
  struct Berendsen {
    bool on = false;
    double from = 0.0, to = 0.0, tau = 400.0;
    long start = 0, end = -1, each = 1, old_each = 1, stop = -1;
    bool active(long k) const {
      if (!on || k < start)
        return false;
      if (end != -1 && k > end)
        return false;
      return ((k - start) % each) == 0;
    }
    // Apply at step k: true and the target when it scales
    bool apply(long k, double dt, double& set_temp) {
      if (!active(k))
        return false;
      set_temp = start == end ? from : from + double(k - start) / double(end - start) * (to - from);
      if (stop == -1) {
        stop = long(double(k) + tau / dt);
        old_each = each;
        each = 1;
        return false;
      }
      if (k <= stop)
        return true;
      stop = -1;
      each = old_each;
      return false;
    }
  } bsched;
  // optional per-kernel timing with HIP events on this system's stream
  bool profiling = false;
  struct Pending {
    int slot;
    hipEvent_t a, b;
    long step; // the integrator step the launch belongs to, or -1
  };
  std::vector<Pending> pending;
  long prof_step = -1; // tag of the launches being recorded

  // Timings of speculative steps that never ran (their kernels returned at
  // once) leave the profile: every entry tagged with a step >= first.
  void prof_discard_from(long first) {
    size_t k = 0;
    for (auto& p : pending) {
      if (p.step >= first) {
        event_pool.push_back(p.a);
        event_pool.push_back(p.b);
      } else {
        pending[k++] = p;
      }
    }
    pending.resize(k);
  }
  std::vector<hipEvent_t> event_pool;
  double prof_ms[kProfSlots] = {0};
  long prof_n[kProfSlots] = {0};

  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    hipchk(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  // brackets work enqueued between begin() and end() on the stream
  hipEvent_t prof_begin() {
    if (!profiling)
      return nullptr;
    hipEvent_t e = take_event();
    hipchk(hipEventRecord(e, st), "hipEventRecord");
    return e;
  }
  void prof_end(int slot, hipEvent_t a) {
    if (!profiling || !a)
      return;
    hipEvent_t e = take_event();
    hipchk(hipEventRecord(e, st), "hipEventRecord");
    pending.push_back({slot, a, e, prof_step});
  }
  void prof_collect() {
    for (auto& p : pending) {
      float ms = 0.0f;
      hipchk(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      prof_ms[p.slot] += ms;
      prof_n[p.slot] += 1;
      event_pool.push_back(p.a);
      event_pool.push_back(p.b);
    }
    pending.clear();
  }

  ~lpmdx_system() {
    if (st)
      (void)hipStreamSynchronize(st);
    for (auto& p : pending) {
      event_pool.push_back(p.a);
      event_pool.push_back(p.b);
    }
    for (auto e : event_pool)
      (void)hipEventDestroy(e);
    if (hctrl)
      (void)hipHostFree(hctrl);
    if (hred)
      (void)hipHostFree(hred);
    pots.clear();
    if (st)
      (void)hipStreamDestroy(st);
  }

  void sync() {
    hipchk(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (!pending.empty())
      prof_collect();
  }

  void launch_check() { hipchk(hipGetLastError(), "kernel launch"); }

  // ------------------------------------------------------------------ lists
  // Effective filter of each potential, following how the reference's managers
  // serve a potential: linkedcell caches each atom's half (full) list with the
  // rcut of the first request after UpdateCell (linkedcell.cc:315-354);
  // verletlist filters its stored list with the caller's rcut (verletlist.cc:151,163).
  double half_list_rcut() const { return pots.empty() ? cm.cutoff : pots[0]->cutoff; }
  double full_list_rcut() const {
    for (auto& p : pots)
      if (p->metal())
        return p->cutoff;
    return half_list_rcut();
  }

  // Cell managers whose lists come from the reference's linked-cell stencil.
  bool stencil_grid() const { return cm.kind == kCmLinkedCell || cm.kind == kCmLcBinary; }
  // Managers that rebuild on every UpdateCell: linkedcell and lcbinary (their
  // UpdateCell re-bins), minimumimage without a skin (it lists afresh on every
  // request).
  bool rebuilds_every_update() const {
    return stencil_grid() || (cm.kind == kCmMinImage && cm.skin < 0.0);
  }

  double list_radius() const {
    if (cm.kind == kCmLinkedCell)
      return std::max(half_list_rcut(), full_list_rcut());
    if (cm.kind == kCmMinImage || cm.kind == kCmLcBinary) {
      // both serve each potential with its own cutoff (minimumimage.cc:74-100,
      // lcbinary.cc:264-306); an export may ask for a wider radius (rl_min)
      double r = pots.empty() ? cm.cutoff : 0.0;
      for (auto& p : pots)
        r = std::max(r, p->cutoff);
      r = std::max(r, rl_min);
      return (cm.kind == kCmMinImage && cm.skin >= 0.0) ? r + cm.skin : r;
    }
    return cm.skin >= 0.0 ? cm.cutoff + cm.skin : cm.extcut;
  }

  // Largest separation a listed pair can have at a force evaluation: the
  // list radius padded as the FP32 build pads it, plus twice the displacement
  // the skin test allows since the build; 0 when that reaches half an edge.
  double wrap_margin() const {
    const double lmax = std::max(box.L[0], std::max(box.L[1], box.L[2]));
    const double hmin = std::min(box.half[0], std::min(box.half[1], box.half[2]));
    const double pos_err = 2.0 * std::ldexp(lmax, -24);
    const double m = rl * (1.0 + 1e-6) + 4.0 * pos_err + 2.0 * budget() + 1e-9 * lmax;
    // below half an edge with room to spare: the force kernels' rint wrap
    // (min_image_near) needs it
    return m < hmin * (1.0 - 1e-6) ? m : 0.0;
  }

  // Exactness budget of the Verlet list: half the gap between the list radius
  // and the largest finite filter radius below it.
  double budget() const {
    double rmax = 0.0;
    for (auto& p : pots)
      if (p->cutoff < rl)
        rmax = std::max(rmax, p->cutoff);
    if (pots.empty())
      rmax = cm.cutoff;
    return 0.5 * (rl - rmax);
  }

  void setup_grid() {
    if (cm.kind == kCmLinkedCell) {
      if (!cm.grid_fixed) {
        // UpdateCell's first call fixes nx ny nz and the stencil (linkedcell.cc:146-241)
        if (cm.mode_auto) {
          cm.nx = lc_auto_n(edge_module(box.L[0]), cm.cutoff);
          cm.ny = lc_auto_n(edge_module(box.L[1]), cm.cutoff);
          cm.nz = lc_auto_n(edge_module(box.L[2]), cm.cutoff);
        }
        if (cm.nx <= 0 || cm.ny <= 0 || cm.nz <= 0)
          fail(LPMDX_EINVAL, "linkedcell: nx, ny and nz must be positive (use mode auto or set them)");
        double d = edge_module(box.L[0]) / double(cm.nx);
        if (edge_module(box.L[1]) / double(cm.ny) < d)
          d = edge_module(box.L[1]) / double(cm.ny);
        if (edge_module(box.L[2]) / double(cm.nz) < d)
          d = edge_module(box.L[2]) / double(cm.nz);
        cm.side = int(ceil((cm.cutoff / d) - 0.5));
        cm.grid_fixed = true;
      }
      grid.mode = kLinkedCell;
      grid.nx = cm.nx;
      grid.ny = cm.ny;
      grid.nz = cm.nz;
      grid.side = cm.side;
    } else if (cm.kind == kCmLcBinary) {
      if (!cm.grid_fixed) {
        // lcbinary.cc:88-145 (auto mode picks n once; the later growth keeps it)
        if (cm.mode_auto) {
          cm.nx = lcb_auto_n(edge_module(box.L[0]), cm.cutoff);
          cm.ny = lcb_auto_n(edge_module(box.L[1]), cm.cutoff);
          cm.nz = lcb_auto_n(edge_module(box.L[2]), cm.cutoff);
          cm.mode_auto = false;
        }
        if (cm.nx <= 0 || cm.ny <= 0 || cm.nz <= 0)
          fail(LPMDX_EINVAL, "lcbinary: nx, ny and nz must be positive");
        lcb_side();
        cm.grid_fixed = true;
      }
      grid.mode = kLinkedCell;
      grid.nx = cm.nx;
      grid.ny = cm.ny;
      grid.nz = cm.nz;
      grid.side = cm.side;
    } else {
      // cells no smaller than the list radius (with a margin for rounding)
      double r = rl * (1.0 + 1e-9);
      int nq[3];
      for (int q = 0; q < 3; ++q) {
        double m = std::isfinite(r) && r > 0 ? std::floor(box.L[q] / r) : 1.0;
        nq[q] = (int)std::max(1.0, std::min(m, 1024.0));
      }
      grid.mode = kExact;
      grid.nx = nq[0];
      grid.ny = nq[1];
      grid.nz = nq[2];
      grid.side = 1;
    }
    long nc = (long)grid.nx * grid.ny * grid.nz;
    if (nc > (1L << 28))
      fail(LPMDX_EINVAL, "cell grid too large");
    grid.ncell = (int)nc;
    grid.brick = grid.mode == kExact ? choose_brick() : 0;
    grid.nbx = grid.nby = grid.nbz = 0;
    nbrick = 0;
    if (grid.brick > 0) {
      const int B = grid.brick;
      grid.nbx = (grid.nx + B - 1) / B;
      grid.nby = (grid.ny + B - 1) / B;
      grid.nbz = (grid.nz + B - 1) / B;
      nbrick = (long)grid.nbx * grid.nby * grid.nbz;
      grid.ncell = (int)(nbrick * B * B * B);
    }
  }

  // lcbinary.cc:150-157: stencil half-width from the smallest subcell edge.
  // The reference wraps a stencil offset once, so a stencil wider than the
  // grid would index outside it there; such a grid is refused.
  void lcb_side() {
    double d = edge_module(box.L[0]) / double(cm.nx);
    if (edge_module(box.L[1]) / double(cm.ny) < d)
      d = edge_module(box.L[1]) / double(cm.ny);
    if (edge_module(box.L[2]) / double(cm.nz) < d)
      d = edge_module(box.L[2]) / double(cm.nz);
    cm.side = int(ceil((cm.cutoff / d) - 0.5));
    if (cm.side >= cm.nx || cm.side >= cm.ny || cm.side >= cm.nz)
      fail(LPMDX_EINVAL, "lcbinary: the stencil is wider than the subcell grid");
  }

  // Brick edge (cells) for the staged force kernels, or 0 for global-index
  // lists.  Staging copies halo / brick atoms per atom (4.6 at B = 3, 8 at
  // B = 2) to save the neighbour-position gathers of the unstaged kernels, so
  // it pays for long lists and well-filled bricks.  Measured on one MI355X:
  // LJ 4M (B 3, 527 atoms a brick, 77 neighbours: 16.6 per staged atom)
  // 0.61 against 0.89 ms; Sutton-Chen 131k (B 2, 310 atoms, 151 neighbours:
  // 18.9) 5 % faster; lennardjonesMod Ar 131k (B 2, 180 atoms, 12.4), the
  // finnissinclair-ext Mo box (B 3, 201 atoms, 6.7) and harmonic springs
  // (B 3, 81 atoms, 2.6) 30-110 % slower.  Hence: at least 256 atoms a brick
  // and 15 listed neighbours per staged atom.  A brick's halo must also fit
  // the LDS of one workgroup with room for two workgroups per CU, and there
  // must be enough bricks to fill 256 CUs.  LPMDX_BRICK=0/2/3 forces the
  // choice.
  int choose_brick() const {
    static const int forced = [] {
      const char* e = getenv("LPMDX_BRICK");
      return e ? atoi(e) : -1;
    }();
    const long cells = (long)grid.nx * grid.ny * grid.nz;
    if (forced == 0 || cells == 0)
      return 0;
    const double per_cell = (double)n / (double)cells;
    // mean list length: atoms of the cells' density within the list radius
    const double side = std::min(box.L[0] / grid.nx, std::min(box.L[1] / grid.ny, box.L[2] / grid.nz));
    const double reach = std::isfinite(rl) ? std::min(1.0, rl / side) : 1.0;
    const double nbar = per_cell * 4.18879 * reach * reach * reach;
    for (int B = kMaxBrick; B >= 2; --B) {
      if (forced > 0 && B != forced)
        continue;
      if (grid.nx < B + 2 || grid.ny < B + 2 || grid.nz < B + 2)
        continue;
      const long nb = (long)((grid.nx + B - 1) / B) * ((grid.ny + B - 1) / B) * ((grid.nz + B - 1) / B);
      // the mean halo with 10 % for density fluctuations
      const double bytes = (double)stage_bytes((int)(1.1 * per_cell * halo_cells(B)), stage_parts());
      const double per_brick = per_cell * B * B * B;
      const double staged_per_atom = (double)halo_cells(B) / (B * B * B);
      if (forced > 0 || (bytes <= kStageLdsLimit / 2 && nb >= 2 * kNumCu && per_brick >= 256.0 &&
                         nbar >= 15.0 * staged_per_atom))
        return B;
    }
    return 0;
  }

  // The brick choice and the staged LDS size depend on the potentials: a new
  // set of potentials that would choose another brick edge, or stage more
  // (rho^{-1/2}, species) than fits, gets new lists.
  void lists_follow_potentials() {
    if (!list_built || stencil_grid())
      return;
    if (choose_brick() != grid.brick || (staged && stage_bytes(halo_cap, stage_parts()) > (size_t)kStageLdsLimit))
      list_built = false;
  }

  // What the staged kernels of the current potentials copy into LDS besides
  // positions: rho^{-1/2} for EAM, species when a species rule or several
  // potentials share the list.
  int stage_parts() const {
    int parts = 0;
    for (auto& p : pots) {
      if (p->metal())
        parts |= kStageW;
      if (p->spc_sum >= 0 || pots.size() > 1)
        parts |= kStageS;
    }
    return parts;
  }

  long force_blocks() const { return staged ? nbrick : blocks_for(n * lpa); }

  int auto_lanes() const {
    if (lpa_req)
      return lpa_req;
    if (grid.brick > 0) {
      // one workgroup per brick, up to 1024 threads: lanes that keep a brick's
      // atoms within one pass
      const double per_brick = (double)n / (double)nbrick;
      int l = 1;
      while (l < 4 && 1.3 * per_brick * (2 * l) <= 1024.0)
        l *= 2;
      return l;
    }
    // fewer lanes per atom issue fewer instructions per neighbour; small
    // systems need more lanes to fill 256 CUs (131 072 Sutton-Chen atoms:
    // 1 lane 7 % faster than 2)
    if (n >= 100000)
      return 1;
    if (n >= 60000)
      return 2;
    if (n >= 15000)
      return 4;
    return 8;
  }

  void sort_atoms() {
    const long nb = blocks_for(n);
    k_cell_keys<<<nb, kBlock, 0, st>>>(x.p, perm.p, n, grid, box, keys.p, vals.p);
    launch_check();
    const int passes = radix_passes((uint32_t)grid.ncell);
    const int nch = (int)((n + kSortChunk - 1) / kSortChunk);
    hist.alloc((size_t)kRadix * nch);
    htot.alloc(kRadix);
    uint32_t *ka = keys.p, *va = vals.p, *kb = keys2.p, *vb = vals2.p;
    for (int ps = 0; ps < passes; ++ps) {
      const int shift = kRadixBits * ps;
      k_radix_count<<<nch, kBlock, 0, st>>>(ka, n, shift, hist.p, nch);
      k_radix_rows<<<kRadix / kSortWaves, kBlock, 0, st>>>(hist.p, nch, htot.p);
      k_radix_scatter<<<nch, kBlock, 0, st>>>(ka, va, n, shift, hist.p, htot.p, nch, kb, vb);
      launch_check();
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    // sorted keys/order now in (ka, va)
    cb.alloc(grid.ncell + 2);
    k_cell_begin<<<blocks_for(n + 1), kBlock, 0, st>>>(ka, n, grid.ncell, cb.p);
    PermuteArgs P{va, perm.p, x.p, v.p, a.p, z.p, ffm.p, mass.p, x2.p, v2.p, a2.p, z2.p,
                  ffm2.p, mass2.p, perm2.p, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, n};
    // extra per-atom vectors (the leapfrog half-step velocities, and during
    // its set-up the saved positions and accelerations) follow the atoms
    std::vector<std::pair<DBuf<double>*, DBuf<double>*>> ex;
    if (lf_ready)
      ex.push_back({&vh, &vh2});
    if (bm_ready)
      ex.push_back({&ab, &ab2});
    for (auto& e : carry)
      ex.push_back(e);
    if (ex.size() > (size_t)kPermuteExtra)
      fail(LPMDX_EDEVICE, "too many per-atom arrays to reorder (internal error)");
    for (size_t e = 0; e < ex.size(); ++e) {
      ex[e].second->alloc(3 * n);
      P.ex[e] = ex[e].first->p;
      P.ex2[e] = ex[e].second->p;
    }
    k_permute<<<nb, kBlock, 0, st>>>(P);
    launch_check();
    for (auto& e : ex)
      std::swap(e.first->p, e.second->p);
    std::swap(x.p, x2.p);
    std::swap(v.p, v2.p);
    std::swap(a.p, a2.p);
    std::swap(z.p, z2.p);
    std::swap(ffm.p, ffm2.p);
    std::swap(mass.p, mass2.p);
    std::swap(perm.p, perm2.p);
    // keep the sorted keys where the build kernel reads them
    if (ka != keys.p) {
      std::swap(keys.p, keys2.p);
      std::swap(vals.p, vals2.p);
    }
  }

  void build_list() {
    // list storage: cap entries per atom, rounded to the lane count
    lpa = auto_lanes();
    stride = ((n + kWave - 1) / kWave) * kWave; // whole tiles for every lane count
    // staged kernels serve up to 4 lanes per atom
    // (and only with a wrap margin: their neighbour loop has no compare-and-select wrap)
    staged = grid.brick > 0 && lpa <= 4 && wrap_margin() > 0.0;
    const size_t np = (size_t)std::max<long>(blocks_for(n * 16), nbrick) * 12;
    partials.alloc(np);
    partials2.alloc(np);
    // halo_max, brick_max
    hipchk(hipMemsetAsync(&ctrl.p->halo_max, 0, 8, st), "memset");
    if (staged) {
      bhalo.alloc((size_t)nbrick * (halo_cells(grid.brick) + 1));
      k_brick_halo<<<nbrick, kHaloThreads, 0, st>>>(cb.p, grid, bhalo.p, &ctrl.p->halo_max, &ctrl.p->brick_max);
      launch_check();
    }
    if (cap == 0) {
      // initial capacity from the density inside the list radius
      double vol = box.L[0] * box.L[1] * box.L[2];
      double rr = std::isfinite(rl) ? rl : std::cbrt(vol);
      double est = (double)n / vol * 4.18879 * rr * rr * rr;
      if (stencil_grid())
        est = std::max(est, (double)n / vol * pow((2 * grid.side + 1) * box.L[0] / grid.nx, 3));
      cap = (int)std::min<double>(n * 8.0 + 64, est * 1.3 + 32);
    }
    for (int attempt = 0; attempt < 4; ++attempt) {
      cap = ((cap + kChunk - 1) / kChunk) * kChunk; // whole 8-entry chunks
      nbr.alloc((size_t)cap * stride);
      cnt.alloc(n);
      // max_count, pad0, sum_count: 16 contiguous bytes
      hipchk(hipMemsetAsync(&ctrl.p->max_count, 0, 16, st), "memset");
      sumslots.alloc((size_t)kSumSlots * kSlotStride);
      hipchk(hipMemsetAsync(sumslots.p, 0, (size_t)kSumSlots * kSlotStride * sizeof(unsigned long long), st), "memset");
      const double rl2 = rl * rl;
      // FP32 positions are off by at most ulp(Lmax)/2; two of them in a distance
      const double lmax = std::max(box.L[0], std::max(box.L[1], box.L[2]));
      const double pos_err = 2.0 * std::ldexp(lmax, -24) + 1e-30;
      const float rl2f = std::isfinite(rl2) ? (float)((rl + 2.0 * pos_err) * (rl + 2.0 * pos_err) * (1.0 + 1e-6))
                                            : INFINITY;
      // FP32 paths need the minimum image to be unambiguous within the list radius
      const double hmin = std::min(box.half[0], std::min(box.half[1], box.half[2]));
      int path = kPathExact64;
      if (std::isfinite(rl2) && hmin > rl + 4.0 * pos_err + 1e-9 * hmin) {
        if (grid.mode == kExact && grid.nx >= 3 && grid.ny >= 3 && grid.nz >= 3)
          path = kPathShift32;
        else
          path = kPathMin32;
      }
      xf.alloc(n + kGroup); // padding read by the candidate groups
      k_to_f32<<<blocks_for(n), kBlock, 0, st>>>(x.p, n, box, xf.p);
      BuildArgs B{};
      B.x = x.p;
      B.xf = xf.p;
      B.skeys = keys.p;
      B.cb = cb.p;
      B.bhalo = staged ? bhalo.p : nullptr;
      B.nbr = nbr.p;
      B.cnt = cnt.p;
      B.max_count = &ctrl.p->max_count;
      B.sum_count = sumslots.p;
      B.sum_spread = 1;
      B.g = grid;
      B.b = box;
      for (int q = 0; q < 3; ++q) {
        B.Lf[q] = (float)box.L[q];
        B.halff[q] = (float)box.half[q];
      }
      B.rl2 = rl2;
      B.rl2f = rl2f;
      B.n = n;
      B.stride = stride;
      B.cap = cap;
      B.lpa = lpa;
      B.path = path;
      hipEvent_t e0 = prof_begin();
      k_build_list<<<blocks_for(n), kBlock, 0, st>>>(B);
      prof_end(kProfBuild, e0);
      k_sum_slots<<<1, kWave, 0, st>>>(sumslots.p, &ctrl.p->sum_count);
      launch_check();
      // max_count, pad0, sum_count, halo_max, brick_max: 24 contiguous bytes
      hipchk(hipMemcpyAsync(&hctrl->max_count, &ctrl.p->max_count, 24, hipMemcpyDeviceToHost, st),
             "memcpy");
      sync();
      if (staged) {
        halo_cap = (int)(((hctrl->halo_max + 63) / 64) * 64);
        const int lanes_atoms = (int)hctrl->brick_max * lpa;
        brick_threads = std::max(kWave, std::min(1024, ((lanes_atoms + kWave - 1) / kWave) * kWave));
        if (stage_bytes(halo_cap, stage_parts()) > (size_t)kStageLdsLimit) {
          // a halo too large for LDS (a dense spot): global indices this time
          // (the grid keeps its brick order; the unstaged kernels take any order)
          staged = false;
          halo_cap = 0;
          continue;
        }
      }
      const long mc = hctrl->max_count;
      max_list = mc;
      sum_list = (long)hctrl->sum_count;
      if (mc <= cap)
        return;
      cap = (int)(mc * 1.2 + 16);
    }
    fail(LPMDX_ECAPACITY, "neighbour list capacity could not be established");
  }

  void rebuild() {
    if (n == 0) {
      list_built = true;
      return;
    }
    rl = list_radius();
    setup_grid();
    hipEvent_t e0 = prof_begin();
    sort_atoms();
    // lcbinary.cc:244-276: a subcell that holds two atoms grows the grid by two
    // subcells per axis, and the atoms are binned again
    while (cm.kind == kCmLcBinary && max_cell_occupancy() > 1) {
      cm.nx += 2;
      cm.ny += 2;
      cm.nz += 2;
      lcb_side();
      setup_grid();
      sort_atoms();
    }
    prof_end(kProfBuild, e0);
    build_list();
    // verletlist.cc:96-97: a list may hold at most maxneighbors-1 entries
    if (cm.kind == kCmVerlet && max_list >= cm.maxneighbors)
      fail(LPMDX_ECAPACITY, "verletlist: maxneighbors is too small!");
    hipchk(hipMemcpyAsync(xb.p, x.p, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    list_built = true;
    n_builds++;
  }

  unsigned int max_cell_occupancy() {
    hipchk(hipMemsetAsync(&ctrl.p->cell_max, 0, sizeof(unsigned), st), "memset");
    k_cell_occupancy<<<blocks_for(grid.ncell), kBlock, 0, st>>>(cb.p, grid.ncell, &ctrl.p->cell_max);
    launch_check();
    hipchk(hipMemcpyAsync(&hctrl->cell_max, &ctrl.p->cell_max, sizeof(unsigned), hipMemcpyDeviceToHost, st),
           "memcpy");
    sync();
    return hctrl->cell_max;
  }

  // exact displacement since the build, for calls through the plugin interface
  double displacement_since_build() {
    // |x - xb| max over atoms via a throw-away kernel
    hipchk(hipMemsetAsync(&ctrl.p->maxd2[0], 0, sizeof(unsigned), st), "memset");
    ForceArgs A{};
    A.x = x.p;
    A.xn = xb.p;
    A.n = n;
    A.b = box;
    A.maxd2 = &ctrl.p->maxd2[0];
    k_disp<<<blocks_for(n), kBlock, 0, st>>>(A);
    launch_check();
    hipchk(hipMemcpyAsync(&hctrl->maxd2[0], &ctrl.p->maxd2[0], sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
    float f;
    memcpy(&f, &hctrl->maxd2[0], sizeof f);
    return std::sqrt((double)f);
  }


  // CellManager::UpdateCell
  void update_cell() {
    if (cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    undrift();
    if (cm.kind == kCmMinImage && fabs(cm.cutoff) < 1E-10) {
      // minimumimage.cc:63-69: no cutoff given: the largest half cell edge
      for (int q = 0; q < 3; ++q)
        if (0.5 * edge_module(box.L[q]) > cm.cutoff)
          cm.cutoff = 0.5 * edge_module(box.L[q]);
    }
    if (rebuilds_every_update()) {
      rebuild();
    } else {
      // rebuilt whenever the list could miss a pair inside the filter radius;
      // while the reference's own list is valid both deliver the same set
      bool must = !list_built || list_radius() != rl;
      if (!must) {
        double b = budget();
        must = displacement_since_build() > b;
      }
      if (must)
        rebuild();
    }
    cm.ucalls++;
    cell_version = pos_version;
  }

  // user positions back into x (undo the look-ahead drift)
  void undrift() {
    if (drifted) {
      std::swap(x.p, xn.p);
      drifted = false;
    }
  }

  // ------------------------------------------------------------------ forces
  // SPC: kernels that read species (a species rule or several potentials of
  // one kind); the others run a branch-free neighbour loop
  static bool reads_species(const ForceArgs& A) { return A.nterm > 1 || A.P.spc_sum >= 0; }

  template <int KIND, bool EVAL> void launch_pair_k(ForceArgs& A, long nb) {
    if (reads_species(A))
      launch_pair_s<KIND, EVAL, true>(A, nb);
    else
      launch_pair_s<KIND, EVAL, false>(A, nb);
  }
  template <int KIND, bool EVAL, bool SPC> void launch_pair_s(ForceArgs& A, long nb) {
    hipEvent_t e0 = prof_begin();
    if (staged) {
      const int parts = SPC ? kStageS : 0;
      const BrickLaunch L{nbrick, brick_threads, stage_bytes(halo_cap, parts), st, bhalo.p, grid.brick, halo_cap};
      switch (lpa) {
      case 1: BrickKernels<KIND, 1, EVAL>::template pair<SPC>(L, A, parts); break;
      case 2: BrickKernels<KIND, 2, EVAL>::template pair<SPC>(L, A, parts); break;
      default: BrickKernels<KIND, 4, EVAL>::template pair<SPC>(L, A, parts); break;
      }
      prof_end(kProfForce, e0);
      return;
    }
    switch (lpa) {
    case 1: k_pair_force<KIND, 1, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    case 2: k_pair_force<KIND, 2, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    case 4: k_pair_force<KIND, 4, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    case 8: k_pair_force<KIND, 8, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    default: k_pair_force<KIND, 16, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    }
    prof_end(kProfForce, e0);
  }
  template <int KIND, bool EVAL> void launch_eam_k(ForceArgs& A, long nb, double rc2_rho) {
    if (A.P.spc_sum >= 0)
      launch_eam_s<KIND, EVAL, true>(A, nb, rc2_rho);
    else
      launch_eam_s<KIND, EVAL, false>(A, nb, rc2_rho);
  }
  // Unstaged Sutton-Chen steps with the common integer exponent pairs run
  // kernels with the exponents fixed at compile time (bit-identical results,
  // about 20 fewer instructions per neighbour).
  template <int KIND, bool EVAL, bool SPC> bool launch_sc_fixed(ForceArgs& A, ForceArgs& R, long nb) {
    if constexpr (KIND != kSuttonChen || EVAL) {
      return false;
    } else {
      const int ex = sc_exp(A.P.ni, A.P.mi);
      if (lpa > 4 || A.P.ni <= 0 || A.P.mi <= 0)
        return false;
      if (ex == sc_exp(9, 6))
        return launch_eam_fixed<KIND, SPC, sc_exp(9, 6)>(A, R, nb);
      if (ex == sc_exp(12, 6))
        return launch_eam_fixed<KIND, SPC, sc_exp(12, 6)>(A, R, nb);
      return false;
    }
  }
  template <int KIND, bool SPC, int EXP> bool launch_eam_fixed(ForceArgs& A, ForceArgs& R, long nb) {
    hipEvent_t e0 = prof_begin();
    if (staged) {
      const int parts = kStageW | (SPC ? kStageS : 0);
      const BrickLaunch Lr{nbrick, brick_threads, stage_bytes(halo_cap, 0), st, bhalo.p, grid.brick, halo_cap};
      const BrickLaunch Lf{nbrick, brick_threads, stage_bytes(halo_cap, parts), st, bhalo.p, grid.brick, halo_cap};
      switch (lpa) {
      case 1: BrickKernels<KIND, 1, false>::template rho<EXP>(Lr, R, w_rho.p, rho.p); break;
      case 2: BrickKernels<KIND, 2, false>::template rho<EXP>(Lr, R, w_rho.p, rho.p); break;
      default: BrickKernels<KIND, 4, false>::template rho<EXP>(Lr, R, w_rho.p, rho.p); break;
      }
      prof_end(kProfRho, e0);
      A.w_rho = w_rho.p;
      hipEvent_t e1 = prof_begin();
      switch (lpa) {
      case 1: BrickKernels<KIND, 1, false>::template eam<SPC, EXP>(Lf, A, parts); break;
      case 2: BrickKernels<KIND, 2, false>::template eam<SPC, EXP>(Lf, A, parts); break;
      default: BrickKernels<KIND, 4, false>::template eam<SPC, EXP>(Lf, A, parts); break;
      }
      prof_end(kProfForce, e1);
      return true;
    }
    switch (lpa) {
    case 1: k_eam_rho<KIND, 1, false, EXP><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    case 2: k_eam_rho<KIND, 2, false, EXP><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    default: k_eam_rho<KIND, 4, false, EXP><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    }
    prof_end(kProfRho, e0);
    A.w_rho = w_rho.p;
    hipEvent_t e1 = prof_begin();
    switch (lpa) {
    case 1: k_eam_force<KIND, 1, false, SPC, EXP><<<nb, kBlock, 0, st>>>(A); break;
    case 2: k_eam_force<KIND, 2, false, SPC, EXP><<<nb, kBlock, 0, st>>>(A); break;
    default: k_eam_force<KIND, 4, false, SPC, EXP><<<nb, kBlock, 0, st>>>(A); break;
    }
    prof_end(kProfForce, e1);
    return true;
  }
  template <int KIND, bool EVAL, bool SPC> void launch_eam_s(ForceArgs& A, long nb, double rc2_rho) {
    ForceArgs R = A;
    R.P.rc2 = rc2_rho;
    R.partials = partials2.p;
    if (launch_sc_fixed<KIND, EVAL, SPC>(A, R, nb))
      return;
    hipEvent_t e0 = prof_begin();
    if (staged) {
      const int parts = kStageW | (SPC ? kStageS : 0);
      const BrickLaunch Lr{nbrick, brick_threads, stage_bytes(halo_cap, 0), st, bhalo.p, grid.brick, halo_cap};
      const BrickLaunch Lf{nbrick, brick_threads, stage_bytes(halo_cap, parts), st, bhalo.p, grid.brick, halo_cap};
      switch (lpa) {
      case 1: BrickKernels<KIND, 1, EVAL>::rho(Lr, R, w_rho.p, rho.p); break;
      case 2: BrickKernels<KIND, 2, EVAL>::rho(Lr, R, w_rho.p, rho.p); break;
      default: BrickKernels<KIND, 4, EVAL>::rho(Lr, R, w_rho.p, rho.p); break;
      }
      prof_end(kProfRho, e0);
      A.w_rho = w_rho.p;
      hipEvent_t e1 = prof_begin();
      switch (lpa) {
      case 1: BrickKernels<KIND, 1, EVAL>::template eam<SPC>(Lf, A, parts); break;
      case 2: BrickKernels<KIND, 2, EVAL>::template eam<SPC>(Lf, A, parts); break;
      default: BrickKernels<KIND, 4, EVAL>::template eam<SPC>(Lf, A, parts); break;
      }
      prof_end(kProfForce, e1);
      return;
    }
    switch (lpa) {
    case 1: k_eam_rho<KIND, 1, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    case 2: k_eam_rho<KIND, 2, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    case 4: k_eam_rho<KIND, 4, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    case 8: k_eam_rho<KIND, 8, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    default: k_eam_rho<KIND, 16, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
    }
    prof_end(kProfRho, e0);
    A.w_rho = w_rho.p;
    hipEvent_t e1 = prof_begin();
    switch (lpa) {
    case 1: k_eam_force<KIND, 1, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    case 2: k_eam_force<KIND, 2, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    case 4: k_eam_force<KIND, 4, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    case 8: k_eam_force<KIND, 8, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    default: k_eam_force<KIND, 16, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
    }
    prof_end(kProfForce, e1);
  }
  template <bool EVAL> void launch_pot(int kind, ForceArgs& A, long nb, double rc2_rho) {
    switch (kind) {
    case kLJ: launch_pair_k<kLJ, EVAL>(A, nb); break;
    case kFastLJ: launch_pair_k<kFastLJ, EVAL>(A, nb); break;
    case kMorse: launch_pair_k<kMorse, EVAL>(A, nb); break;
    case kBuckingham: launch_pair_k<kBuckingham, EVAL>(A, nb); break;
    case kSuttonChen: launch_eam_k<kSuttonChen, EVAL>(A, nb, rc2_rho); break;
    case kFinnisSinclair: launch_eam_k<kFinnisSinclair, EVAL>(A, nb, rc2_rho); break;
    case kGupta: launch_eam_k<kGupta, EVAL>(A, nb, rc2_rho); break;
    case kHarmonic: launch_pair_k<kHarmonic, EVAL>(A, nb); break;
    case kLJMod: launch_pair_k<kLJMod, EVAL>(A, nb); break;
    case kFinnisSinclairExt: launch_eam_k<kFinnisSinclairExt, EVAL>(A, nb, rc2_rho); break;
    }
    launch_check();
  }

  // Filters as the reference's managers serve each potential (see half_list_rcut).
  void filters(const PotParams& P, double& rc2_force, double& rc2_rho) const {
    const double own = P.cutoff * P.cutoff;
    if (cm.kind == kCmLinkedCell) {
      const double h = half_list_rcut() * half_list_rcut();
      const double f = full_list_rcut() * full_list_rcut();
      // pair potentials also test their own cutoff (pairpotential.cc:56);
      // metal potentials take the list as it comes (metalpotential.cc:140-145)
      rc2_force = P.metal() ? h : std::min(h, own);
      rc2_rho = f;
    } else if (cm.kind == kCmMinImage || cm.kind == kCmLcBinary) {
      // every request lists with the potential's own cutoff
      rc2_force = own;
      rc2_rho = own;
    } else {
      // the stored list may hold a few pairs just beyond its radius (the build
      // admits with a small FP32 margin); a potential without a finite cutoff
      // takes the list radius itself
      rc2_force = std::min(own, rl * rl);
      rc2_rho = rc2_force;
    }
  }

  // Evaluate all potentials at x; mode kForcesOnly / kInitDrift / kStep.
  // With eval, energies land in partials and are reduced into `red`.
  // gate: the previous step's displacement slot of a speculative step (the
  // kernels return at once when it failed the skin test), or null
  void forces(int mode, bool eval, double dt, unsigned int* maxd2, const unsigned int* gate = nullptr) {
    const double gate_b = gate ? budget() : 0.0;
    if (pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    const long nb = force_blocks();
    const bool fused = pots.size() == 1;
    for (size_t q = 0; q < pots.size(); ++q) {
      const PotParams& P = *pots[q];
      double rcf, rcr;
      filters(P, rcf, rcr);
      ForceArgs A{};
      A.x = x.p;
      A.xn = xn.p;
      A.v = v.p;
      A.a = fused ? a.p : asum.p;
      A.z = z.p;
      A.ffm = ffm.p;
      A.nbr = nbr.p;
      A.cnt = cnt.p;
      A.partials = partials.p;
      A.maxd2 = maxd2;
      A.gate = gate;
      A.gate_b = gate_b;
      A.xb = xb.p;
      A.vh = vh.p;
      A.ab = ab.p;
      A.margin = wrap_margin();
      if (staged && !(A.margin > 0.0))
        fail(LPMDX_EDEVICE, "staged force kernels without a wrap margin (internal error)");
      A.P = device_view(P, rcf, rcr);
      A.b = box;
      for (int k = 0; k < 3; ++k)
        A.ll[k] = ll[k];
      A.dt = dt;
      A.n = n;
      A.stride = stride;
      A.cap = cap;
      A.mode = fused ? mode : kForcesOnly;
      A.accumulate = q > 0;
      if (P.metal()) {
        w_rho.alloc(n);
        rho.alloc(n);
        xw.alloc(n);
        A.xw = xw.p;
      }
      if (eval)
        launch_pot<true>(P.kind, A, nb, rcr);
      else
        launch_pot<false>(P.kind, A, nb, rcr);
      if (eval) {
        // per-potential totals accumulate into red[0..11]
        k_reduce_partials<<<1, kBlock, 0, st>>>(partials.p, nb, 12, red.p + 12 * (1 + 2 * q));
        if (P.metal())
          k_reduce_partials<<<1, kBlock, 0, st>>>(partials2.p, nb, 12, red.p + 12 * (2 + 2 * q));
        else
          k_fill_double<<<1, 32, 0, st>>>(red.p + 12 * (2 + 2 * q), 12, 0.0);
        launch_check();
      }
    }
    if (!fused && mode != kForcesOnly) {
      ForceArgs A{};
      A.x = x.p;
      A.xn = xn.p;
      A.v = v.p;
      A.a = a.p;
      A.maxd2 = maxd2;
      A.gate = gate;
      A.gate_b = gate_b;
      A.xb = xb.p;
      A.vh = vh.p;
      A.ab = ab.p;
      A.b = box;
      for (int k = 0; k < 3; ++k)
        A.ll[k] = ll[k];
      A.dt = dt;
      A.n = n;
      A.mode = mode;
      k_integrate<<<blocks_for(n), kBlock, 0, st>>>(A, asum.p);
      launch_check();
    } else if (!fused) {
      hipchk(hipMemcpyAsync(a.p, asum.p, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    }
    n_forces++;
  }


  // energies of the last evaluated forces: E, W, S[9]
  void read_energies(double* e, double* w, double* S) {
    const size_t np = pots.size();
    hipchk(hipMemcpyAsync(hred, red.p + 12, 24 * np * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
    double E = 0.0, W = 0.0, T[9] = {0};
    for (size_t q = 0; q < np; ++q) {
      const double* pr = hred + 24 * q;
      const double* em = pr + 12;
      // CombinedPotential::energy sums each potential's energycache
      double ep = pr[0] + em[1];
      E += ep;
      W += pr[2];
      for (int k = 0; k < 9; ++k)
        T[k] += pr[3 + k];
    }
    if (e)
      *e = E;
    if (w)
      *w = W;
    if (S)
      for (int k = 0; k < 9; ++k)
        S[k] = T[k];
    last_e[0] = E;
    last_e[1] = W;
    for (int k = 0; k < 9; ++k)
      last_e[2 + k] = T[k];
    last_e_valid = true;
  }

  // One potential's UpdateForces (pairpotential.cc:35-84 / metalpotential.cc:62-180
  // for potential k alone): its accelerations into asum (internal order),
  // its energy, virial and stress into red.
  void forces_term(int k) {
    const PotParams& P = *pots[k];
    double rcf, rcr;
    filters(P, rcf, rcr);
    const long nb = force_blocks();
    asum.alloc(3 * n);
    ForceArgs A{};
    A.x = x.p;
    A.xn = xn.p;
    A.v = v.p;
    A.a = asum.p;
    A.z = z.p;
    A.ffm = ffm.p;
    A.nbr = nbr.p;
    A.cnt = cnt.p;
    A.partials = partials.p;
    A.maxd2 = ctrl.p->maxd2;
    A.xb = xb.p;
    A.margin = wrap_margin();
    if (staged && !(A.margin > 0.0))
      fail(LPMDX_EDEVICE, "staged force kernels without a wrap margin (internal error)");
    A.P = device_view(P, rcf, rcr);
    A.b = box;
    for (int q = 0; q < 3; ++q)
      A.ll[q] = ll[q];
    A.n = n;
    A.stride = stride;
    A.cap = cap;
    A.mode = kForcesOnly;
    A.accumulate = 0;
    if (P.metal()) {
      w_rho.alloc(n);
      rho.alloc(n);
      xw.alloc(n);
      A.xw = xw.p;
    }
    launch_pot<true>(P.kind, A, nb, rcr);
    k_reduce_partials<<<1, kBlock, 0, st>>>(partials.p, nb, 12, red.p + 12 * (1 + 2 * k));
    if (P.metal())
      k_reduce_partials<<<1, kBlock, 0, st>>>(partials2.p, nb, 12, red.p + 12 * (2 + 2 * k));
    else
      k_fill_double<<<1, 32, 0, st>>>(red.p + 12 * (2 + 2 * k), 12, 0.0);
    launch_check();
    hipchk(hipMemcpyAsync(hred, red.p + 12 * (1 + 2 * k), 24 * sizeof(double), hipMemcpyDeviceToHost, st),
           "memcpy");
    sync();
    n_forces++;
  }

  double kinetic(double p[3]) {
    const long nb = blocks_for(n);
    partials.alloc(std::max<size_t>((size_t)nb * 4, partials.n));
    k_kinetic<<<nb, kBlock, 0, st>>>(v.p, mass.p, n, partials.p);
    k_reduce_partials<<<1, kBlock, 0, st>>>(partials.p, nb, 4, red.p);
    launch_check();
    hipchk(hipMemcpyAsync(hred, red.p, 4 * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
    if (p)
      for (int q = 0; q < 3; ++q)
        p[q] = hred[1 + q];
    return hred[0] * kKin2eV;
  }

  // SetTemperature on the device, ordered on the stream (no host round trip);
  // partials must hold 4 * blocks_for(n) doubles
  // q: -1 for SetTemperature, dt / tau for a Berendsen step
  void scale_to_temperature(double temp, const unsigned int* gate = nullptr, double q = -1.0) {
    const long nb = blocks_for(n);
    k_kinetic<<<nb, kBlock, 0, st>>>(v.p, mass.p, n, partials.p);
    k_reduce_partials<<<1, kBlock, 0, st>>>(partials.p, nb, 4, red.p);
    k_set_temperature<<<nb, kBlock, 0, st>>>(v.p, red.p, n, temp, q, gate, gate ? budget() : 0.0);
    launch_check();
  }

  // After a step's kernel in the two-step loops: when the schedule is active
  // in step k, scale v and make the look-ahead positions (in x; the step's
  // positions are in xn) again from the scaled velocities.
  // Velocity modifiers of step k (tempscaling, then berendsen); true when
  // any scaled v.
  bool velocity_modifiers(long k, double dt, const unsigned int* gate) {
    bool any = false;
    if (tsched.active(k)) {
      scale_to_temperature(tsched.value(k), gate);
      any = true;
    }
    double t = 0.0;
    if (bsched.apply(k, dt, t)) {
      scale_to_temperature(t, gate, dt / bsched.tau);
      any = true;
    }
    return any;
  }

  // After a chunk that ran `ran` of its steps: the berendsen state as the
  // steps that ran left it (from the state saved at the chunk start).
  void replay_berendsen(const Berendsen& saved, long first_step, long ran, double dt) {
    bsched = saved;
    double t;
    for (long k = first_step; k < first_step + ran; ++k)
      bsched.apply(k, dt, t);
  }

  void modifiers_after_step(long k, bool beeman, double dt, unsigned int* maxd2, const unsigned int* gate) {
    if (!velocity_modifiers(k, dt, gate))
      return;
    ForceArgs A{};
    A.gate = gate;
    A.gate_b = gate ? budget() : 0.0;
    A.x = xn.p;
    A.xn = x.p;
    A.v = v.p;
    A.a = a.p;
    A.ab = ab.p;
    A.maxd2 = maxd2; // the larger of the two look-aheads' displacements: conservative
    A.xb = xb.p;
    A.b = box;
    A.n = n;
    A.dt = dt;
    A.mode = beeman ? kBeemanInit : kStep;
    for (int q = 0; q < 3; ++q)
      A.ll[q] = ll[q];
    k_drift<<<blocks_for(n), kBlock, 0, st>>>(A);
    launch_check();
  }

  // ------------------------------------------------------------------ VV
  // beeman: the same two-step scheme with Beeman's updates (beeman.cc:63-84)
  // in place of velocity Verlet's; ab carries its extra history.
  void vv(double dt, long steps, long obs_every, double* ep_t, double* ek_t, bool beeman = false) {
    const char* who = beeman ? "beeman" : "velocityverlet";
    if (pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!(dt == dt) || !std::isfinite(dt))
      fail(LPMDX_ENONFINITE, "dt is not finite");
    if (beeman && !bm_ready)
      fail(LPMDX_ESTATE, "beeman: lpmdx_beeman_init was not called");
    if (beeman && !(dt == bm_dt))
      fail(LPMDX_EINVAL, "beeman: dt differs from the one lpmdx_beeman_init used");
    if (n == 0 || steps <= 0)
      return;
    // a look-ahead made by the other scheme is taken back (a stays valid)
    if (drifted && drift_beeman != beeman)
      undrift();
    const int init_mode = beeman ? kBeemanInit : kInitDrift;
    const int step_mode = beeman ? kBeemanStep : kStep;
    if (tsched.on || bsched.on)
      partials.alloc(std::max<size_t>((size_t)blocks_for(n) * 4, partials.n));
    unsigned int* md = ctrl.p->maxd2;
    auto read_disp = [&](long s) {
      float f;
      memcpy(&f, &hctrl->maxd2[s], sizeof f);
      return std::sqrt((double)f);
    };
    // reach the look-ahead state: x = next positions, (v, a) of the last step
    if (!drifted) {
      if (!forces_ok)
        update_cell();
      hipchk(hipMemsetAsync(md, 0, sizeof(unsigned), st), "memset");
      if (!forces_ok) {
        forces(init_mode, false, dt, md);
      } else {
        ForceArgs A{};
        A.x = x.p;
        A.xn = xn.p;
        A.v = v.p;
        A.a = a.p;
        A.ab = ab.p;
        A.maxd2 = md;
        A.xb = xb.p;
        A.b = box;
        A.n = n;
        A.dt = dt;
        A.mode = beeman ? kBeemanInit : kStep;
        for (int k = 0; k < 3; ++k)
          A.ll[k] = ll[k];
        k_drift<<<blocks_for(n), kBlock, 0, st>>>(A);
        launch_check();
      }
      std::swap(x.p, xn.p);
      drifted = true;
      drift_beeman = beeman;
      forces_ok = true;
      hipchk(hipMemcpyAsync(&hctrl->maxd2[0], md, sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy");
      sync();
      disp_cur = read_disp(0);
    }
    // Steps run in chunks without a host round trip; after a chunk the exact
    // displacement of every step's positions from the build positions says
    // whether the chunk's lists stayed valid (every atom within half the skin
    // of its build position).  A chunk that went past it is rolled back and
    // re-run up to its last valid step; the list is rebuilt there.
    // Steps run in chunks without a host round trip.  Every step after a
    // chunk's first is speculative: its kernels run only while the positions
    // its forces read passed the skin test (every atom within half the skin
    // of its build position, read from the previous step's displacement slot
    // on the device), and return at once otherwise.  After the chunk the host
    // reads the slots, keeps the steps that ran, and rebuilds the list there;
    // nothing has to be rolled back.
    long done = 0;
    while (done < steps) {
      prof_step = -1; // rebuilds are no step's launches
      long m = steps - done;
      if (obs_every > 0)
        m = std::min(m, obs_every - (done % obs_every));
      m = std::min<long>(m, kMaxChunk);
      const bool every = rebuilds_every_update();
      bool need = false;
      if (every) {
        m = 1;
        need = true; // UpdateCell rebuilds every call
      } else {
        need = !list_built || list_radius() != rl || disp_cur > budget();
        m = std::min(m, chunk);
      }
      // positions x are where the next force evaluation happens
      drifted = false;
      if (need) {
        rebuild();
        disp_cur = 0.0;
      }
      // cleared per-step maxima (m <= kMaxChunk)
      k_chunk_begin<<<blocks_for(m), kBlock, 0, st>>>(md, m);
      launch_check();
      const Berendsen bsaved = bsched;
      for (long s = 0; s < m; ++s) {
        const bool obs = obs_every > 0 && ((done + s + 1) % obs_every) == 0;
        const unsigned int* gate = (every || s == 0) ? nullptr : md + s - 1;
        prof_step = n_steps + s;
        forces(step_mode, obs, dt, md + s, gate);
        std::swap(x.p, xn.p);
        modifiers_after_step(n_steps + s + 1, beeman, dt, md + s, gate);
      }
      prof_step = -1;
      hipchk(hipMemcpyAsync(hctrl->maxd2, md, m * sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy");
      // no profile collection yet: the timings of steps that did not run go first
      hipchk(hipStreamSynchronize(st), "hipStreamSynchronize");
      // the force of step s saw the positions drifted by step s-1; the first
      // step that failed the test ran no kernel, nor did any after it
      const double b = budget();
      double d = disp_cur, growth = 0.0;
      long ran = m;
      for (long s = 0; s < m; ++s) {
        if (!every && d > b) {
          ran = s;
          break;
        }
        const double dn = read_disp(s);
        growth = std::max(growth, dn - d);
        d = dn;
      }
      // each launch swapped the position buffers; the skipped ones wrote none
      if ((m - ran) % 2 == 1)
        std::swap(x.p, xn.p);
      prof_discard_from(n_steps + ran);
      if (ran < m)
        replay_berendsen(bsaved, n_steps + 1, ran, dt);
      if (!pending.empty())
        prof_collect();
      if (ran < m)
        n_cut_chunks++;
      disp_cur = d;
      pos_version++;
      // next chunk: double while nothing moves, else the steps the growth
      // rate leaves before the budget (from a fresh list after a cut), plus
      // two so that the chunk usually ends past the limit in one round trip
      long next = std::min<long>(2 * chunk, kMaxChunk);
      if (growth > 0.0) {
        const double room = disp_cur > b ? b : b - disp_cur;
        next = std::min<long>(next, (long)(room / growth) + 2);
      }
      chunk = std::max<long>(1, next);
      drifted = true;
      drift_beeman = beeman;
      cm.ucalls += 2 * ran;
      done += ran;
      n_steps += ran;
      if (obs_every > 0 && (done % obs_every) == 0) {
        double e;
        read_energies(&e, nullptr, nullptr);
        // the step's state: v in v, positions in xn
        double ek = kinetic(nullptr);
        long idx = done / obs_every - 1;
        if (ep_t)
          ep_t[idx] = e;
        if (ek_t)
          ek_t[idx] = ek;
      }
    }
  }

  // ---------------------------------------------------------------- leapfrog
  // Leapfrog::Initialize (plugins/leapfrog.cc:58-67) through
  // Integrator::UseOldConfig / GoBack (liblpmd/simulation/integrator.cc:16-42):
  // the old configuration is FittedInside(x - v dt) with velocities v - a dt;
  // forces at the old positions give a_old, and the half-step velocities are
  // vh = (v - a dt) + (a_old * 0.5) dt.  x and a are left as they were.
  void lf_init(double dt) {
    if (pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!std::isfinite(dt))
      fail(LPMDX_ENONFINITE, "dt is not finite");
    undrift();
    lf_dt = dt;
    lf_ready = false;
    if (n == 0) {
      lf_ready = true;
      return;
    }
    old_config_forces(dt, true);
    k_lf_halfkick<<<blocks_for(n), kBlock, 0, st>>>(vh.p, a.p, n, dt);
    launch_check();
    restore_after_old_config();
    lf_ready = true;
    sync();
  }

  // Integrator::UseOldConfig / GoBack (integrator.cc:16-42) and one force
  // call on the old configuration FittedInside(x - v dt): leaves a = a_old at
  // the old positions in x, keep_x / keep_a holding x and a (and, with_vh,
  // vh = v - a dt), all in the atom order after any reorder of that call.
  void old_config_forces(double dt, bool with_vh) {
    carry.clear();
    keep_x.alloc(3 * n);
    keep_a.alloc(3 * n);
    if (with_vh)
      vh.alloc(3 * n);
    k_lf_goback<<<blocks_for(n), kBlock, 0, st>>>(x.p, v.p, a.p, keep_x.p, keep_a.p, with_vh ? vh.p : nullptr, n,
                                                  dt, ll[0], ll[1], ll[2]);
    launch_check();
    // lists and forces at the old positions; the saved arrays follow the
    // atoms through any reorder
    if (with_vh)
      carry.push_back({&vh, &vh2});
    carry.push_back({&keep_x, &keep_x2});
    carry.push_back({&keep_a, &keep_a2});
    list_built = false;
    forces_ok = false;
    try {
      update_cell();
      forces(kForcesOnly, false, 0.0, nullptr); // a <- a_old
    } catch (...) {
      carry.clear();
      throw;
    }
    carry.clear();
  }

  // x and a back from keep_x / keep_a after old_config_forces
  void restore_after_old_config() {
    hipchk(hipMemcpyAsync(x.p, keep_x.p, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    hipchk(hipMemcpyAsync(a.p, keep_a.p, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    list_built = false; // the lists were built for the old positions
    forces_ok = false;
  }

  // ---------------------------------------------------------------- beeman
  // Beeman::Initialize (plugins/beeman.cc:55-61): the forces at the old
  // configuration of Integrator::GoBack give the first step's a_old (ab).
  // x, v and a are left as they were.  The steps are vv(..., beeman = true).
  void bm_init(double dt) {
    if (pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!std::isfinite(dt))
      fail(LPMDX_ENONFINITE, "dt is not finite");
    undrift();
    bm_dt = dt;
    bm_ready = false;
    if (n == 0) {
      bm_ready = true;
      return;
    }
    old_config_forces(dt, false);
    ab.alloc(3 * n);
    hipchk(hipMemcpyAsync(ab.p, a.p, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    restore_after_old_config();
    bm_ready = true;
    sync();
  }

  // `steps` calls of OneStepIntegrator::Advance with Leapfrog::Advance
  // (onestepintegrator.cc:13-46, leapfrog.cc:69-77): forces at x, then per
  // atom vhalf = vh + a dt, x = FittedInside(x + vhalf dt),
  // v = 0.5 (vhalf + vh), vh = vhalf (fused into the force kernels).
  // Speculative chunks as in vv(); the skin test reads the positions each
  // step's force sees.
  void lf(double dt, long steps, long obs_every, double* ep_t, double* ek_t) {
    if (pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!lf_ready)
      fail(LPMDX_ESTATE, "leapfrog: lpmdx_leapfrog_init was not called");
    if (!(dt == lf_dt))
      fail(LPMDX_EINVAL, "leapfrog: dt differs from the one lpmdx_leapfrog_init used");
    if (n == 0 || steps <= 0)
      return;
    undrift();
    if (tsched.on || bsched.on)
      partials.alloc(std::max<size_t>((size_t)blocks_for(n) * 4, partials.n));
    unsigned int* md = ctrl.p->maxd2;
    auto read_disp = [&](long s) {
      float f;
      memcpy(&f, &hctrl->maxd2[s], sizeof f);
      return std::sqrt((double)f);
    };
    if (list_built && !rebuilds_every_update()) {
      // displacement of the current positions from the build positions
      ForceArgs A{};
      A.x = x.p;
      A.xn = xb.p;
      A.maxd2 = md;
      A.b = box;
      A.n = n;
      hipchk(hipMemsetAsync(md, 0, sizeof(unsigned), st), "memset");
      k_disp<<<blocks_for(n), kBlock, 0, st>>>(A);
      launch_check();
      hipchk(hipMemcpyAsync(&hctrl->maxd2[0], md, sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy");
      sync();
      disp_cur = read_disp(0);
    }
    // speculative chunks as in vv()
    long done = 0;
    while (done < steps) {
      prof_step = -1; // rebuilds are no step's launches
      long m = steps - done;
      if (obs_every > 0)
        m = std::min(m, obs_every - (done % obs_every));
      m = std::min<long>(m, kMaxChunk);
      const bool every = rebuilds_every_update();
      bool need;
      if (every) {
        m = 1;
        need = true;
      } else {
        need = !list_built || list_radius() != rl || disp_cur > budget();
        m = std::min(m, chunk);
      }
      if (need) {
        rebuild();
        disp_cur = 0.0;
      }
      k_chunk_begin<<<blocks_for(m), kBlock, 0, st>>>(md, m);
      launch_check();
      const Berendsen bsaved = bsched;
      for (long s = 0; s < m; ++s) {
        const bool obs = obs_every > 0 && ((done + s + 1) % obs_every) == 0;
        const unsigned int* gate = (every || s == 0) ? nullptr : md + s - 1;
        prof_step = n_steps + s;
        forces(kLeapfrog, obs, dt, md + s, gate);
        std::swap(x.p, xn.p);
        // the next step reads the half-step velocities only, as in the
        // reference, so the scaled v changes no trajectory
        velocity_modifiers(n_steps + s + 1, dt, gate);
      }
      prof_step = -1;
      hipchk(hipMemcpyAsync(hctrl->maxd2, md, m * sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy");
      // no profile collection yet: the timings of steps that did not run go first
      hipchk(hipStreamSynchronize(st), "hipStreamSynchronize");
      // step s's force saw the positions step s-1 made (step 0: the chunk's)
      const double b = budget();
      double d = disp_cur, growth = 0.0;
      long ran = m;
      for (long s = 0; s < m; ++s) {
        if (!every && d > b) {
          ran = s;
          break;
        }
        const double dn = read_disp(s);
        growth = std::max(growth, dn - d);
        d = dn;
      }
      if ((m - ran) % 2 == 1)
        std::swap(x.p, xn.p);
      prof_discard_from(n_steps + ran);
      if (ran < m)
        replay_berendsen(bsaved, n_steps + 1, ran, dt);
      if (!pending.empty())
        prof_collect();
      if (ran < m)
        n_cut_chunks++;
      disp_cur = d;
      pos_version++;
      long next = std::min<long>(2 * chunk, kMaxChunk);
      if (growth > 0.0) {
        const double room = disp_cur > b ? b : b - disp_cur;
        next = std::min<long>(next, (long)(room / growth) + 2);
      }
      chunk = std::max<long>(1, next);
      cm.ucalls += ran;
      done += ran;
      n_steps += ran;
      forces_ok = false; // a is the force of the positions before x
      if (obs_every > 0 && (done % obs_every) == 0) {
        double e;
        read_energies(&e, nullptr, nullptr);
        const double ek = kinetic(nullptr);
        const long idx = done / obs_every - 1;
        if (ep_t)
          ep_t[idx] = e;
        if (ek_t)
          ek_t[idx] = ek;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// C ABI

namespace {

void need(const void* p, const char* what) {
  if (!p)
    fail(LPMDX_EINVAL, std::string(what) + " is NULL");
}

void finite(const double* p, long m, const char* what) {
  for (long k = 0; k < m; ++k)
    if (!std::isfinite(p[k]))
      fail(LPMDX_ENONFINITE, std::string(what) + " holds a non-finite value");
}

void ensure_device() {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0)
    fail(LPMDX_EDEVICE, "no HIP device available");
}

} // namespace

extern "C" {

const char* lpmdx_last_error(void) { return g_err.c_str(); }

int lpmdx_abi_version(void) { return LPMDX_ABI_VERSION; }

int lpmdx_device_count(int* count) {
  return guard([&] {
    need(count, "count");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess)
      c = 0;
    *count = c;
  });
}

int lpmdx_system_create(lpmdx_system** out, long n, const double box[3], const int* z,
                        const double* mass) {
  return guard([&] {
    need(out, "out");
    need(box, "box");
    *out = nullptr;
    if (n < 0 || n > (1L << 30))
      fail(LPMDX_EINVAL, "atom count out of range");
    if (n > 0)
      need(z, "z");
    finite(box, 3, "box");
    for (int q = 0; q < 3; ++q)
      if (!(box[q] > 0.0))
        fail(LPMDX_EINVAL, "box edges must be positive");
    ensure_device();
    auto S = std::make_unique<lpmdx_system>();
    S->n = n;
    for (int q = 0; q < 3; ++q) {
      S->box.L[q] = box[q];
      S->box.half[q] = 0.5 * box[q];
      S->box.inv[q] = 1.0 / box[q];
      S->ll[q] = edge_module(box[q]);
    }
    hipchk(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking), "hipStreamCreate");
    hipchk(hipHostMalloc(&S->hctrl, sizeof(Ctrl)), "hipHostMalloc");
    hipchk(hipHostMalloc(&S->hred, 64 * 12 * sizeof(double)), "hipHostMalloc");
    const size_t m = n > 0 ? n : 1;
    for (DBuf<double>* b : {&S->x, &S->xn, &S->v, &S->a, &S->asum, &S->xb, &S->x2, &S->v2, &S->a2})
      b->alloc(3 * m);
    for (DBuf<double>* b : {&S->ffm, &S->mass, &S->ffm2, &S->mass2})
      b->alloc(m);
    for (DBuf<int>* b : {&S->z, &S->perm, &S->z2, &S->perm2})
      b->alloc(m);
    for (DBuf<uint32_t>* b : {&S->keys, &S->vals, &S->keys2, &S->vals2})
      b->alloc(m);
    S->ctrl.alloc(1);
    S->red.alloc(12 * 64);
    hipchk(hipMemsetAsync(S->ctrl.p, 0, sizeof(Ctrl), S->st), "memset");
    std::vector<double> ms(m, 1.0), ff(m, 0.0);
    std::vector<int> zz(m, 0), pm(m);
    for (long i = 0; i < n; ++i) {
      zz[i] = z[i];
      double mi = mass ? mass[i] : ElemMass(z[i]);
      if (!(mi > 0.0) || !std::isfinite(mi))
        fail(LPMDX_EINVAL, "no mass for atomic number " + std::to_string(z[i]));
      ms[i] = mi;
      ff[i] = kForceFactor / mi; // forcefactor / Mass(), pairpotential.cc:62
      pm[i] = (int)i;
    }
    hipchk(hipMemcpyAsync(S->mass.p, ms.data(), m * sizeof(double), hipMemcpyHostToDevice, S->st), "memcpy");
    hipchk(hipMemcpyAsync(S->ffm.p, ff.data(), m * sizeof(double), hipMemcpyHostToDevice, S->st), "memcpy");
    hipchk(hipMemcpyAsync(S->z.p, zz.data(), m * sizeof(int), hipMemcpyHostToDevice, S->st), "memcpy");
    hipchk(hipMemcpyAsync(S->perm.p, pm.data(), m * sizeof(int), hipMemcpyHostToDevice, S->st), "memcpy");
    for (DBuf<double>* b : {&S->x, &S->v, &S->a})
      hipchk(hipMemsetAsync(b->p, 0, 3 * m * sizeof(double), S->st), "memset");
    S->sync();
    *out = S.release();
  });
}

int lpmdx_system_destroy(lpmdx_system* sys) {
  return guard([&] { delete sys; });
}

long lpmdx_system_size(const lpmdx_system* sys) { return sys ? sys->n : -1; }

}

namespace {

// caller order <-> internal order through perm (internal -> caller)
void to_internal(lpmdx_system* S, const double* h, double* d) {
  const long n = S->n;
  std::vector<int> perm(n);
  hipchk(hipMemcpyAsync(perm.data(), S->perm.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
  S->sync();
  std::vector<double> tmp(3 * n);
  for (long i = 0; i < n; ++i)
    for (int q = 0; q < 3; ++q)
      tmp[3 * i + q] = h[3 * perm[i] + q];
  hipchk(hipMemcpyAsync(d, tmp.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, S->st), "memcpy");
  S->sync();
}

void to_caller(lpmdx_system* S, const double* d, double* h) {
  const long n = S->n;
  std::vector<int> perm(n);
  std::vector<double> tmp(3 * n);
  hipchk(hipMemcpyAsync(perm.data(), S->perm.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
  hipchk(hipMemcpyAsync(tmp.data(), d, 3 * n * sizeof(double), hipMemcpyDeviceToHost, S->st), "memcpy");
  S->sync();
  for (long i = 0; i < n; ++i)
    for (int q = 0; q < 3; ++q)
      h[3 * perm[i] + q] = tmp[3 * i + q];
}

} // namespace

extern "C" {

int lpmdx_system_set_state(lpmdx_system* S, const double* x, const double* v, const double* a) {
  return guard([&] {
    need(S, "system");
    if (S->n == 0)
      return;
    if (x)
      finite(x, 3 * S->n, "positions");
    if (v)
      finite(v, 3 * S->n, "velocities");
    if (a)
      finite(a, 3 * S->n, "accelerations");
    S->undrift();
    if (x) {
      to_internal(S, x, S->x.p);
      S->forces_ok = false;
      S->pos_version++;
    }
    if (v)
      to_internal(S, v, S->v.p);
    if (a) {
      to_internal(S, a, S->a.p);
      S->forces_ok = false;
    }
  });
}

int lpmdx_system_get_state(lpmdx_system* S, double* x, double* v, double* a) {
  return guard([&] {
    need(S, "system");
    if (S->n == 0)
      return;
    const double* xu = S->drifted ? S->xn.p : S->x.p;
    if (x)
      to_caller(S, xu, x);
    if (v)
      to_caller(S, S->v.p, v);
    if (a)
      to_caller(S, S->a.p, a);
  });
}

int lpmdx_set_cellmanager(lpmdx_system* S, const char* name, const char* args) {
  return guard([&] {
    need(S, "system");
    need(name, "name");
    std::string nm(name);
    CellMgr c;
    if (nm == "linkedcell") {
      // plugins/linkedcell.cc:17-41 (the mode keyword alone decides auto)
      Keywords k(args, {{"cutoff", "7.0"}, {"nx", "0"}, {"ny", "0"}, {"nz", "0"}, {"mode", "noauto"},
                        {"warn-outside", "true"}});
      c.kind = kCmLinkedCell;
      c.cutoff = k.d("cutoff");
      c.nx = k.i("nx");
      c.ny = k.i("ny");
      c.nz = k.i("nz");
      std::string mode = k.s("mode");
      c.mode_auto = (mode == "auto" || mode == "AUTO");
      c.warn_outside = k.b("warn-outside");
    } else if (nm == "verletlist") {
      // plugins/verletlist.cc:12-29; extcut = 1.08*cutoff (the unread extcutoff
      // keyword stays unread), `skin` switches to a skin-triggered list
      Keywords k(args, {{"cutoff", "0.0"}, {"extcutoff", "0.0"}, {"maxneighbors", "80"}, {"each", "5"}});
      c.kind = kCmVerlet;
      c.cutoff = k.d("cutoff");
      c.maxneighbors = k.i("maxneighbors");
      c.each = std::max(1, k.i("each"));
      c.extcut = 1.08 * c.cutoff;
      if (k.has("skin")) {
        c.skin = k.d("skin");
        if (!(c.skin >= 0.0))
          fail(LPMDX_EINVAL, "verletlist: skin must be non-negative");
      }
      if (fabs(c.cutoff) < 1E-10)
        fail(LPMDX_EINVAL, "verletlist: No cutoff was defined");
    } else if (nm == "minimumimage") {
      // plugins/minimumimage.cc:12-19; `skin` is this engine's extension, as
      // for verletlist (without it the lists are rebuilt on every UpdateCell)
      Keywords k(args, {{"cutoff", "0.0"}});
      c.kind = kCmMinImage;
      c.cutoff = k.d("cutoff");
      if (k.has("skin")) {
        c.skin = k.d("skin");
        if (!(c.skin >= 0.0))
          fail(LPMDX_EINVAL, "minimumimage: skin must be non-negative");
      }
    } else if (nm == "lcbinary") {
      // plugins/lcbinary.cc:9-33: automatic subdivisions when one is 0 or with mode auto
      Keywords k(args, {{"cutoff", "7.0"}, {"nx", "0"}, {"ny", "0"}, {"nz", "0"}, {"mode", "noauto"}});
      c.kind = kCmLcBinary;
      c.cutoff = k.d("cutoff");
      c.nx = k.i("nx");
      c.ny = k.i("ny");
      c.nz = k.i("nz");
      const std::string mode = k.s("mode");
      c.mode_auto = c.nx == 0 || c.ny == 0 || c.nz == 0 || mode == "auto" || mode == "AUTO";
    } else {
      fail(LPMDX_EINVAL, "unknown cell manager '" + nm + "'");
    }
    S->undrift();
    S->cm = c;
    S->list_built = false;
    S->forces_ok = false;
    S->update_cell(); // Configuration::SetCellManager calls UpdateCell
  });
}

int lpmdx_update_cell(lpmdx_system* S) {
  return guard([&] {
    need(S, "system");
    S->update_cell();
  });
}

int lpmdx_potential_append(lpmdx_system* S, const char* name, const char* args, int s1, int s2) {
  return guard([&] {
    need(S, "system");
    need(name, "name");
    S->pots.push_back(make_potential(name, args, s1, s2));
    S->forces_ok = false;
    S->lists_follow_potentials();
    S->undrift();
  });
}

int lpmdx_potential_clear(lpmdx_system* S) {
  return guard([&] {
    need(S, "system");
    S->pots.clear();
    S->forces_ok = false;
    S->lists_follow_potentials();
    S->undrift();
  });
}

int lpmdx_update_forces(lpmdx_system* S, double* epot, double* virial, double stress[9]) {
  return guard([&] {
    need(S, "system");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (S->pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (S->n == 0) {
      if (epot)
        *epot = 0.0;
      if (virial)
        *virial = 0.0;
      if (stress)
        for (int k = 0; k < 9; ++k)
          stress[k] = 0.0;
      return;
    }
    S->update_cell();
    S->forces(kForcesOnly, true, 0.0, S->ctrl.p->maxd2);
    S->read_energies(epot, virial, stress);
    S->forces_ok = true;
  });
}

int lpmdx_potential_forces(lpmdx_system* S, int k, double* acc, double* epot, double* virial, double stress[9]) {
  return guard([&] {
    need(S, "system");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (k < 0 || k >= (int)S->pots.size())
      fail(LPMDX_EINVAL, "no potential with index " + std::to_string(k));
    double e = 0.0, w = 0.0, t[9] = {0};
    if (S->n > 0) {
      S->undrift();
      if (!S->list_built || S->cell_version != S->pos_version)
        S->update_cell();
      S->forces_term(k);
      const double* pr = S->hred;
      e = pr[0] + pr[12 + 1];
      w = pr[2];
      for (int q = 0; q < 9; ++q)
        t[q] = pr[3 + q];
      if (acc)
        to_caller(S, S->asum.p, acc);
    }
    if (epot)
      *epot = e;
    if (virial)
      *virial = w;
    if (stress)
      for (int q = 0; q < 9; ++q)
        stress[q] = t[q];
  });
}

int lpmdx_velocityverlet(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t,
                         double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->vv(dt, steps, obs_every, epot_t, ekin_t);
  });
}

int lpmdx_leapfrog_init(lpmdx_system* S, double dt) {
  return guard([&] {
    need(S, "system");
    S->lf_init(dt);
  });
}

int lpmdx_leapfrog(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t, double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->lf(dt, steps, obs_every, epot_t, ekin_t);
  });
}

int lpmdx_beeman_init(lpmdx_system* S, double dt) {
  return guard([&] {
    need(S, "system");
    S->bm_init(dt);
  });
}

int lpmdx_beeman(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t, double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->vv(dt, steps, obs_every, epot_t, ekin_t, true);
  });
}

int lpmdx_set_temperature(lpmdx_system* S, double temp) {
  return guard([&] {
    need(S, "system");
    if (!std::isfinite(temp) || temp < 0.0)
      fail(LPMDX_EINVAL, "temperature must be finite and non-negative");
    if (S->n == 0)
      return;
    S->undrift(); // the look-ahead positions were made from the old velocities
    S->partials.alloc(std::max<size_t>((size_t)blocks_for(S->n) * 4, S->partials.n));
    S->scale_to_temperature(temp);
    S->sync();
  });
}

int lpmdx_tempscaling(lpmdx_system* S, double from, double to, long start, long end, long each) {
  return guard([&] {
    need(S, "system");
    if (each == 0) {
      S->tsched.on = false;
      return;
    }
    if (!std::isfinite(from) || !std::isfinite(to) || from < 0.0 || to < 0.0)
      fail(LPMDX_EINVAL, "tempscaling: temperatures must be finite and non-negative");
    if (each < 0 || start < 0 || end < -1 || (end != -1 && end < start))
      fail(LPMDX_EINVAL, "tempscaling: need each >= 1, start >= 0 and end == -1 or end >= start");
    if (end == -1 && from != to)
      fail(LPMDX_EINVAL, "tempscaling: a ramp (from != to) needs its last step as end");
    S->tsched.on = true;
    S->tsched.from = from;
    S->tsched.to = to;
    S->tsched.start = start;
    S->tsched.end = end;
    S->tsched.each = each;
  });
}

int lpmdx_berendsen(lpmdx_system* S, double from, double to, double tau, long start, long end, long each) {
  return guard([&] {
    need(S, "system");
    if (each == 0) {
      S->bsched = {};
      return;
    }
    if (!std::isfinite(from) || !std::isfinite(to) || from < 0.0 || to < 0.0 || !std::isfinite(tau) || !(tau > 0.0))
      fail(LPMDX_EINVAL, "berendsen: temperatures must be finite and non-negative, tau positive");
    if (each < 0 || start < 0 || end < -1 || (end != -1 && end < start))
      fail(LPMDX_EINVAL, "berendsen: need each >= 1, start >= 0 and end == -1 or end >= start");
    if (end == -1 && from != to)
      fail(LPMDX_EINVAL, "berendsen: a ramp (from != to) needs its last step as end");
    S->bsched = {};
    S->bsched.on = true;
    S->bsched.from = from;
    S->bsched.to = to;
    S->bsched.tau = tau;
    S->bsched.start = start;
    S->bsched.end = end;
    S->bsched.each = each;
  });
}

int lpmdx_observables(lpmdx_system* S, double out[LPMDX_NOBS]) {
  return guard([&] {
    need(S, "system");
    need(out, "out");
    for (int k = 0; k < LPMDX_NOBS; ++k)
      out[k] = 0.0;
    if (S->n == 0)
      return;
    double p[3];
    const double ke = S->kinetic(p);
    const double vol = S->ll[0] * S->ll[1] * S->ll[2];
    out[LPMDX_OBS_KINETIC] = ke;
    out[LPMDX_OBS_TEMP] = (2.0 / 3.0) * ke / (kBoltzmann * double(S->n));
    for (int q = 0; q < 3; ++q)
      out[LPMDX_OBS_PX + q] = p[q];
    if (S->last_e_valid) {
      out[LPMDX_OBS_POTENTIAL] = S->last_e[0];
      out[LPMDX_OBS_VIRIAL] = S->last_e[1];
      out[LPMDX_OBS_PRESSURE] = (kPressFactor / vol) * ((2.0 / 3.0) * ke + (1.0 / 3.0) * S->last_e[1]);
      for (int k = 0; k < 9; ++k)
        out[LPMDX_OBS_STRESS + k] = (kPressFactor / vol) * S->last_e[2 + k];
    }
  });
}

int lpmdx_system_stats(const lpmdx_system* S, long out[8]) {
  return guard([&] {
    need(S, "system");
    need(out, "out");
    out[0] = S->n_builds;
    out[1] = S->n_forces;
    out[2] = S->n_steps;
    out[3] = S->cap;
    out[4] = S->max_list;
    out[5] = S->lpa;
    out[6] = S->staged ? S->grid.brick : 0;
    out[7] = S->sum_list;
  });
}

int lpmdx_system_profile(lpmdx_system* S, int enable) {
  return guard([&] {
    need(S, "system");
    S->sync();
    S->profiling = enable != 0;
    for (int k = 0; k < kProfSlots; ++k) {
      S->prof_ms[k] = 0.0;
      S->prof_n[k] = 0;
    }
  });
}

int lpmdx_system_profile_read(lpmdx_system* S, double ms[4], long launches[4]) {
  return guard([&] {
    need(S, "system");
    need(ms, "ms");
    need(launches, "launches");
    S->sync();
    for (int k = 0; k < kProfSlots; ++k) {
      ms[k] = S->prof_ms[k];
      launches[k] = S->prof_n[k];
    }
  });
}

int lpmdx_system_set_lanes(lpmdx_system* S, int lanes) {
  return guard([&] {
    need(S, "system");
    if (!(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16))
      fail(LPMDX_EINVAL, "lanes per atom must be 0, 1, 2, 4, 8 or 16");
    S->lpa_req = lanes;
    S->list_built = false;
    S->cap = 0;
  });
}

int lpmdx_neighbor_lists(lpmdx_system* S, int full, double rcut, long* counts, long cap, long* js,
                         double* r2, long* total) {
  return guard([&] {
    need(S, "system");
    need(counts, "counts");
    need(total, "total");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!S->list_built)
      S->update_cell();
    S->undrift();
    const long n = S->n;
    if (n == 0) {
      *total = 0;
      return;
    }
    if (S->cm.kind == kCmMinImage && rcut < 1.0E-10)
      fail(LPMDX_EINVAL, "MinimumImage cutoff equals zero..."); // minimumimage.cc:76-77
    if (S->cm.kind == kCmLinkedCell && rcut > S->rl) {
      // the cached lists are built on the first request's rcut (linkedcell.cc:315-354)
      S->rl = rcut;
      S->build_list();
    }
    if ((S->cm.kind == kCmMinImage || S->cm.kind == kCmLcBinary) && rcut > S->rl) {
      // these list afresh with the caller's radius: lists of that radius (the
      // next UpdateCell returns to the potentials' radius)
      S->rl_min = rcut;
      S->rebuild();
      S->rl_min = 0.0;
    }
    std::vector<int> perm(n), cntv(n), nb((size_t)S->cap * S->stride);
    std::vector<double> x(3 * n);
    hipchk(hipMemcpyAsync(perm.data(), S->perm.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(cntv.data(), S->cnt.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(nb.data(), S->nbr.p, nb.size() * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(x.data(), S->x.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, S->st), "memcpy");
    const int H = S->staged ? halo_cells(S->grid.brick) : 0;
    std::vector<int2> bh(S->staged ? (size_t)S->nbrick * (H + 1) : 0);
    std::vector<uint32_t> keys(S->staged ? n : 0);
    if (S->staged) {
      hipchk(hipMemcpyAsync(bh.data(), S->bhalo.p, bh.size() * sizeof(int2), hipMemcpyDeviceToHost, S->st),
             "memcpy");
      hipchk(hipMemcpyAsync(keys.data(), S->keys.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, S->st), "memcpy");
    }
    S->sync();
    // staged lists hold slots of the atom's brick halo: back to sorted indices
    const int B3 = S->grid.brick * S->grid.brick * S->grid.brick;
    auto atom_of = [&](long i, long e) -> long {
      if (!S->staged)
        return e;
      const int2* row = bh.data() + (size_t)(keys[i] / B3) * (H + 1);
      int h = 0; // the last halo cell whose first slot is <= e
      while (h + 1 < H && row[h + 1].x <= e)
        ++h;
      return (long)row[h].y + (e - row[h].x);
    };
    std::vector<long> iperm(n);
    for (long i = 0; i < n; ++i)
      iperm[perm[i]] = i;
    const double rc2 = rcut * rcut;
    long k = 0;
    std::vector<std::pair<long, double>> row;
    for (long o = 0; o < n; ++o) {
      const long i = iperm[o];
      row.clear();
      for (int c = 0; c < cntv[i]; ++c) {
        const long jj = atom_of(i, nb[list_index(i, c, S->lpa, S->cap)]);
        const long oj = perm[jj];
        if (!full && S->stencil_grid() && oj > o)
          continue; // linkedcell / lcbinary half lists keep j < i (linkedcell.cc:323, lcbinary.cc:291)
        if (!full && (S->cm.kind == kCmVerlet || S->cm.kind == kCmMinImage) && oj < o)
          continue; // verletlist / minimumimage half lists keep j > i (verletlist.cc:157, minimumimage.cc:90)
        double d[3];
        for (int q = 0; q < 3; ++q) {
          d[q] = x[3 * jj + q] - x[3 * i + q];
          if (d[q] >= S->box.half[q])
            d[q] -= S->box.L[q];
          else if (d[q] < -S->box.half[q])
            d[q] += S->box.L[q];
        }
        const double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (rr < rc2)
          row.emplace_back(oj, rr);
      }
      if (S->cm.kind == kCmVerlet || S->cm.kind == kCmMinImage)
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<long, double>& p, const std::pair<long, double>& q) {
                           return p.first < q.first;
                         });
      counts[o] = (long)row.size();
      for (auto& e : row) {
        if (k < cap) {
          if (js)
            js[k] = e.first;
          if (r2)
            r2[k] = e.second;
        }
        ++k;
      }
    }
    *total = k;
  });
}

} // extern "C"

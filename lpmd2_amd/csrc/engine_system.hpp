// Host engine of the MI355X LPMD force-and-integration path: one lpmdx_system
// is one Configuration resident in HBM (positions, velocities, accelerations
// in AoS FP64, internally kept in cell-sorted order), with its cell manager,
// its potentials and its own HIP stream.  The C ABI over it is in engine.hip;
// the force-kernel launchers of each potential kind are in engine_launch.hpp.
#pragma once

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


// timing slots of lpmdx_system_profile_read
enum ProfSlot { kProfForce = 0, kProfRho = 1, kProfBuild = 2, kProfOther = 3, kProfSlots = 4 };

// Linkedcell / lcbinary stencils of systems up to this many atoms are listed
// one wavefront per atom (k_build_list_lc_wave) rather than one thread per atom.
constexpr long kWaveBuildMaxAtoms = 32768;

namespace {

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

// End of a chunk of m speculative steps whose list is rebuilt before the host
// has read the displacement slots: the steps that ran (a step whose gate was
// shut closed its own slot, and every later step's, so the closed slots are a
// suffix of md[0, m); step 0 always runs) wrote their positions alternately
// into the two buffers the host swapped after every launch.  When an odd
// number of launches returned at once, the last positions written are in the
// buffer the host now calls xn: they are copied into x, so that the host's x
// holds them whatever the count.
__global__ __launch_bounds__(kBlock) void k_settle_positions(const unsigned int* __restrict__ md, long m,
                                                              double* __restrict__ x,
                                                              const double* __restrict__ xn, long n3) {
  // ran = the first closed slot (binary search over the closed suffix), or m
  long lo = 1, hi = m;
  while (lo < hi) {
    const long mid = (lo + hi) / 2;
    if (md[mid] == kGateClosed)
      hi = mid;
    else
      lo = mid + 1;
  }
  if (((m - lo) & 1) == 0)
    return;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t < n3)
    x[t] = xn[t];
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
// n_eff: the atoms the temperature counts; fix / skip: atoms whose flags
// have a bit of skip set are left as they are (null: none).
__global__ __launch_bounds__(kBlock) void k_set_temperature(double* __restrict__ v, const double* __restrict__ red,
                                                            long n, double temp, double q,
                                                            const unsigned int* gate, double gate_b, long n_eff,
                                                            const int* __restrict__ fix, int skip) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  if (gate && (*gate == kGateClosed || sqrt((double)__uint_as_float(*gate)) > gate_b))
    return;
  if (fix && (fix[i] & skip))
    return;
  const double ti = (2.0 / 3.0) * (red[0] * kKin2eV) / (kBoltzmann * double(n_eff));
  if (!(ti > 0.0))
    return;
  // q < 0: SetTemperature; q = dt / tau: the berendsen modifier's factor
  // (plugins/berendsen.cc:62-70)
  const double f = q < 0.0 ? sqrt(temp / ti) : sqrt(1.0 + q * (temp / ti - 1.0));
#pragma unroll
  for (int q = 0; q < 3; ++q)
    v[3 * i + q] *= f;
}

// v of slot k = vc of the atom's caller index perm[k] (caller-order
// velocities put in place on the stream); gate as k_set_temperature's.
__global__ __launch_bounds__(kBlock) void k_gather_velocities(double* __restrict__ v, const double* __restrict__ vc,
                                                              const int* __restrict__ perm, long n,
                                                              const unsigned int* gate, double gate_b) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  if (gate && (*gate == kGateClosed || sqrt((double)__uint_as_float(*gate)) > gate_b))
    return;
  const long c = perm[i];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    v[3 * i + q] = vc[3 * c + q];
}

// Slot-order [3n] rows into caller order: out[perm[i]] = d[i].
__global__ __launch_bounds__(kBlock) void k_to_caller(const double* __restrict__ d, const int* __restrict__ perm,
                                                      long n, double* __restrict__ out) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const long c = perm[i];
  const double r[3] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    out[3 * c + q] = r[q];
}

// Caller-order [3n] rows into slot order: d[i] = h[perm[i]].
__global__ __launch_bounds__(kBlock) void k_from_caller(const double* __restrict__ h, const int* __restrict__ perm,
                                                        long n, double* __restrict__ d) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const long c = perm[i];
  const double r[3] = {h[3 * c], h[3 * c + 1], h[3 * c + 2]};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    d[3 * i + q] = r[q];
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
    if constexpr (eam_force_capped<KIND, EVAL, EXP>()) {
      static const bool ok = allow_stage_lds((const void*)k_eam_force_bk<KIND, LPA, EVAL, SPC, EXP>);
      (void)ok;
      k_eam_force_bk<KIND, LPA, EVAL, SPC, EXP><<<L.bricks, L.threads, L.lds, L.st>>>(A, L.bhalo, L.brick, L.cap,
                                                                                        parts);
    } else {
      static const bool ok = allow_stage_lds((const void*)k_eam_force_bk_wide<KIND, LPA, EVAL, SPC, EXP>);
      (void)ok;
      k_eam_force_bk_wide<KIND, LPA, EVAL, SPC, EXP><<<L.bricks, L.threads, L.lds, L.st>>>(A, L.bhalo, L.brick,
                                                                                             L.cap, parts);
    }
  }
};

// ---------------------------------------------------------------------------
// the system

struct lpmdx_system {
  int device = 0; // the HIP device current when the system was created
  long n = 0;
  Box box{};
  double ll[3] = {0, 0, 0};
  hipStream_t st = nullptr;
  // state, internal order
  DBuf<double> x, xn, v, a, ffm, mass, asum, xb;
  DBuf<int> z, perm; // perm[k]: caller index of the atom in slot k
  // fixed-atom flags (kFixVel | kFixPos | kFixTagged per atom, internal
  // order) while some atom carries the fixedvel or fixedpos tag
  DBuf<int> fix, fix2;
  bool any_fix = false; // some atom has kFixVel or kFixPos
  bool fixpos_any = false; // some atom has kFixPos
  long n_tagged = 0;    // atoms with kFixTagged
  // SimulationEngine's velocitiesSet (simulationbuilder.cc:54-55,184,290) and
  // the system's drand48 stream for InitVelocities; masses in caller order
  bool velocities_set = false;
  Drand48 rng48;
  long vinit_step = -1; // the step whose modifier drew the velocities
  std::vector<double> hmass, hvinit;
  // atomic numbers and the fixed-atom flags last set, in caller order
  std::vector<int> hz, hfix;
  DBuf<double> vinit;
  DBuf<double> eatom; // per-atom energies of lpmdx_atom_energies
  const int* fix_arg() const { return any_fix ? fix.p : nullptr; }
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
  // caller-order copies of x, v, a on the device and their pinned host stage
  DBuf<double> cstage;
  double* hstage = nullptr;
  size_t hstage_n = 0;
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
  // Nose-Hoover (plugins/nosehoover.cc): ab holds old.a, nn the
  // friction-modified acceleration of the position update still to be ended,
  // nh = {friction before the step's increments, increment per atom}
  DBuf<double> nn, nn2, nh;
  bool nh_ready = false, drift_nose = false;
  double nh_dt = 0.0, nh_q = 0.0, nh_t = 0.0, nh_kb = 0.0;
  // moving atoms (no fixedpos "true"): their count, and per caller index the
  // number of moving atoms before it (filled while any atom is held)
  long n_moving = 0;
  DBuf<int> mv_rank;
  double bm_dt = 0.0;
  double lf_dt = 0.0;
  // more arrays that reorders carry along (with their scratch twins)
  std::vector<std::pair<DBuf<double>*, DBuf<double>*>> carry;
  DBuf<double> keep_x, keep_x2, keep_a, keep_a2;
  long pos_version = 0;   // bumped whenever the user positions change
  long cell_version = -1; // pos_version at the last UpdateCell
  double disp_cur = 0.0; // largest |x - xb| of the positions of the next force evaluation
  long chunk = 8;        // steps of the next velocity-Verlet chunk
  bool chunk_by_growth = false; // chunk was sized by the displacement growth (expected to end past the budget)
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
  // optional per-kernel timing with HIP events on this system's stream:
  // every launch outside the integrator steps (list builds), and the
  // launches of every prof_every-th step
  bool profiling = false;
  long prof_every = 1;
  struct Pending {
    int slot;
    hipEvent_t a, b;
    long step;  // the integrator step the launch belongs to, or -1
    bool own_a; // false: a is the previous entry's b (one event between two launches)
  };
  std::vector<Pending> pending;
  long prof_step = -1; // tag of the launches being recorded

  // Timings of speculative steps that never ran (their kernels returned at
  // once) leave the profile: every entry tagged with a step >= first.
  void prof_discard_from(long first) {
    size_t k = 0;
    for (auto& p : pending) {
      if (p.step >= first) {
        if (p.own_a)
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
    // Timing-only events: the host reads them after a stream synchronisation,
    // so they need no system-scope release, whose L2 writeback and
    // invalidation would add several microseconds to every bracketed launch.
    hipEvent_t e;
    hipchk(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
    return e;
  }
  // brackets work enqueued between begin() and end() on the stream
  hipEvent_t prof_begin() {
    if (!profiling || (prof_step >= 0 && prof_step % prof_every != 0))
      return nullptr;
    hipEvent_t e = take_event();
    hipchk(hipEventRecord(e, st), "hipEventRecord");
    return e;
  }
  // own_a: a came from prof_begin (false: from prof_next, owned by the
  // entry whose end it is)
  void prof_end(int slot, hipEvent_t a, bool own_a = true) {
    if (!profiling || !a)
      return;
    hipEvent_t e = take_event();
    hipchk(hipEventRecord(e, st), "hipEventRecord");
    pending.push_back({slot, a, e, prof_step, own_a});
  }
  // End of a bracketed launch whose successor is enqueued right after it:
  // the end event is returned as the successor's begin event (one event
  // between two launches instead of two); close the successor with
  // prof_end(slot, event, false).
  hipEvent_t prof_next(int slot, hipEvent_t a, bool own_a = true) {
    if (!profiling || !a)
      return nullptr;
    prof_end(slot, a, own_a);
    return pending.back().b;
  }
  void prof_collect() {
    for (auto& p : pending) {
      float ms = 0.0f;
      hipchk(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      prof_ms[p.slot] += ms;
      prof_n[p.slot] += 1;
      if (p.own_a)
        event_pool.push_back(p.a);
      event_pool.push_back(p.b);
    }
    pending.clear();
  }

  ~lpmdx_system() {
    if (st)
      (void)hipStreamSynchronize(st);
    for (auto& p : pending) {
      if (p.own_a)
        event_pool.push_back(p.a);
      event_pool.push_back(p.b);
    }
    for (auto e : event_pool)
      (void)hipEventDestroy(e);
    if (hctrl)
      (void)hipHostFree(hctrl);
    if (hred)
      (void)hipHostFree(hred);
    if (hstage)
      (void)hipHostFree(hstage);
    pots.clear();
    if (st)
      (void)hipStreamDestroy(st);
  }

  // Room for k caller-order [3n] arrays in the device and pinned stages.
  void stage(int k) {
    const size_t m = (size_t)3 * n * k;
    cstage.alloc(m);
    if (hstage_n < m) {
      if (hstage)
        (void)hipHostFree(hstage);
      hstage = nullptr;
      hstage_n = 0;
      hipchk(hipHostMalloc(&hstage, m * sizeof(double)), "hipHostMalloc");
      hstage_n = m;
    }
  }

  // Copies up to three slot-order [3n] device arrays into caller-order host
  // arrays (null entries skipped): one gather kernel each into a device
  // stage, one copy into pinned memory, then into the caller's arrays.
  void download(const double* const src[3], double* const dst[3]) {
    int k = 0;
    for (int j = 0; j < 3; ++j)
      k += (src[j] && dst[j]) ? 1 : 0;
    if (n == 0 || k == 0)
      return;
    const size_t m = (size_t)3 * n;
    stage(k);
    int s = 0;
    for (int j = 0; j < 3; ++j)
      if (src[j] && dst[j]) {
        k_to_caller<<<blocks_for(n), kBlock, 0, st>>>(src[j], perm.p, n, cstage.p + m * s);
        ++s;
      }
    launch_check();
    hipchk(hipMemcpyAsync(hstage, cstage.p, m * k * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy");
    sync();
    s = 0;
    for (int j = 0; j < 3; ++j)
      if (src[j] && dst[j]) {
        memcpy(dst[j], hstage + m * s, m * sizeof(double));
        ++s;
      }
  }

  // The reverse: caller-order host arrays into slot-order device arrays.
  void upload(const double* const src[3], double* const dst[3]) {
    int k = 0;
    for (int j = 0; j < 3; ++j)
      k += (src[j] && dst[j]) ? 1 : 0;
    if (n == 0 || k == 0)
      return;
    const size_t m = (size_t)3 * n;
    stage(k);
    int s = 0;
    for (int j = 0; j < 3; ++j)
      if (src[j] && dst[j]) {
        memcpy(hstage + m * s, src[j], m * sizeof(double));
        ++s;
      }
    hipchk(hipMemcpyAsync(cstage.p, hstage, m * k * sizeof(double), hipMemcpyHostToDevice, st), "memcpy");
    s = 0;
    for (int j = 0; j < 3; ++j)
      if (src[j] && dst[j]) {
        k_from_caller<<<blocks_for(n), kBlock, 0, st>>>(cstage.p + m * s, perm.p, n, dst[j]);
        ++s;
      }
    launch_check();
    // the pinned stage is reused by the next call
    sync();
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
                  ffm2.p, mass2.p, perm2.p, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, n,
                  nullptr, nullptr};
    const bool move_fix = any_fix || n_tagged > 0;
    if (move_fix) {
      fix2.alloc(n);
      P.fix = fix.p;
      P.fix2 = fix2.p;
    }
    // extra per-atom vectors (the leapfrog half-step velocities, and during
    // its set-up the saved positions and accelerations) follow the atoms
    std::vector<std::pair<DBuf<double>*, DBuf<double>*>> ex;
    if (lf_ready)
      ex.push_back({&vh, &vh2});
    if (bm_ready || nh_ready)
      ex.push_back({&ab, &ab2});
    if (nh_ready)
      ex.push_back({&nn, &nn2});
    for (auto& e : carry)
      ex.push_back(e);
    if (ex.size() > (size_t)kPermuteExtra)
      fail(LPMDX_EDEVICE, "too many per-atom arrays to reorder (internal error)");
    for (size_t e = 0; e < ex.size(); ++e) {
      ex[e].second->alloc(3 * n);
      P.ex[e] = ex[e].first->p;
      P.ex2[e] = ex[e].second->p;
    }
    // the list build's inputs: the build positions and their FP32 copy
    xf.alloc(n + kGroup); // padding read by the candidate groups
    P.xb = xb.p;
    P.xf = xf.p;
    P.b = box;
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
    if (move_fix)
      std::swap(fix.p, fix2.p);
    // keep the sorted keys where the build kernel reads them
    if (ka != keys.p) {
      std::swap(keys.p, keys2.p);
      std::swap(vals.p, vals2.p);
    }
  }

  // Set-up of a list build: lanes, the staging choice, room for the partial
  // sums, the brick halo tables and a first capacity estimate.
  void build_prepare() {
    // list storage: cap entries per atom, rounded to the lane count
    lpa = auto_lanes();
    stride = ((n + kWave - 1) / kWave) * kWave; // whole tiles for every lane count
    // staged kernels serve up to 4 lanes per atom
    // (and only with a wrap margin: their neighbour loop has no compare-and-select wrap)
    staged = grid.brick > 0 && lpa <= 4 && wrap_margin() > 0.0;
    const size_t np = (size_t)std::max<long>(blocks_for(n * 16), nbrick) * 12;
    partials.alloc(np);
    partials2.alloc(np);
    // the build's counters, halo_max and brick_max among them (k_brick_halo's)
    sumslots.alloc((size_t)kSumSlots * kSlotStride);
    k_build_reset<<<blocks_for(kSumSlots * kSlotStride), kBlock, 0, st>>>(&ctrl.p->max_count, 1, sumslots.p,
                                                                          kSumSlots * kSlotStride);
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
  }
  // One build with the current capacity, ending with the copy of its control
  // words (max_count, pad0, sum_count, halo_max, brick_max) to the host; no
  // host synchronisation.  reset: clear the counters again (a repeated
  // attempt; build_prepare cleared them for the first).  fresh_xf: the sort
  // just wrote the FP32 copy of the positions (k_permute).
  void build_attempt(bool reset, bool fresh_xf) {
    cap = ((cap + kChunk - 1) / kChunk) * kChunk; // whole 8-entry chunks
    nbr.alloc((size_t)cap * stride);
    cnt.alloc(n);
    if (reset)
      k_build_reset<<<blocks_for(kSumSlots * kSlotStride), kBlock, 0, st>>>(&ctrl.p->max_count, 0, sumslots.p,
                                                                            kSumSlots * kSlotStride);
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
    if (!fresh_xf) {
      xf.alloc(n + kGroup); // padding read by the candidate groups
      k_to_f32<<<blocks_for(n), kBlock, 0, st>>>(x.p, n, box, xf.p);
    }
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
    if (grid.mode == kLinkedCell && !staged && n <= kWaveBuildMaxAtoms) {
      // few atoms on a linkedcell stencil: a wavefront per atom
      if (path == kPathExact64)
        k_build_list_lc_wave<true><<<blocks_for(n * kWave), kBlock, 0, st>>>(B);
      else
        k_build_list_lc_wave<false><<<blocks_for(n * kWave), kBlock, 0, st>>>(B);
    } else if (grid.mode == kExact && path == kPathShift32) {
      k_build_list_shift32<<<blocks_for(n), kBlock, 0, st>>>(B);
    } else {
      k_build_list<<<blocks_for(n), kBlock, 0, st>>>(B);
    }
    prof_end(kProfBuild, e0);
    k_sum_slots<<<1, kWave, 0, st>>>(sumslots.p, &ctrl.p->sum_count);
    launch_check();
    // max_count, pad0, sum_count, halo_max, brick_max: 24 contiguous bytes
    hipchk(hipMemcpyAsync(&hctrl->max_count, &ctrl.p->max_count, 24, hipMemcpyDeviceToHost, st),
           "memcpy");
  }
  // After the stream synchronised past build_attempt(): true when the list
  // fits its capacity (and the staged kernels' halo fits LDS), with halo_cap,
  // brick_threads, max_list and sum_list set; false after the capacity or the
  // staging choice was changed for another attempt.
  bool build_check() {
    if (staged) {
      halo_cap = (int)(((hctrl->halo_max + 63) / 64) * 64);
      const int lanes_atoms = (int)hctrl->brick_max * lpa;
      brick_threads = std::max(kWave, std::min(1024, ((lanes_atoms + kWave - 1) / kWave) * kWave));
      if (stage_bytes(halo_cap, stage_parts()) > (size_t)kStageLdsLimit) {
        // a halo too large for LDS (a dense spot): global indices this time
        // (the grid keeps its brick order; the unstaged kernels take any order)
        staged = false;
        halo_cap = 0;
        return false;
      }
    }
    const long mc = hctrl->max_count;
    max_list = mc;
    sum_list = (long)hctrl->sum_count;
    if (mc <= cap)
      return true;
    cap = (int)(mc * 1.2 + 16);
    return false;
  }

  // fresh_xf: see build_attempt
  void build_list(bool fresh_xf = false) {
    build_prepare();
    for (int attempt = 0; attempt < 4; ++attempt) {
      build_attempt(attempt > 0, fresh_xf);
      sync();
      if (build_check())
        return;
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
    // the sort wrote the build positions xb and the FP32 copy the build reads
    build_list(true);
    // verletlist.cc:96-97: a list may hold at most maxneighbors-1 entries
    if (cm.kind == kCmVerlet && max_list >= cm.maxneighbors)
      fail(LPMDX_ECAPACITY, "verletlist: maxneighbors is too small!");
    list_built = true;
    n_builds++;
  }

  // rebuild() enqueued behind work still running, with no host
  // synchronisation: the sort (which writes the build positions) and one
  // build attempt at the current capacity.  rebuild_settle() finishes it once the
  // stream has synchronised.  Not for lcbinary, whose grid growth needs the
  // host between sorts.
  bool can_rebuild_ahead() const { return n > 0 && cm.kind != kCmLcBinary && list_built && cap > 0; }
  void rebuild_ahead() {
    rl = list_radius();
    setup_grid();
    hipEvent_t e0 = prof_begin();
    sort_atoms();
    prof_end(kProfBuild, e0);
    build_prepare();
    build_attempt(false, true);
  }
  // After the synchronisation: the list of rebuild_ahead() is kept when it fit
  // its capacity; otherwise the build runs again at the same positions with
  // the larger capacity (rebuild(), which synchronises).
  void rebuild_settle() {
    if (!build_check()) {
      rebuild();
      return;
    }
    if (cm.kind == kCmVerlet && max_list >= cm.maxneighbors)
      fail(LPMDX_ECAPACITY, "verletlist: maxneighbors is too small!");
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

  // Force launches of one potential kind (engine_launch.hpp; instantiated
  // per kind in engine_kinds_*.hip)
  template <int KIND, bool EVAL> void launch_kind(ForceArgs& A, long nb, double rc2_rho);
  template <int KIND, bool EVAL, bool SPC> void launch_pair_s(ForceArgs& A, long nb);
  template <int KIND, bool EVAL, bool SPC> bool launch_sc_fixed(ForceArgs& A, ForceArgs& R, long nb);
  template <int KIND, bool SPC, int EXP> bool launch_eam_fixed(ForceArgs& A, ForceArgs& R, long nb);
  template <int KIND, bool EVAL, bool SPC> void launch_eam_s(ForceArgs& A, long nb, double rc2_rho);
  template <bool EVAL> void launch_pot(int kind, ForceArgs& A, long nb, double rc2_rho) {
    switch (kind) {
    case kLJ: launch_kind<kLJ, EVAL>(A, nb, rc2_rho); break;
    case kFastLJ: launch_kind<kFastLJ, EVAL>(A, nb, rc2_rho); break;
    case kMorse: launch_kind<kMorse, EVAL>(A, nb, rc2_rho); break;
    case kBuckingham: launch_kind<kBuckingham, EVAL>(A, nb, rc2_rho); break;
    case kSuttonChen: launch_kind<kSuttonChen, EVAL>(A, nb, rc2_rho); break;
    case kFinnisSinclair: launch_kind<kFinnisSinclair, EVAL>(A, nb, rc2_rho); break;
    case kGupta: launch_kind<kGupta, EVAL>(A, nb, rc2_rho); break;
    case kHarmonic: launch_kind<kHarmonic, EVAL>(A, nb, rc2_rho); break;
    case kLJMod: launch_kind<kLJMod, EVAL>(A, nb, rc2_rho); break;
    case kFinnisSinclairExt: launch_kind<kFinnisSinclairExt, EVAL>(A, nb, rc2_rho); break;
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
      A.nn = nn.p;
      A.fix = fix_arg();
      A.margin = wrap_margin();
      for (int q = 0; q < 3; ++q)
        A.margin_hi[q] = box.L[q] - A.margin;
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
      A.nn = nn.p;
      A.fix = fix_arg();
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
  void forces_term(int k, double* eatom = nullptr) {
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
    for (int q = 0; q < 3; ++q)
      A.margin_hi[q] = box.L[q] - A.margin;
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
    A.eatom = eatom;
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

  // Kinetic energy (eV) and momentum; tagged: the tag-aware sums of
  // SetSimulationTags (simulationbuilder.cc:84-124), which leave out atoms
  // carrying fixedvel or fixedpos when any atom does.
  double kinetic(double p[3], bool tagged = false) {
    const long nb = blocks_for(n);
    partials.alloc(std::max<size_t>((size_t)nb * 4, partials.n));
    const int* fx = (tagged && n_tagged > 0) ? fix.p : nullptr;
    k_kinetic<<<nb, kBlock, 0, st>>>(v.p, mass.p, n, partials.p, fx, kFixTagged);
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
  // SetTemperature is tag-aware (tempscaling.cc:70 passes whether any atom
  // carries fixedvel or fixedpos): tagged atoms are neither counted nor
  // scaled.  The Berendsen factor uses the plain Temperature of every atom
  // and scales every atom (berendsen.cc:84-91).
  void scale_to_temperature(double temp, const unsigned int* gate = nullptr, double q = -1.0) {
    const long nb = blocks_for(n);
    const int* fx = (q < 0.0 && n_tagged > 0) ? fix.p : nullptr;
    const long n_eff = fx ? n - n_tagged : n;
    k_kinetic<<<nb, kBlock, 0, st>>>(v.p, mass.p, n, partials.p, fx, kFixTagged);
    k_reduce_partials<<<1, kBlock, 0, st>>>(partials.p, nb, 4, red.p);
    k_set_temperature<<<nb, kBlock, 0, st>>>(v.p, red.p, n, temp, q, gate, gate ? budget() : 0.0, n_eff, fx,
                                             kFixTagged);
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
      // SetTemperature draws the velocities first while none were set
      // (simulationbuilder.cc:54-55)
      if (!velocities_set) {
        draw_velocities(gate);
        vinit_step = k;
      }
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

  // Host state of the velocity modifiers at the start of a chunk.
  struct ModState {
    Berendsen b;
    bool vset;
    Drand48 rng;
  };
  ModState mod_state() const { return {bsched, velocities_set, rng48}; }

  // After a chunk that ran `ran` of its steps: the berendsen state as the
  // steps that ran left it (from the state saved at the chunk start), and
  // the velocities drawn for a step that did not run taken back.
  void replay_modifiers(const ModState& saved, long first_step, long ran, double dt) {
    bsched = saved.b;
    double t;
    for (long k = first_step; k < first_step + ran; ++k)
      bsched.apply(k, dt, t);
    if (!saved.vset && vinit_step >= first_step + ran) {
      velocities_set = false;
      rng48 = saved.rng;
      vinit_step = -1;
    }
  }

  // SimulationEngine::InitVelocities (simulationbuilder.cc:167-185) from the
  // system's drand48 stream, drawn on the host in caller order and put in
  // place on the stream (after any kernel already queued; gate as for the
  // force kernels of a speculative step).
  void draw_velocities(const unsigned int* gate = nullptr) {
    hvinit.assign(3 * n, 0.0);
    init_velocities(n, hmass.data(), rng48, hvinit.data());
    vinit.alloc(3 * n);
    hipchk(hipMemcpyAsync(vinit.p, hvinit.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, st), "memcpy");
    k_gather_velocities<<<blocks_for(n), kBlock, 0, st>>>(v.p, vinit.p, perm.p, n, gate, gate ? budget() : 0.0);
    launch_check();
    velocities_set = true;
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
    A.fix = fix_arg();
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
  void vv(double dt, long steps, long obs_every, double* ep_t, double* ek_t, bool beeman = false,
          bool nose = false) {
    const char* who = nose ? "nosehoover" : beeman ? "beeman" : "velocityverlet";
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
    if (nose && !nh_ready)
      fail(LPMDX_ESTATE, "nosehoover: lpmdx_nosehoover_init was not called");
    if (nose && !(dt == nh_dt))
      fail(LPMDX_EINVAL, "nosehoover: dt differs from the one lpmdx_nosehoover_init used");
    if (n == 0 || steps <= 0)
      return;
    // a look-ahead made by the other scheme is taken back (a stays valid)
    if (drifted && (drift_beeman != beeman || drift_nose != nose))
      undrift();
    if (nose && (!nn.p || !ab.p || !nh.p))
      fail(LPMDX_EDEVICE, "nosehoover: history arrays missing (internal error)");
    const int init_mode = nose ? kNHStart : beeman ? kBeemanInit : kInitDrift;
    const int step_mode = nose ? kNHStep : beeman ? kBeemanStep : kStep;
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
      if (nose) {
        if (!forces_ok)
          forces(init_mode, false, dt, md);
        nh_position(dt, md, nullptr);
      } else if (!forces_ok) {
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
        A.fix = fix_arg();
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
      drift_nose = nose;
      forces_ok = true;
      hipchk(hipMemcpyAsync(&hctrl->maxd2[0], md, sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy");
      sync();
      disp_cur = read_disp(0);
    }
    // Steps run in chunks without a host round trip.  Every step after a
    // chunk's first is speculative: its kernels run only while the positions
    // its forces read passed the skin test (every atom within half the skin
    // of its build position, read from the previous step's displacement slot
    // on the device), and return at once otherwise.  After the chunk the host
    // reads the slots, keeps the steps that ran, and rebuilds the list there;
    // nothing has to be rolled back.
    // A chunk sized to end past the budget (and every step of a manager that
    // rebuilds on every update) has its rebuild enqueued behind it before the
    // host reads anything: the positions of the last step that ran are
    // settled into x on the device, the list is rebuilt there, and one
    // synchronisation serves both.  When the chunk ended inside the budget
    // after all, the list was rebuilt a little early, which is still exact.
    long done = 0;
    bool prebuilt = false; // the list for the next chunk was built ahead
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
      if (prebuilt) {
        need = false;
        prebuilt = false;
      }
      // positions x are where the next force evaluation happens
      drifted = false;
      if (need) {
        rebuild();
        disp_cur = 0.0;
      }
      // the next chunk's list is built ahead when this chunk is expected to
      // end past the budget (or the manager rebuilds on every update), and
      // steps follow it in this call without an observation in between
      const bool ahead = can_rebuild_ahead() && (every || (m == chunk && chunk_by_growth)) && done + m < steps &&
                         (obs_every <= 0 || (done + m) % obs_every != 0);
      // cleared per-step maxima (m <= kMaxChunk)
      k_chunk_begin<<<blocks_for(m), kBlock, 0, st>>>(md, m);
      launch_check();
      const ModState msaved = mod_state();
      for (long s = 0; s < m; ++s) {
        const bool obs = obs_every > 0 && ((done + s + 1) % obs_every) == 0;
        const unsigned int* gate = (every || s == 0) ? nullptr : md + s - 1;
        prof_step = n_steps + s;
        forces(step_mode, obs, dt, md + s, gate);
        if (nose) {
          // the step's modifiers act before the next position update, whose
          // friction reads the temperature of the velocities they leave
          velocity_modifiers(n_steps + s + 1, dt, gate);
          nh_position(dt, md + s, gate);
          std::swap(x.p, xn.p);
        } else {
          std::swap(x.p, xn.p);
          modifiers_after_step(n_steps + s + 1, beeman, dt, md + s, gate);
        }
      }
      prof_step = -1;
      if (ahead) {
        if (m > 1) {
          k_settle_positions<<<blocks_for(3 * n), kBlock, 0, st>>>(md, m, x.p, xn.p, 3 * n);
          launch_check();
        }
        rebuild_ahead();
      }
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
      // (a list built ahead found the positions settled into x already)
      if (!ahead && (m - ran) % 2 == 1)
        std::swap(x.p, xn.p);
      prof_discard_from(n_steps + ran);
      if (ran < m)
        replay_modifiers(msaved, n_steps + 1, ran, dt);
      if (!pending.empty())
        prof_collect();
      if (ran < m)
        n_cut_chunks++;
      disp_cur = d;
      if (ahead) {
        rebuild_settle();
        disp_cur = 0.0;
        prebuilt = true;
      }
      pos_version++;
      // next chunk: double while nothing moves, else the steps the growth
      // rate leaves before the budget (from a fresh list after a cut), plus
      // two so that the chunk usually ends past the limit in one round trip
      long next = std::min<long>(2 * chunk, kMaxChunk);
      chunk_by_growth = false;
      if (growth > 0.0) {
        const double room = (disp_cur > b || prebuilt) ? b : b - disp_cur;
        const long by_growth = (long)(room / growth) + 2;
        if (by_growth < next) {
          next = by_growth;
          chunk_by_growth = true;
        }
      }
      chunk = std::max<long>(1, next);
      drifted = true;
      drift_beeman = beeman;
      drift_nose = nose;
      cm.ucalls += 2 * ran;
      done += ran;
      n_steps += ran;
      if (obs_every > 0 && (done % obs_every) == 0) {
        double e;
        read_energies(&e, nullptr, nullptr);
        // the step's state: v in v, positions in xn
        double ek = kinetic(nullptr, true);
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

  // Nose-Hoover: the position update of the step whose forces a holds (x the
  // positions, v the velocities after the last step): the kinetic sum of v,
  // the step's friction (k_nh_friction), then k_nh_drift writes xn and nn.
  void nh_position(double dt, unsigned int* maxd2, const unsigned int* gate) {
    const long nb = blocks_for(n);
    partials.alloc(std::max<size_t>((size_t)nb * 4, partials.n));
    k_kinetic<<<nb, kBlock, 0, st>>>(v.p, mass.p, n, partials.p);
    k_reduce_partials<<<1, kBlock, 0, st>>>(partials.p, nb, 4, red.p);
    k_nh_friction<<<1, 1, 0, st>>>(red.p, nh.p, n, fixpos_any ? n_moving : n, nh_q, nh_t, dt, nh_kb, gate,
                                   gate ? budget() : 0.0);
    ForceArgs A{};
    A.x = x.p;
    A.xn = xn.p;
    A.v = v.p;
    A.a = a.p;
    A.ab = ab.p;
    A.nn = nn.p;
    A.nh = nh.p;
    A.perm = perm.p;
    A.fix = fix_arg();
    A.maxd2 = maxd2;
    A.gate = gate;
    A.gate_b = gate ? budget() : 0.0;
    A.xb = xb.p;
    A.b = box;
    A.n = n;
    A.dt = dt;
    // a stepping mode: flush_maxd2 records the displacement only outside kForcesOnly
    A.mode = kNHStep;
    for (int q = 0; q < 3; ++q)
      A.ll[q] = ll[q];
    if (fixpos_any && !mv_rank.p)
      fail(LPMDX_EDEVICE, "nosehoover: moving-atom ranks missing (internal error)");
    k_nh_drift<<<nb, kBlock, 0, st>>>(A, fixpos_any ? mv_rank.p : nullptr);
    launch_check();
  }

  // NoseHoover::Initialize (plugins/nosehoover.cc:62-68): the forces at the
  // old configuration of Integrator::GoBack become old.a (in ab), the
  // friction starts at zero; x, v and a are left as they were.
  void nh_init(double dt, double fmass, double t, double kboltzmann) {
    if (pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!std::isfinite(dt) || !std::isfinite(fmass) || !std::isfinite(t) || !std::isfinite(kboltzmann))
      fail(LPMDX_ENONFINITE, "nosehoover: dt, fmass, t and kboltzmann must be finite");
    if (fmass == 0.0)
      fail(LPMDX_EINVAL, "nosehoover: fmass must not be zero");
    undrift();
    nh_ready = false;
    nh_dt = dt;
    nh_q = fmass;
    nh_t = t;
    nh_kb = kboltzmann;
    nh.alloc(2);
    k_fill_double<<<1, 32, 0, st>>>(nh.p, 2, 0.0);
    launch_check();
    if (n == 0) {
      nh_ready = true;
      return;
    }
    old_config_forces(dt, false);
    ab.alloc(3 * n);
    nn.alloc(3 * n);
    hipchk(hipMemcpyAsync(ab.p, a.p, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    restore_after_old_config();
    nh_ready = true;
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
      const ModState msaved = mod_state();
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
        replay_modifiers(msaved, n_steps + 1, ran, dt);
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
        const double ek = kinetic(nullptr, true);
        const long idx = done / obs_every - 1;
        if (ep_t)
          ep_t[idx] = e;
        if (ek_t)
          ek_t[idx] = ek;
      }
    }
  }
};

// The launchers of every kind are instantiated in the engine_kinds_*.hip
// files only (engine_launch.hpp); no other file instantiates them.
#define LPMDX_EXTERN_KIND(K)                                                          \
  extern template void lpmdx_system::launch_kind<K, true>(ForceArgs&, long, double); \
  extern template void lpmdx_system::launch_kind<K, false>(ForceArgs&, long, double);
LPMDX_EXTERN_KIND(kLJ)
LPMDX_EXTERN_KIND(kFastLJ)
LPMDX_EXTERN_KIND(kMorse)
LPMDX_EXTERN_KIND(kBuckingham)
LPMDX_EXTERN_KIND(kSuttonChen)
LPMDX_EXTERN_KIND(kFinnisSinclair)
LPMDX_EXTERN_KIND(kGupta)
LPMDX_EXTERN_KIND(kHarmonic)
LPMDX_EXTERN_KIND(kLJMod)
LPMDX_EXTERN_KIND(kFinnisSinclairExt)

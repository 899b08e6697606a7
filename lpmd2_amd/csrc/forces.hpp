// Pair and EAM force kernels with the fused velocity-Verlet update.
//
// Precision split: the displacement d = x_j - x_i, its minimum image and
// r2 = |d|^2 are formed in FP64 exactly as OrthogonalCell::Displacement and
// SquareModule form them, so cutoff decisions match the reference bit for bit;
// r2 is then cast to FP32 for the pair and embedding arithmetic; each pair's
// FP32 force scalar multiplies the FP64 displacement and the per-atom sums
// accumulate in FP64 (this bounds the summation error below what compensated
// FP32 sums reach, in fewer instructions on CDNA4); accelerations, the position
// and velocity update, rho and the global energy/virial/stress sums are FP64.
//
// Full neighbour lists: every atom accumulates its own force, so there is no
// scatter and no atomics; each pair's energy, virial and stress enter half per
// direction.
#pragma once

#include <type_traits>

#include "cells.hpp"
#include "common.hpp"

namespace lpmdx {

enum PotKind {
  kLJ = 1,
  kFastLJ = 2,
  kMorse = 3,
  kBuckingham = 4,
  kSuttonChen = 5,
  kFinnisSinclair = 6,
  kGupta = 7,
  kHarmonic = 8,
  kLJMod = 9,
  kFinnisSinclairExt = 10
};
constexpr int kNumKinds = 10; // kinds are 1 .. kNumKinds

// the two-pass (embedded-atom) kinds
__host__ __device__ constexpr bool metal_kind(int k) {
  return k == kSuttonChen || k == kFinnisSinclair || k == kGupta || k == kFinnisSinclairExt;
}

struct DevPot {
  int kind;
  int spc_sum;  // Potential::AppliesTo: negative = all pairs
  double rc2;   // pair filter r2 < rc2 in FP64 (list cutoff and own cutoff)
  float f[12];  // kind-specific FP32 parameters (see PotParams in engine.hip)
  int ni, mi;   // integer exponents (Sutton-Chen, lennardjonesMod), 0 if not integral
  double g[6];  // FP64 parameters (finnissinclair-ext: c0 .. c4, c)
  // fastlj table (FP32 copies of the reference's FP64 tables)
  const float* tab_e;
  const float* tab_f;
  double tab_cut; // cutoff, for k = floor(bins * r / cutoff) in FP64
  int bins;
};

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// 2^x on the hardware exponential (v_exp_f32, 1 ulp).  The kernels take e^y
// as 2^(y log2 e) with log2 e folded into the parameters on the host, so an
// exponential is one instruction: the FP32 argument's own rounding (a few
// ulp of its largest term) already bounds the error, which the library
// expf's range reduction and overflow selects (about eleven more
// instructions) would not lower.
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }

// t^k for 0 <= k <= 31 without branches (square-and-multiply over the bits of k)
__device__ __forceinline__ float powu(float t, int k) {
  const float t2 = t * t, t4 = t2 * t2, t8 = t4 * t4, t16 = t8 * t8;
  float r = (k & 1) ? t : 1.0f;
  r *= (k & 2) ? t2 : 1.0f;
  r *= (k & 4) ? t4 : 1.0f;
  r *= (k & 8) ? t8 : 1.0f;
  r *= (k & 16) ? t16 : 1.0f;
  return r;
}

__device__ __forceinline__ float powk(float t, int ki, float kf) {
  return ki > 0 ? powu(t, ki) : powf(t, kf);
}

// t^K for a compile-time 0 < K < 32: the same squares multiplied in the same
// order as powu (by 1.0 first, which is exact), so the result is bit-identical.
template <int K> __device__ __forceinline__ float powc(float t) {
  static_assert(K > 0 && K < 32, "exponent out of range");
  const float t2 = t * t, t4 = t2 * t2, t8 = t4 * t4, t16 = t8 * t8;
  float r = 1.0f;
  if constexpr ((K & 1) != 0)
    r = t;
  if constexpr ((K & 2) != 0)
    r *= t2;
  if constexpr ((K & 4) != 0)
    r *= t4;
  if constexpr ((K & 8) != 0)
    r *= t8;
  if constexpr ((K & 16) != 0)
    r *= t16;
  return r;
}

// Sutton-Chen exponent pair fixed at compile time: EXP = 32 n + m, 0 = read
// them at run time.
constexpr int sc_exp(int n, int m) { return 32 * n + m; }

template <int EXP> __device__ __forceinline__ float sc_pow_n(const DevPot& P, float t) {
  if constexpr (EXP != 0)
    return powc<EXP / 32>(t);
  else
    return powk(t, P.ni, P.f[2]);
}

template <int EXP> __device__ __forceinline__ float sc_pow_m(const DevPot& P, float t) {
  if constexpr (EXP != 0)
    return powc<EXP % 32>(t);
  else
    return powk(t, P.mi, P.f[3]);
}

// Pair potentials: force on i is fscal * rij (rij = x_j - x_i), pair energy e.
// Pair forms follow plugins/{lennardjones,fastlj,morse,buckingham}.cc; the
// FP32 evaluation uses hardware reciprocal / reciprocal square root (1 ulp).
template <int KIND>
__device__ __forceinline__ void pair_eval(const DevPot& P, double r2d, float r2, float& fscal,
                                          float& e) {
  if constexpr (KIND == kLJ) {
    // lennardjones.cc:53-67: E = 4 eps (s12 - s6), ff = -48 eps/r2 (s12 - s6/2)
    const float inv = rcp(r2);
    const float s2 = P.f[0] * inv;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    fscal = -P.f[1] * inv * (s12 - 0.5f * s6);
    e = P.f[2] * (s12 - s6);
  } else if constexpr (KIND == kFastLJ) {
    // fastlj.cc:81-95: table entry floor(bins*r/cutoff), no interpolation
    const double r = sqrt(r2d);
    long k = (long)floor(P.bins * r / P.tab_cut);
    if (k > P.bins - 1)
      k = P.bins - 1;
    fscal = P.tab_f[k];
    e = P.tab_e[k];
  } else if constexpr (KIND == kMorse) {
    // morse.cc:53-62: E = D (1 - e^{-a(r-re)})^2, ff = 2 a D / r (1 - ex) ex
    const float ir = rsq(r2);
    const float r = r2 * ir;
    const float ex = exp2_hw(-P.f[4] * (r - P.f[1]));
    const float om = 1.0f - ex;
    fscal = P.f[3] * ir * om * ex;
    e = P.f[2] * om * om;
  } else if constexpr (KIND == kBuckingham) {
    // buckingham.cc:58-75: E = B1 e^{-r/Ro} - B2/r^6, ff = 6 B2/r^8 - B1 e^{-r/Ro}/(Ro r)
    const float ir = rsq(r2);
    const float r = r2 * ir;
    const float ir2 = ir * ir;
    const float ir6 = ir2 * ir2 * ir2;
    const float ex = P.f[0] * exp2_hw(-r * P.f[5]);
    fscal = P.f[4] * ir6 * ir2 - ex * P.f[3] * ir;
    e = ex - P.f[1] * ir6;
  } else if constexpr (KIND == kHarmonic) {
    // harmonic.cc:50-60: E = k/2 (r - a)^2, ff = k (r - a) / r; f = {k, a, k/2}
    const float ir = rsq(r2);
    const float t = r2 * ir - P.f[1];
    fscal = P.f[0] * t * ir;
    e = P.f[2] * t * t;
  } else if constexpr (KIND == kLJMod) {
    // lennardjonesMod.cc:65-79 with t = sigma^2 / r^2: (sigma/r)^m = t^{m/2},
    // E = 4 eps (t^{m/2} - c t^{n/2}), ff = -4 eps / r^2 (m t^{m/2} - c n t^{n/2});
    // f = {sigma^2, 4 eps, m, c n, c, m/2, n/2}, mi / ni = m/2, n/2 when integral
    const float inv = rcp(r2);
    const float t = P.f[0] * inv;
    const float tm = powk(t, P.mi, P.f[5]);
    const float tn = powk(t, P.ni, P.f[6]);
    fscal = -P.f[1] * inv * (P.f[2] * tm - P.f[3] * tn);
    e = P.f[1] * (tm - P.f[4] * tn);
  }
}

// EAM: rho_ij(r) and, for the force pass, the pair energy V and the scalar
// fscal with force on i = fscal * rij:
//   fscal = (dV/dr + (w_i + w_j) G(r)) / r,  w = rho^{-1/2},
// where (w_i + w_j) G(r) r-hat is the reference's ManyBodies term.
template <int KIND, int EXP = 0> __device__ __forceinline__ float eam_rho(const DevPot& P, float r2) {
  if constexpr (KIND == kSuttonChen) { // suttonchen.cc:78-81: a^m r^-m
    return sc_pow_m<EXP>(P, P.f[1] * rsq(r2));
  } else if constexpr (KIND == kGupta) { // gupta.cc:69
    const float r = r2 * rsq(r2);
    return exp2_hw(P.f[6] - P.f[7] * r);
  } else if constexpr (KIND == kFinnisSinclairExt) { // finnissinclair-ext.cc:93-99
    const float r = r2 * rsq(r2);
    const float d = P.f[6];
    if (r <= d) {
      const float t = r - d;
      return (t * t) * (1.0f + P.f[8] * t * t);
    }
    return 0.0f;
  } else { // finnissinclair.cc:82-87
    const float r = r2 * rsq(r2);
    const float d = P.f[6];
    if (r <= d) {
      const float t = r - d;
      return (t * t) * (1.0f + P.f[8] * t);
    }
    return 0.0f;
  }
}

template <int KIND, int EXP = 0>
__device__ __forceinline__ void eam_pair(const DevPot& P, float r2, float wsum, float& fscal,
                                         float& V) {
  const float ir = rsq(r2);
  if constexpr (KIND == kSuttonChen) {
    // suttonchen.cc:73-95, f = {e, a, n, m, c, -, -, -, -n e, m c e / 2}
    const float t = P.f[1] * ir;
    const float tn = sc_pow_n<EXP>(P, t);
    const float tm = sc_pow_m<EXP>(P, t);
    V = P.f[0] * tn;
    fscal = (ir * ir) * (P.f[8] * tn + P.f[9] * wsum * tm);
  } else if constexpr (KIND == kGupta) {
    // gupta.cc:67-82, f = {A, r0, p, B, q, p L, 2q L, 2q/r0 L, p/r0 L, -A p/r0, B q/r0},
    // L = log2 e
    const float r = r2 * ir;
    const float ep = exp2_hw(P.f[5] - P.f[8] * r);
    const float eq = exp2_hw(P.f[6] - P.f[7] * r);
    V = P.f[0] * ep;
    fscal = ir * (P.f[9] * ep + P.f[10] * wsum * eq);
  } else if constexpr (KIND == kFinnisSinclairExt) {
    // finnissinclair-ext.cc:85-123, f = {c0, c1, c2, A, c3, c, d, c4, B^2, 4 B^2, -A/2, K = A},
    // g = {c0, c1, c2, c3, c4, c}.  The pair polynomial runs in FP64: its
    // terms cancel (Mo: c0 48, c1 r -96 and c2 r^2 46 sum to -0.2 near the
    // first shell), which FP32 would leave good to 1e-4 only.
    const float r = r2 * ir;
    const float c = P.f[5], d = P.f[6];
    float dV = 0.0f, G = 0.0f;
    V = 0.0f;
    if (r <= c) {
      const double rd = (double)r, t = rd - P.g[5], rr = rd * rd;
      const double poly = P.g[0] + P.g[1] * rd + P.g[2] * rr + P.g[3] * rr * rd + P.g[4] * rr * rr;
      V = (float)(t * t * poly);
      dV = (float)(2.0 * t * poly + (P.g[1] + 2.0 * P.g[2] * rd + 3.0 * P.g[3] * rr + 4.0 * P.g[4] * rr * rd) * t * t);
    }
    if (r <= d) {
      const float t = r - d;
      G = P.f[10] * (2.0f * t + P.f[9] * t * t * t);
    }
    fscal = (dV + wsum * G) * ir;
  } else {
    // finnissinclair.cc:75-108, f = {c0, c1, c2, A, B, c, d, -, B/d, 3B/d, -A/2}
    const float r = r2 * ir;
    const float c = P.f[5], d = P.f[6];
    float dV = 0.0f, G = 0.0f;
    V = 0.0f;
    if (r <= c) {
      const float t = r - c;
      const float poly = P.f[0] + P.f[1] * r + P.f[2] * r * r;
      V = t * t * poly;
      dV = 2.0f * t * poly + t * t * (P.f[1] + 2.0f * P.f[2] * r);
    }
    if (r <= d) {
      const float t = r - d;
      G = P.f[10] * (2.0f * t + P.f[9] * t * t);
    }
    fscal = (dV + wsum * G) * ir;
  }
}

// Integration modes of the force kernels.
enum StepMode {
  kForcesOnly = 0, // a = F/m (or a += with accumulate)
  kInitDrift = 1,  // a = F/m, then next positions xn = drift(x, v, a)
  kStep = 2,       // a = F/m, v += dt/2 (a_old + a), xn = drift(x, v, a)
  kLeapfrog = 3,   // a = F/m, vh' = vh + a dt, xn = FittedInside(x + vh' dt), v = (vh' + vh) / 2, vh = vh'
  kBeemanInit = 4, // a = F/m, then xn = Beeman position update with a_old = ab
  kBeemanStep = 5, // a = F/m, Beeman velocity update from (a, old a, ab), ab = old a, xn = position update
  // Nose-Hoover (plugins/nosehoover.cc:66-96): the force kernel ends the step
  // whose positions it sees; the next step's position update follows in
  // k_nh_drift, once the step's friction is known
  kNHStart = 6, // a = F/m
  kNHStep = 7   // a = F/m, v += (a/3 + 5 nn/6 - ab/6) dt, ab = nn
};

// Beeman's position update before FittedInside (beeman.cc:63-76), in its order
// of operations: x + v dt + (2/3) a dt^2 - (1/6) a_old dt^2.
__device__ __forceinline__ double beeman_position(double x, double v, double a, double a_old, double dt) {
  return ((x + v * dt) + ((2.0 / 3.0) * a) * dt * dt) - ((1.0 / 6.0) * a_old) * dt * dt;
}

struct ForceArgs {
  const double* x;   // [3n] current positions
  double* xn;        // [3n] next positions (modes 1, 2)
  double* v;         // [3n]
  double* a;         // [3n] in: a_old (mode 2); out: new accelerations
  const int* z;      // [n] atomic numbers
  const double* ffm; // [n] forcefactor / m
  const int* nbr;
  const int* cnt;
  const float* w_rho;   // EAM: [n] rho^{-1/2}
  // EAM, unstaged force pass: [n] (x, y, z, rho^{-1/2}) of each atom, written
  // by the rho pass, so one neighbour is two 16-byte loads instead of three
  double4* xw;
  double* partials;     // [nblocks * 12] when evaluating energies
  unsigned int* maxd2;  // max over atoms of |x_next - xb|^2 (float bits), or null
  const unsigned int* gate; // the previous step's maxd2 slot (speculative steps), or null
  double gate_b;            // the skin budget that slot is held against
  const double* xb;     // [3n] positions at the last list build
  double* vh;           // [3n] leapfrog half-step velocities (kLeapfrog)
  double* ab;           // [3n] Beeman: the acceleration of the step before a's (kBeeman*);
                        // Nose-Hoover: the friction-modified acceleration of the last position update
  double* nn;           // [3n] Nose-Hoover: that of the position update still to be ended
  const double* nh;     // Nose-Hoover: {friction before this step's increments, increment per atom}
  const int* perm;      // caller index of each slot (Nose-Hoover: the order of the increments)
  // Largest separation a listed pair can have now (list radius + padding +
  // twice the displacement allowed since the build), or 0.  An atom at least
  // this far from every face has no listed neighbour across the boundary.
  double margin;
  // L - margin per axis (the interior's upper bounds), set with margin: kernel
  // arguments, so the interior test holds no per-thread copies
  double margin_hi[3];
  // per-atom fixed-atom flags (kFixVel, kFixPos; internal order) of a
  // system where some atom carries the fixedvel or fixedpos tag, else null
  const int* fix;
  // per-atom energies (Potential::AtomEnergy) of an evaluating launch, or
  // null: the pair energies over the atom's full list, plus F(rho_i) for EAM
  double* eatom;
  DevPot P;
  const DevPot* terms; // several potentials of one kind on one list (species pairs), or null
  int one_term;        // with terms: each species pair meets at most one of them
  int nterm;
  Box b;
  double ll[3]; // |cell vector q| as FittedInside reads it
  double dt;
  long n;
  long stride;
  int cap; // list entries per atom (a multiple of the lanes per atom)
  int mode;
  int accumulate; // add to a instead of overwriting (several potentials)
};

// |p - xb_i|^2 through the minimum image: how far atom i has moved since the
// list was built (the list stays valid while every atom is within half the
// skin of its build position).
__device__ __forceinline__ float build_displacement2(const ForceArgs& A, long i, double px, double py,
                                                     double pz) {
  if (!A.xb)
    return 0.0f;
  const double* q = A.xb + 3 * i;
  double dx = px - q[0], dy = py - q[1], dz = pz - q[2];
  min_image(A.b, dx, dy, dz);
  return (float)sqmod(dx, dy, dz);
}

// Fixed atoms of an integrator step (TwoStepIntegrator::Advance,
// twostepintegrator.cc:12-68; OneStepIntegrator::Advance, onestepintegrator.cc:13-46):
// a fixedvel atom's acceleration is set to zero after every force evaluation
// of a step, a fixedpos atom whose tag reads "true" is neither moved nor
// has its velocity (or any history the scheme keeps) advanced.
constexpr int kFixVel = 1;
constexpr int kFixPos = 2;
// the atom carries fixedvel or fixedpos with any value: the tag-aware
// kinetic energy, temperature and momentum (properties.h:22-81) and
// SetTemperature (simulationbuilder.cc:53-71) leave it out
constexpr int kFixTagged = 4;

// Velocity Verlet tail for atom i given its new acceleration (FP64):
// plugins/velocityverlet.cc:60-75, in the reference's order of operations.
__device__ __forceinline__ void vv_tail(const ForceArgs& A, long i, double ax, double ay,
                                        double az, float& d2) {
  double* a = A.a + 3 * i;
  d2 = 0.0f;
  if (A.mode != kForcesOnly && A.fix) {
    const int f = A.fix[i];
    if (f & kFixVel)
      ax = ay = az = 0.0;
    if (f & kFixPos) {
      // UpdateForces still sets a; nothing else of the atom changes
      a[0] = ax;
      a[1] = ay;
      a[2] = az;
      const double* x = A.x + 3 * i;
      const double p[3] = {x[0], x[1], x[2]};
      d2 = build_displacement2(A, i, p[0], p[1], p[2]);
      double* xn = A.xn + 3 * i;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        xn[q] = p[q];
      return;
    }
  }
  if (A.mode == kForcesOnly) {
    if (A.accumulate) {
      a[0] += ax;
      a[1] += ay;
      a[2] += az;
    } else {
      a[0] = ax;
      a[1] = ay;
      a[2] = az;
    }
    return;
  }
  double* v = A.v + 3 * i;
  const double* x = A.x + 3 * i;
  if (A.mode == kLeapfrog) {
    // Leapfrog::Advance (plugins/leapfrog.cc:69-77): vhalf = old v + a dt,
    // x = FittedInside(x + vhalf dt), v = 0.5 (vhalf + old v), old v = vhalf
    double* vh = A.vh + 3 * i;
    const double dt = A.dt;
    const double am[3] = {ax, ay, az};
    // every load before the first store (the stores could alias them, which
    // would serialise each load behind the previous component's stores)
    const double vo[3] = {vh[0], vh[1], vh[2]};
    const double xo[3] = {x[0], x[1], x[2]};
    double p[3], vn[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      vn[q] = vo[q] + am[q] * dt;
      p[q] = xo[q] + vn[q] * dt;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[q] = 0.5 * (vn[q] + vo[q]);
      vh[q] = vn[q];
      a[q] = am[q];
    }
    d2 = build_displacement2(A, i, p[0], p[1], p[2]);
    double* xn = A.xn + 3 * i;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      xn[q] = fitted_inside(p[q], A.ll[q]);
    return;
  }
  if (A.mode == kNHStart || A.mode == kNHStep) {
    // Nose-Hoover: NoseHoover::AdvanceVelocity (nosehoover.cc:90-96) of the
    // step whose positions these forces see, with old.a = nn (the
    // friction-modified acceleration of its position update) and aux = ab;
    // the history then moves on (ab = nn).  The start only sets a.
    const double an[3] = {ax, ay, az};
    if (A.mode == kNHStep && !(A.fix && (A.fix[i] & kFixPos))) {
      double* ab = A.ab + 3 * i;
      const double* nn = A.nn + 3 * i;
      const double dt = A.dt;
      const double vo[3] = {v[0], v[1], v[2]};
      const double no[3] = {nn[0], nn[1], nn[2]};
      const double ap[3] = {ab[0], ab[1], ab[2]};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        v[q] = ((vo[q] + ((1.0 / 3.0) * an[q]) * dt) + ((5.0 / 6.0) * no[q]) * dt) - ((1.0 / 6.0) * ap[q]) * dt;
        ab[q] = no[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
      a[q] = an[q];
    return;
  }
  if (A.mode == kBeemanInit || A.mode == kBeemanStep) {
    // TwoStepIntegrator with Beeman (beeman.cc:63-84).  kBeemanStep ends the
    // step whose positions are x (AdvanceVelocity with the new a, the old a
    // and ab), shifts the history (ab = old a) and begins the next step
    // (AdvancePosition); kBeemanInit only begins one, with a_old = ab.
    double* ab = A.ab + 3 * i;
    const double dt = A.dt;
    const bool step = A.mode == kBeemanStep;
    const double an[3] = {ax, ay, az};
    const double ao[3] = {a[0], a[1], a[2]};
    const double ap[3] = {ab[0], ab[1], ab[2]};
    const double vo[3] = {v[0], v[1], v[2]};
    const double xo[3] = {x[0], x[1], x[2]};
    double vn[3], p[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      vn[q] = step ? ((vo[q] + ((1.0 / 3.0) * an[q]) * dt) + ((5.0 / 6.0) * ao[q]) * dt) - ((1.0 / 6.0) * ap[q]) * dt
                   : vo[q];
      p[q] = beeman_position(xo[q], vn[q], an[q], step ? ao[q] : ap[q], dt);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (step) {
        v[q] = vn[q];
        ab[q] = ao[q];
      }
      a[q] = an[q];
    }
    d2 = build_displacement2(A, i, p[0], p[1], p[2]);
    double* xn = A.xn + 3 * i;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      xn[q] = fitted_inside(p[q], A.ll[q]);
    return;
  }
  double vx = v[0], vy = v[1], vz = v[2];
  if (A.mode == kStep) {
    const double h = 0.5 * A.dt;
    vx = vx + (a[0] + ax) * h;
    vy = vy + (a[1] + ay) * h;
    vz = vz + (a[2] + az) * h;
    v[0] = vx;
    v[1] = vy;
    v[2] = vz;
  }
  a[0] = ax;
  a[1] = ay;
  a[2] = az;
  const double dt = A.dt;
  double px = (x[0] + vx * dt) + ax * 0.5 * dt * dt;
  double py = (x[1] + vy * dt) + ay * 0.5 * dt * dt;
  double pz = (x[2] + vz * dt) + az * 0.5 * dt * dt;
  d2 = build_displacement2(A, i, px, py, pz);
  double* xn = A.xn + 3 * i;
  xn[0] = fitted_inside(px, A.ll[0]);
  xn[1] = fitted_inside(py, A.ll[1]);
  xn[2] = fitted_inside(pz, A.ll[2]);
}

// Speculative steps.  A chunk of steps is launched without waiting for the
// skin test; a step's kernels run only while the positions its forces read
// stay within the list's budget, i.e. the previous step's largest
// displacement d2 (the gate slot) has sqrt(d2) <= gate_b, the test the host
// makes.  Otherwise they return at once and close their own step's slot, so
// every later step of the chunk returns too and the state stays that of the
// last valid step.  Uniform over the grid (every thread reads the same slot),
// so the return comes before any barrier.
constexpr unsigned int kGateClosed = 0xFFFFFFFFu;

__device__ __forceinline__ bool gate_shut(const ForceArgs& A) {
  if (!A.gate)
    return false;
  const unsigned int g = *A.gate;
  if (g != kGateClosed && !(sqrt((double)__uint_as_float(g)) > A.gate_b))
    return false;
  if (blockIdx.x == 0 && threadIdx.x == 0 && A.maxd2)
    atomicMax(A.maxd2, kGateClosed);
  return true;
}

__device__ __forceinline__ void flush_maxd2(const ForceArgs& A, float d2) {
  float m = wave_max(d2);
  if ((threadIdx.x & (kWave - 1)) == 0 && A.mode != kForcesOnly && m > 0.0f)
    atomic_max_sparse(A.maxd2, __float_as_uint(m)); // non-negative floats order as their bits
}

// Energy/virial/stress accumulators of one lane (FP64), half per direction.
struct Tally {
  double e = 0.0, w = 0.0, s[6] = {0, 0, 0, 0, 0, 0};
  __device__ __forceinline__ void add(float fscal, float dx, float dy, float dz, float r2, float epair) {
    // stress[a][b] += -rij[a] * ff[b], ff = fscal * rij (pairpotential.cc:67-73)
    const double f = 0.5 * (double)fscal;
    e += 0.5 * (double)epair;
    w -= f * (double)r2;
    s[0] -= f * (double)(dx * dx);
    s[1] -= f * (double)(dx * dy);
    s[2] -= f * (double)(dx * dz);
    s[3] -= f * (double)(dy * dy);
    s[4] -= f * (double)(dy * dz);
    s[5] -= f * (double)(dz * dz);
  }
};

template <int LPA> __device__ __forceinline__ float group_sum_f(float v) {
#pragma unroll
  for (int o = LPA / 2; o > 0; o >>= 1)
    v += __shfl_xor(v, o, kWave);
  return v;
}
template <int LPA> __device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = LPA / 2; o > 0; o >>= 1)
    v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ void write_partials(const ForceArgs& A, Tally& t, double extra) {
  __shared__ double sh[9 * kMaxWaves]; // up to 1024-thread workgroups
  double v[9] = {t.e, extra, t.w, t.s[0], t.s[1], t.s[2], t.s[3], t.s[4], t.s[5]};
  block_sum<9>(v, sh);
  if (threadIdx.x == 0) {
    double* o = A.partials + (long)blockIdx.x * 12;
    // E_pair, E_embed, W, S (full 3x3, symmetric)
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
    o[3] = v[3];
    o[4] = v[4];
    o[5] = v[5];
    o[6] = v[4];
    o[7] = v[6];
    o[8] = v[7];
    o[9] = v[5];
    o[10] = v[7];
    o[11] = v[8];
  }
}

// Neighbour loop shared by the force kernels: lane l of atom i visits the
// 4-entry word groups l, l + LPA, ... of its list, issuing the index and
// position loads of a group before any arithmetic, so the gathers overlap.
// body(j, dx, dy, dz, r2) sees the FP64 minimum-image displacement.
constexpr int kUnroll = 4;

// LDS copy of a brick's halo (slots as in the brick's lists): positions
// interleaved, P[3 slot + q], so a neighbour's three coordinates are adjacent
// 8-byte words at one address; W and S are null unless the kernel reads
// neighbour rho^{-1/2} or species.
struct Stage {
  const double* P;
  const float* W;
  const int* S;
};

// The list entries of one batch: kUnroll consecutive entries k0 .. k0+3 of
// this lane (k0 a multiple of 4, so they are one aligned 16-byte word group
// of a chunk), as loaded.
// Batches of list entries in flight ahead of the one being evaluated (one
// batch ahead left list loads exposed for the unstaged EAM kernels at one
// lane per atom; three or four measured no better than two).
constexpr int kListAhead = 2;

template <int LPA> __device__ __forceinline__ int4 list_batch(const int* __restrict__ nb, int k0) {
  static_assert(kUnroll == 4 && kChunk % kUnroll == 0, "one int4 per batch");
  constexpr int T = kWave / LPA;
  return *reinterpret_cast<const int4*>(nb + (long)(k0 / kChunk) * (T * kChunk) + (k0 % kChunk));
}

// Minimum image of the neighbour loop: the rint form on the axes of a bit
// mask (kWrapX | kWrapY | kWrapZ; none: no listed pair crosses a face), for
// displacements clear of half an edge, or the reference's compare-and-select
// form on every axis.
enum WrapMode { kWrapNone = 0, kWrapX = 1, kWrapY = 2, kWrapZ = 4, kWrapNear = 7, kWrapSelect = 8 };

template <int LPA, int WRAP, bool STAGED, bool NEED_S, bool NEED_W, bool PREFETCH, bool PACKED,
          typename Body>
__device__ __forceinline__ void neighbour_loop(const ForceArgs& A, const Stage& H, long i, int l, double xi,
                                               double yi, double zi, Body& body) {
  // a list that outgrew its capacity (reported as an error) is read no further
  // than the capacity, so no entry outside the atom's chunks is ever used
  const int c = min(A.cnt[i], A.cap);
  constexpr int T = kWave / LPA;
  // lane l takes the 4-entry word groups l, l + LPA, l + 2 LPA, ... of the list
  constexpr int step = kUnroll * LPA;
  // entry k of this atom: chunk k/8 of the tile row, word k%8 (see list_index)
  const int* __restrict__ nb = A.nbr + (i / T) * (long)T * A.cap + (i % T) * kChunk;
  // entries past the list end read a valid place (the atom itself, or halo
  // slot 0) and are masked by `valid`
  const int pad = STAGED ? 0 : (int)i;
  // PREFETCH: the next batch's entries are loaded one batch ahead and read
  // only when that batch starts, so the list load is in flight while this
  // batch runs.  Otherwise the batch's entries and then its position gathers
  // are loaded at its start (measured faster for the unstaged pair kernels,
  // where the gathers, not the list, hold a batch up and the arithmetic per
  // batch is too short to cover a list load).
  int4 qn[kListAhead];
#pragma unroll
  for (int d = 0; d < kListAhead; ++d) {
    qn[d] = make_int4(pad, pad, pad, pad);
    if (PREFETCH && kUnroll * l + d * step < c)
      qn[d] = list_batch<LPA>(nb, kUnroll * l + d * step);
  }
  auto next = [&](int k0) {
    int4 q;
    if constexpr (PREFETCH) {
      q = qn[0];
#pragma unroll
      for (int d = 0; d + 1 < kListAhead; ++d)
        qn[d] = qn[d + 1];
      if (k0 + kListAhead * step < c)
        qn[kListAhead - 1] = list_batch<LPA>(nb, k0 + kListAhead * step);
    } else {
      q = list_batch<LPA>(nb, k0);
    }
    return q;
  };
  // one batch; FULL: all kUnroll entries are in the list (no masks)
  auto batch = [&](int k0, const int4& q, auto full) {
    constexpr bool FULL = decltype(full)::value;
    const int js[kUnroll] = {FULL || k0 < c ? q.x : pad, FULL || k0 + 1 < c ? q.y : pad,
                             FULL || k0 + 2 < c ? q.z : pad, FULL || k0 + 3 < c ? q.w : pad};
    double px[kUnroll], py[kUnroll], pz[kUnroll];
    float pw[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if constexpr (STAGED) {
        // slot j at byte 24 j: one 24-bit multiply-add for the address
        const double* p = reinterpret_cast<const double*>(reinterpret_cast<const char*>(H.P) +
                                                          __mul24(js[u], 3 * (int)sizeof(double)));
        px[u] = p[0];
        py[u] = p[1];
        pz[u] = p[2];
      } else if constexpr (PACKED) {
        const double4 p = A.xw[js[u]];
        px[u] = p.x;
        py[u] = p.y;
        pz[u] = p.z;
        pw[u] = (float)p.w;
      } else {
        const double* p = A.x + 3 * (long)js[u];
        px[u] = p[0];
        py[u] = p[1];
        pz[u] = p[2];
      }
    }
    // the batch's gathers go out together, before any arithmetic uses them
    // (left to itself the scheduler delays the last one behind the first
    // three, which exposes its latency)
    if constexpr (!STAGED)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int j = js[u];
      int sj = 0;
      float wj = 0.0f;
      if constexpr (NEED_S)
        sj = STAGED ? H.S[j] : A.z[j];
      if constexpr (NEED_W)
        wj = STAGED ? H.W[j] : (PACKED ? pw[u] : A.w_rho[j]);
      double dx = px[u] - xi, dy = py[u] - yi, dz = pz[u] - zi;
      if constexpr (WRAP == kWrapSelect) {
        min_image(A.b, dx, dy, dz);
      } else {
        if constexpr ((WRAP & kWrapX) != 0)
          dx = wrap1_near(dx, A.b.L[0], A.b.inv[0]);
        if constexpr ((WRAP & kWrapY) != 0)
          dy = wrap1_near(dy, A.b.L[1], A.b.inv[1]);
        if constexpr ((WRAP & kWrapZ) != 0)
          dz = wrap1_near(dz, A.b.L[2], A.b.inv[2]);
      }
      body(FULL || k0 + u < c, sj, wj, dx, dy, dz, sqmod(dx, dy, dz));
    }
  };
  if constexpr (STAGED) {
    // whole batches, then this lane's partly filled last batch, if any
    // (measured 2 % faster for the staged LJ kernel)
    int k0 = kUnroll * l;
    for (; k0 + kUnroll <= c; k0 += step)
      batch(k0, next(k0), std::true_type{});
    if (k0 < c)
      batch(k0, next(k0), std::false_type{});
  } else {
    for (int k0 = kUnroll * l; k0 < c; k0 += step)
      batch(k0, next(k0), std::false_type{});
  }
}

// Runs body(valid, s_j, w_j, dx, dy, dz, r2) over atom i's list, with
// neighbour species s_j (NEED_S) and rho^{-1/2} w_j (NEED_W).  A.margin > 0
// bounds every listed pair's separation below half an edge: where no atom of
// the wavefront is within it of a face, no displacement can reach half an
// edge and the minimum image is the identity (OrthogonalCell::Displacement
// would not wrap).  The same holds axis by axis: the staged kernels, which
// run only with a margin, take the rint wrap (every raw displacement lies
// within the margin of 0 or +-L, where it gives the reference's result;
// measured 16 % faster for the staged LJ kernel, whose VGPR count it also
// lowers) on the one axis whose faces some atom of the wavefront is near, or
// on all three; the others take the compare-and-select wrap on every axis
// (measured faster there: 7 % for the unstaged EAM rho pass).  Each branch is
// uniform over the wavefront.
template <int LPA, bool STAGED, bool NEED_S, bool NEED_W, bool PREFETCH, bool PACKED = false, typename Body>
__device__ __forceinline__ void for_neighbours(const ForceArgs& A, const Stage& H, long i, int l, double xi,
                                               double yi, double zi, Body&& body) {
  const double m = A.margin;
  const bool in_x = m > 0.0 && xi >= m && xi <= A.margin_hi[0];
  const bool in_y = m > 0.0 && yi >= m && yi <= A.margin_hi[1];
  const bool in_z = m > 0.0 && zi >= m && zi <= A.margin_hi[2];
  if constexpr (STAGED) {
    const int near = (__all(in_x) ? 0 : kWrapX) | (__all(in_y) ? 0 : kWrapY) | (__all(in_z) ? 0 : kWrapZ);
    if (near == kWrapNone)
      neighbour_loop<LPA, kWrapNone, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
    else if (near == kWrapX)
      neighbour_loop<LPA, kWrapX, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
    else if (near == kWrapY)
      neighbour_loop<LPA, kWrapY, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
    else if (near == kWrapZ)
      neighbour_loop<LPA, kWrapZ, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
    else
      neighbour_loop<LPA, kWrapNear, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
  } else {
    if (__all(in_x && in_y && in_z))
      neighbour_loop<LPA, kWrapNone, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
    else
      neighbour_loop<LPA, kWrapSelect, STAGED, NEED_S, NEED_W, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi, body);
  }
}

// Per-atom force: FP32 pair scalar times the FP64 displacement, summed in FP64
// (a product of an FP32 and an FP64 value plus an FP64 add per component).
struct ForceSum {
  double x = 0.0, y = 0.0, z = 0.0;
  __device__ __forceinline__ void add(float fs, double dx, double dy, double dz) {
    const double f = (double)fs;
    x = fma(f, dx, x);
    y = fma(f, dy, y);
    z = fma(f, dz, z);
  }
};

// Sum of the pair terms of one kind that apply to a pair (CombinedPotential
// adds every potential whose species rule matches, combinedpotential.h:48-49).
// SPC = false: one potential for every pair (no species rule), branch free.
template <int KIND, bool SPC>
__device__ __forceinline__ void pair_terms(const ForceArgs& A, bool valid, int si, int sj, double r2,
                                           float& fs, float& e) {
  if (!SPC) {
    const bool in = valid && r2 < A.P.rc2;
    float f, en;
    pair_eval<KIND>(A.P, r2, valid ? (float)r2 : 1.0f, f, en);
    fs = in ? f : 0.0f;
    e = in ? en : 0.0f;
  } else if (A.nterm <= 1) {
    const bool in = valid && r2 < A.P.rc2 && (A.P.spc_sum < 0 || si + sj == A.P.spc_sum);
    float f, en;
    pair_eval<KIND>(A.P, r2, valid ? (float)r2 : 1.0f, f, en);
    fs = in ? f : 0.0f;
    e = in ? en : 0.0f;
  } else if (A.one_term) {
    // at most one term per pair: its parameters selected per lane, one
    // evaluation.  Every field pair_eval reads is selected (the FP32
    // parameters, the integer exponents of lennardjonesMod, the table of
    // fastlj / tabulatedpair); the selects of fields a kind never reads are
    // dead code for that kind.
    const int ss = si + sj;
    DevPot Q = A.terms[0];
    bool any = Q.spc_sum < 0 || ss == Q.spc_sum;
    for (int q = 1; q < A.nterm; ++q) {
      const DevPot& T = A.terms[q];
      const bool m = T.spc_sum < 0 || ss == T.spc_sum;
#pragma unroll
      for (int k = 0; k < 12; ++k)
        Q.f[k] = m ? T.f[k] : Q.f[k];
      Q.rc2 = m ? T.rc2 : Q.rc2;
      Q.ni = m ? T.ni : Q.ni;
      Q.mi = m ? T.mi : Q.mi;
      Q.tab_e = m ? T.tab_e : Q.tab_e;
      Q.tab_f = m ? T.tab_f : Q.tab_f;
      Q.tab_cut = m ? T.tab_cut : Q.tab_cut;
      Q.bins = m ? T.bins : Q.bins;
      any = any || m;
    }
    const bool in = valid && any && r2 < Q.rc2;
    float f, en;
    pair_eval<KIND>(Q, r2, valid ? (float)r2 : 1.0f, f, en);
    fs = in ? f : 0.0f;
    e = in ? en : 0.0f;
  } else {
    fs = 0.0f;
    e = 0.0f;
    for (int q = 0; q < A.nterm; ++q) {
      const DevPot& P = A.terms[q];
      const bool in = valid && r2 < P.rc2 && (P.spc_sum < 0 || si + sj == P.spc_sum);
      float f, en;
      pair_eval<KIND>(P, r2, valid ? (float)r2 : 1.0f, f, en);
      fs += in ? f : 0.0f;
      e += in ? en : 0.0f;
    }
  }
}

// Per-atom bodies of the force kernels; lane l of the LPA lanes serving atom i.
// Every lane of a wavefront must call them (they end with wavefront reductions).
// Energies, virial and stress (EVAL) add into the caller's tal.  SPC: the
// potentials read species (a species rule, or several potentials of the kind
// on one list); without it the neighbour loop has no branch.
template <int KIND, int LPA, bool EVAL, bool STAGED, bool SPC>
__device__ __forceinline__ void pair_atom(const ForceArgs& A, const Stage& H, long i, int l, bool active,
                                          Tally& tal) {
  ForceSum F;
  double ea = 0.0; // the atom's pair energies (PairPotential::AtomEnergy)
  if (active) {
    const double xi = A.x[3 * i], yi = A.x[3 * i + 1], zi = A.x[3 * i + 2];
    const int si = SPC ? A.z[i] : 0;
    for_neighbours<LPA, STAGED, SPC, false, STAGED>(A, H, i, l, xi, yi, zi,
                                            [&](bool valid, int sj, float, double dx, double dy, double dz,
                                                double r2) {
      float fs, e;
      pair_terms<KIND, SPC>(A, valid, si, sj, r2, fs, e);
      F.add(fs, dx, dy, dz);
      if constexpr (EVAL) {
        tal.add(fs, (float)dx, (float)dy, (float)dz, (float)r2, e);
        ea += (double)e;
      }
    });
  }
  if constexpr (EVAL)
    if (A.eatom) {
      ea = group_sum<LPA>(ea);
      if (active && l == 0)
        A.eatom[i] = ea;
    }
  const double Fx = group_sum<LPA>(F.x), Fy = group_sum<LPA>(F.y), Fz = group_sum<LPA>(F.z);
  float d2 = 0.0f;
  if (active && l == 0) {
    const double m = A.ffm[i];
    vv_tail(A, i, Fx * m, Fy * m, Fz * m, d2);
  }
  flush_maxd2(A, d2);
}

// EAM pass 1: rho_i over the full list within the list cutoff (no species
// check, metalpotential.cc:121-132), accumulated in FP64; writes
// w_i = rho_i^{-1/2} (FP32) and rho_i, and, when evaluating, the embedding
// energy F(rho_i) = -K sqrt(rho_i) into the caller's emb.
// With A.xw set, also writes the atom's packed (x, y, z, w) record.
template <int KIND, int LPA, bool EVAL, bool STAGED, bool PREFETCH = true, int EXP = 0>
__device__ __forceinline__ void eam_rho_atom(const ForceArgs& A, const Stage& H, long i, int l, bool active,
                                             float* w_out, double* rho_out, double& emb) {
  double rho = 0.0;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (active) {
    xi = A.x[3 * i];
    yi = A.x[3 * i + 1];
    zi = A.x[3 * i + 2];
    const double rc2 = A.P.rc2;
    for_neighbours<LPA, STAGED, false, false, PREFETCH>(A, H, i, l, xi, yi, zi,
                                              [&](bool valid, int, float, double, double, double, double r2) {
      const bool in = valid && r2 < rc2;
      const float v = eam_rho<KIND, EXP>(A.P, valid ? (float)r2 : 1.0f);
      rho += in ? (double)v : 0.0;
    });
  }
  rho = group_sum<LPA>(rho);
  if (active && l == 0) {
    const float w = (float)(1.0 / sqrt(rho));
    w_out[i] = w;
    if (A.xw)
      A.xw[i] = make_double4(xi, yi, zi, (double)w);
    if (rho_out)
      rho_out[i] = rho;
    if constexpr (EVAL) {
      const double fe = -(double)A.P.f[11] * sqrt(rho);
      emb += fe;
      if (A.eatom)
        A.eatom[i] = fe; // MetalPotential::AtomEnergy's F(rho_i); the force pass adds the pairs
    }
  }
}

// EAM pass 2: pair + many-body forces over the full list with AppliesTo
// (metalpotential.cc:140-169), fused with the integrator tail.
// PACKED: neighbour positions and rho^{-1/2} from A.xw.
template <int KIND, int LPA, bool EVAL, bool STAGED, bool SPC, bool PREFETCH = true, int EXP = 0,
          bool PACKED = false>
__device__ __forceinline__ void eam_force_atom(const ForceArgs& A, const Stage& H, long i, int l, bool active,
                                               Tally& tal) {
  ForceSum F;
  double ea = 0.0; // the atom's pair energies where the species rule applies
  if (active) {
    const double xi = A.x[3 * i], yi = A.x[3 * i + 1], zi = A.x[3 * i + 2];
    const int si = SPC ? A.z[i] : 0;
    const float wi = A.w_rho[i];
    const double rc2 = A.P.rc2;
    for_neighbours<LPA, STAGED, SPC, true, PREFETCH, PACKED>(A, H, i, l, xi, yi, zi,
                                           [&](bool valid, int sj, float wj, double dx, double dy, double dz,
                                               double r2) {
      const bool in = valid && r2 < rc2 && (!SPC || A.P.spc_sum < 0 || si + sj == A.P.spc_sum);
      const float fr2 = valid ? (float)r2 : 1.0f;
      float fs, V;
      eam_pair<KIND, EXP>(A.P, fr2, wi + wj, fs, V);
      fs = in ? fs : 0.0f;
      F.add(fs, dx, dy, dz);
      if constexpr (EVAL) {
        tal.add(fs, (float)dx, (float)dy, (float)dz, fr2, in ? V : 0.0f);
        ea += in ? (double)V : 0.0;
      }
    });
  }
  if constexpr (EVAL)
    if (A.eatom) {
      ea = group_sum<LPA>(ea);
      if (active && l == 0)
        A.eatom[i] += ea;
    }
  const double Fx = group_sum<LPA>(F.x), Fy = group_sum<LPA>(F.y), Fz = group_sum<LPA>(F.z);
  float d2 = 0.0f;
  if (active && l == 0) {
    const double m = A.ffm[i];
    vv_tail(A, i, Fx * m, Fy * m, Fz * m, d2);
  }
  flush_maxd2(A, d2);
}

template <int KIND, int LPA, bool EVAL, bool SPC>
static __global__ __launch_bounds__(kBlock) void k_pair_force(ForceArgs A) {
  if (gate_shut(A))
    return;
  const long t = xcd_block() * kBlock + threadIdx.x;
  Tally tal;
  pair_atom<KIND, LPA, EVAL, false, SPC>(A, Stage{}, t / LPA, (int)(t % LPA), t / LPA < A.n, tal);
  if constexpr (EVAL)
    write_partials(A, tal, 0.0);
}

template <int KIND, int LPA, bool EVAL, int EXP = 0>
static __global__ __launch_bounds__(kBlock) void k_eam_rho(ForceArgs A, float* w_out, double* rho_out) {
  if (gate_shut(A))
    return;
  const long t = xcd_block() * kBlock + threadIdx.x;
  double emb = 0.0;
  eam_rho_atom<KIND, LPA, EVAL, false, true, EXP>(A, Stage{}, t / LPA, (int)(t % LPA), t / LPA < A.n, w_out,
                                                  rho_out, emb);
  if constexpr (EVAL) {
    Tally tal;
    write_partials(A, tal, emb);
  }
}

template <int KIND, int LPA, bool EVAL, bool SPC, int EXP = 0>
static __global__ __launch_bounds__(kBlock) void k_eam_force(ForceArgs A) {
  if (gate_shut(A))
    return;
  const long t = xcd_block() * kBlock + threadIdx.x;
  Tally tal;
  eam_force_atom<KIND, LPA, EVAL, false, SPC, true, EXP, true>(A, Stage{}, t / LPA, (int)(t % LPA), t / LPA < A.n,
                                                               tal);
  if constexpr (EVAL)
    write_partials(A, tal, 0.0);
}

// ---- brick-staged kernels: one workgroup per brick of a bricked grid.  The
// workgroup copies its brick's halo into LDS (interleaved FP64 positions,
// plus rho^{-1/2} and species when the kernel reads them: dynamic LDS of
// cap * (24 + 4 [W] + 4 [S]) bytes, cap >= the largest halo of the launch),
// then its threads take the brick's atoms, LPA lanes each, blockDim / LPA
// atoms per pass; the neighbour lists hold halo slots.
enum StageParts { kStageW = 1, kStageS = 2 };
// dynamic LDS of one staged workgroup: 160 KiB less the kernels' static LDS
constexpr int kStageLdsLimit = 160 * 1024 - 4096;

struct StageLds {
  double* P; // 3 * cap
  float* W;
  int* S;
};

__device__ __forceinline__ StageLds stage_lds(int cap, int parts) {
  extern __shared__ double lds_dyn[];
  StageLds L;
  L.P = lds_dyn;
  char* rest = reinterpret_cast<char*>(L.P + 3 * (size_t)cap);
  L.W = (parts & kStageW) ? reinterpret_cast<float*>(rest) : nullptr;
  if (parts & kStageW)
    rest += sizeof(float) * (size_t)cap;
  L.S = (parts & kStageS) ? reinterpret_cast<int*>(rest) : nullptr;
  return L;
}

__host__ __device__ inline size_t stage_bytes(int cap, int parts) {
  return (size_t)cap * (3 * sizeof(double) + ((parts & kStageW) ? sizeof(float) : 0) +
                        ((parts & kStageS) ? sizeof(int) : 0));
}

// Copies the halo of one brick (its row of the k_brick_halo table; bricks of
// `brick` cells per edge) into L and returns the brick's first sorted atom and
// atom count.
// Every thread takes slots threadIdx.x + u * blockDim.x, finds each slot's
// halo cell by binary search over the cells' first slots, and issues the
// loads of kStageBatch slots before storing any.  Slots past the halo's end
// load the last slot's atom (no branch around a load, which would make the
// compiler wait for each load before the next) and store nothing.  The slots
// in flight hold about 9 registers each: 4 of them keep the staging inside
// the register budget of the force kernels' occupancy (8 spilled there;
// measured with 4: LJ 4M force pass 597 -> 558 us, 2 no better, 3 and 6
// slower).
constexpr int kStageBatch = 4;

__device__ __forceinline__ void stage_brick(const ForceArgs& A, const int2* __restrict__ row, int brick,
                                            const StageLds& L, const float* __restrict__ wsrc, int& a0,
                                            int& na) {
  __shared__ int2 sr[kMaxHaloCells + 1];
  const int E = brick + 2, H = E * E * E;
  for (int h = threadIdx.x; h <= H; h += blockDim.x)
    sr[h] = row[h];
  __syncthreads();
  const int total = sr[H].x;
  for (int base = threadIdx.x; base < total; base += kStageBatch * blockDim.x) {
    int k[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int slot = min(base + u * (int)blockDim.x, total - 1);
      // the last halo cell whose first slot is <= slot (sr[0].x == 0)
      int lo = 0, hi = H - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sr[mid].x <= slot)
          lo = mid;
        else
          hi = mid - 1;
      }
      k[u] = sr[lo].y + (slot - sr[lo].x);
    }
    double vx[kStageBatch], vy[kStageBatch], vz[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const double* p = A.x + 3 * (long)k[u];
      vx[u] = p[0];
      vy[u] = p[1];
      vz[u] = p[2];
    }
    float vw[kStageBatch];
    int vs[kStageBatch];
    if (L.W) {
#pragma unroll
      for (int u = 0; u < kStageBatch; ++u)
        vw[u] = wsrc[k[u]];
    }
    if (L.S) {
#pragma unroll
      for (int u = 0; u < kStageBatch; ++u)
        vs[u] = A.z[k[u]];
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int slot = base + u * blockDim.x;
      if (slot < total) {
        L.P[3 * slot] = vx[u];
        L.P[3 * slot + 1] = vy[u];
        L.P[3 * slot + 2] = vz[u];
      }
    }
    if (L.W) {
#pragma unroll
      for (int u = 0; u < kStageBatch; ++u) {
        const int slot = base + u * blockDim.x;
        if (slot < total)
          L.W[slot] = vw[u];
      }
    }
    if (L.S) {
#pragma unroll
      for (int u = 0; u < kStageBatch; ++u) {
        const int slot = base + u * blockDim.x;
        if (slot < total)
          L.S[slot] = vs[u];
      }
    }
  }
  // the brick's own first cell is halo cell (1, 1, 1)
  a0 = sr[(E + 1) * E + 1].y;
  na = sr[H].y;
  __syncthreads();
}

// Runs body(i, l, active) over the brick's atoms after staging its halo.
template <int LPA, typename Body>
__device__ __forceinline__ void for_brick_atoms(const ForceArgs& A, const int2* bhalo, int brick, int cap,
                                                int parts, const float* wsrc, Stage& S, Body&& body) {
  const StageLds L = stage_lds(cap, parts);
  int a0, na;
  stage_brick(A, bhalo + xcd_block() * (halo_cells(brick) + 1), brick, L, wsrc, a0, na);
  S = Stage{L.P, L.W, L.S};
  const int per = blockDim.x / LPA;
  const int l = (int)(threadIdx.x % LPA);
  for (int base = 0; base < na; base += per) {
    const int k = base + (int)(threadIdx.x / LPA);
    body((long)a0 + k, l, k < na);
  }
}

// Six waves per SIMD: the register allocator is held to 80 VGPRs (at 81 the
// occupancy falls to five waves and the LJ 4M step runs 19 % slower).
template <int KIND, int LPA, bool EVAL, bool SPC>
static __global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) void k_pair_force_bk(ForceArgs A, const int2* bhalo,
                                                                                  int brick, int cap, int parts) {
  if (gate_shut(A))
    return;
  Tally tal;
  Stage S;
  for_brick_atoms<LPA>(A, bhalo, brick, cap, parts, nullptr, S, [&](long i, int l, bool active) {
    pair_atom<KIND, LPA, EVAL, true, SPC>(A, S, i, l, active, tal);
  });
  if constexpr (EVAL)
    write_partials(A, tal, 0.0);
}

// The force pass of the EAM kinds whose pair arithmetic fits (Sutton-Chen with compile-time
// exponents, Gupta; not the evaluating variants) is held to six waves per SIMD
// (80 VGPRs, k_eam_force_bk): at its unconstrained 107 VGPRs only one
// 768-thread workgroup fits a CU, and the 512 bricks of the SC 131k box run in
// two rounds.  Measured there (with the 4-slot staging): force pass 57.7 ->
// 44.0 us, rho pass 37.6 -> 34.8 us, step 116 -> 99 us; Gupta Cu 131k step
// 124 -> 110 us.  The other kinds spill in the neighbour loop at 80 VGPRs
// (Sutton-Chen with run-time exponents: force pass 70 -> 94 us) and keep the
// unconstrained kernel (k_eam_force_bk_wide).
template <int KIND, int EVAL, int EXP> constexpr bool eam_force_capped() {
  return !EVAL && ((KIND == kSuttonChen && EXP != 0) || KIND == kGupta);
}

template <int KIND, int LPA, bool EVAL, int EXP = 0>
static __global__ __launch_bounds__(1024) void k_eam_rho_bk(ForceArgs A, const int2* bhalo,
                                                                               int brick, int cap, float* w_out,
                                                                               double* rho_out) {
  if (gate_shut(A))
    return;
  double emb = 0.0;
  Stage S;
  for_brick_atoms<LPA>(A, bhalo, brick, cap, 0, nullptr, S, [&](long i, int l, bool active) {
    eam_rho_atom<KIND, LPA, EVAL, true, true, EXP>(A, S, i, l, active, w_out, rho_out, emb);
  });
  if constexpr (EVAL) {
    Tally tal;
    write_partials(A, tal, emb);
  }
}

template <int KIND, int LPA, bool EVAL, bool SPC, int EXP>
__device__ __forceinline__ void eam_force_brick(const ForceArgs& A, const int2* bhalo, int brick, int cap,
                                                int parts) {
  if (gate_shut(A))
    return;
  Tally tal;
  Stage S;
  for_brick_atoms<LPA>(A, bhalo, brick, cap, parts, A.w_rho, S, [&](long i, int l, bool active) {
    eam_force_atom<KIND, LPA, EVAL, true, SPC, true, EXP>(A, S, i, l, active, tal);
  });
  if constexpr (EVAL)
    write_partials(A, tal, 0.0);
}

template <int KIND, int LPA, bool EVAL, bool SPC, int EXP = 0>
static __global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) void k_eam_force_bk(
    ForceArgs A, const int2* bhalo, int brick, int cap, int parts) {
  eam_force_brick<KIND, LPA, EVAL, SPC, EXP>(A, bhalo, brick, cap, parts);
}

template <int KIND, int LPA, bool EVAL, bool SPC, int EXP = 0>
static __global__ __launch_bounds__(1024) void k_eam_force_bk_wide(ForceArgs A, const int2* bhalo, int brick,
                                                                    int cap, int parts) {
  eam_force_brick<KIND, LPA, EVAL, SPC, EXP>(A, bhalo, brick, cap, parts);
}

// Velocity Verlet with precomputed accelerations (several potentials):
// a_new holds the summed accelerations.
static __global__ __launch_bounds__(kBlock) void k_integrate(ForceArgs A, const double* __restrict__ anew) {
  if (gate_shut(A))
    return;
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float d2 = 0.0f;
  if (i < A.n)
    vv_tail(A, i, anew[3 * i], anew[3 * i + 1], anew[3 * i + 2], d2);
  flush_maxd2(A, d2);
}

// Drift only: xn = FittedInside((x + v dt) + a/2 dt^2) (velocityverlet.cc:60-67),
// or with mode kBeemanInit Beeman's position update with a_old = ab.
// Fixed atoms (A.fix): a fixedvel atom's acceleration is zeroed first (the
// step's force evaluation may be one made outside a step), a fixedpos atom
// keeps its position.
static __global__ __launch_bounds__(kBlock) void k_drift(ForceArgs A) {
  if (gate_shut(A))
    return;
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float d2 = 0.0f;
  if (i < A.n) {
    const double dt = A.dt;
    double p[3];
    const int f = A.fix ? A.fix[i] : 0;
    if (f & kFixVel)
      for (int q = 0; q < 3; ++q)
        A.a[3 * i + q] = 0.0;
    for (int q = 0; q < 3; ++q) {
      if (f & kFixPos) {
        p[q] = A.x[3 * i + q];
        A.xn[3 * i + q] = p[q];
        continue;
      }
      if (A.mode == kBeemanInit)
        p[q] = beeman_position(A.x[3 * i + q], A.v[3 * i + q], A.a[3 * i + q], A.ab[3 * i + q], dt);
      else
        p[q] = (A.x[3 * i + q] + A.v[3 * i + q] * dt) + A.a[3 * i + q] * 0.5 * dt * dt;
      A.xn[3 * i + q] = fitted_inside(p[q], A.ll[q]);
    }
    d2 = build_displacement2(A, i, p[0], p[1], p[2]);
  }
  flush_maxd2(A, d2);
}

// Nose-Hoover's position update (NoseHoover::AdvancePosition,
// plugins/nosehoover.cc:73-88) once the step's friction increment is known:
// atom i (caller index c) sees the friction after r + 1 increments,
// f = F + (r + 1) D, with A.nh = {F, D} and r = rank[c], the number of moving
// atoms before c (r = c with rank null: no atom is held); newacc = a - v f,
// the position ((x + v dt) + (2/3 newacc) dt^2) - (1/6 ab) dt^2 folded into
// the cell, and nn = newacc (the next velocity update's old.a).  a keeps the
// force.  A fixedpos "true" atom is skipped (twostepintegrator.cc:35-45): it
// stays where it is and adds no increment.
static __global__ __launch_bounds__(kBlock) void k_nh_drift(ForceArgs A, const int* __restrict__ rank) {
  if (gate_shut(A))
    return;
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float d2 = 0.0f;
  if (i < A.n) {
    const int fl = A.fix ? A.fix[i] : 0;
    // a fixedvel atom's acceleration is zero when the position update runs
    // (the step's force call may have been made outside a step)
    if (fl & kFixVel)
      for (int q = 0; q < 3; ++q)
        A.a[3 * i + q] = 0.0;
    double p[3];
    if (fl & kFixPos) {
      for (int q = 0; q < 3; ++q) {
        p[q] = A.x[3 * i + q];
        A.xn[3 * i + q] = p[q];
      }
    } else {
      const int c = A.perm[i];
      const double f = A.nh[0] + (double)((rank ? rank[c] : c) + 1) * A.nh[1];
      const double dt = A.dt;
      for (int q = 0; q < 3; ++q) {
        const double na = A.a[3 * i + q] - A.v[3 * i + q] * f;
        p[q] = beeman_position(A.x[3 * i + q], A.v[3 * i + q], na, A.ab[3 * i + q], dt);
        A.nn[3 * i + q] = na;
        A.xn[3 * i + q] = fitted_inside(p[q], A.ll[q]);
      }
    }
    d2 = build_displacement2(A, i, p[0], p[1], p[2]);
  }
  flush_maxd2(A, d2);
}

// The friction of a Nose-Hoover step from the kinetic sum red[0]
// (sum of m v^2 / 2): nh[0] moves past the previous step's increments, one
// per moving atom (n_moving), nh[1] = D = ((3 N / Q) k_B (T - T_set)) dt
// with T = (2/3) K / (k_B' N) over all N atoms (properties.h:39-53, k_B' its
// KBOLTZMANN) and k_B the session's kboltzmann (nosehoover.cc:69,78-79).
// gate as the force kernels'.
static __global__ void k_nh_friction(const double* __restrict__ red, double* __restrict__ nh, long n,
                                     long n_moving, double q, double temp, double dt, double kboltzmann,
                                     const unsigned int* gate, double gate_b) {
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  if (gate && (*gate == kGateClosed || sqrt((double)__uint_as_float(*gate)) > gate_b))
    return;
  const double T = (2.0 / 3.0) * (red[0] * kKin2eV) / (kBoltzmann * double(n));
  nh[0] = nh[0] + double(n_moving) * nh[1];
  nh[1] = ((3.0 * double(n) / q) * kboltzmann * (T - temp)) * dt;
}

// Kinetic energy and momentum partials (properties.h:22-81).  skip: the
// tag-aware sums (tag = true there) leave out the atoms whose flags have a
// bit of `skip` set; null: every atom.
static __global__ __launch_bounds__(kBlock) void k_kinetic(const double* __restrict__ v,
                                                    const double* __restrict__ mass, long n,
                                                    double* __restrict__ partials,
                                                    const int* __restrict__ fix = nullptr, int skip = 0) {
  __shared__ double sh[4 * (kBlock / kWave)];
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double t[4] = {0, 0, 0, 0};
  if (i < n && !(fix && (fix[i] & skip))) {
    const double m = mass[i];
    const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
    t[0] = 0.5 * m * sqmod(vx, vy, vz);
    t[1] = vx * m;
    t[2] = vy * m;
    t[3] = vz * m;
  }
  block_sum<4>(t, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q)
      partials[(long)blockIdx.x * 4 + q] = t[q];
}

// Fixed-order sum of nblk rows of `width` partials: out[q] = sum_b p[b*width+q].
static __global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ p, long nblk,
                                                            int width, double* __restrict__ out) {
  __shared__ double sh[kBlock];
  for (int q = 0; q < width; ++q) {
    double s = 0.0;
    for (long b = threadIdx.x; b < nblk; b += kBlock)
      s += p[b * width + q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o)
        sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      out[q] = sh[0];
    __syncthreads();
  }
}

} // namespace lpmdx

// Pair and EAM force kernels with the fused velocity-Verlet update.
//
// Precision split: the displacement d = x_j - x_i, its minimum image and
// r2 = |d|^2 are formed in FP64 exactly as OrthogonalCell::Displacement and
// SquareModule form them, so cutoff decisions match the reference bit for bit;
// d and r2 are then cast to FP32 for the pair and embedding arithmetic; the
// per-atom force sums are compensated FP32; accelerations, the position and
// velocity update, rho and the global energy/virial/stress sums are FP64.
//
// Full neighbour lists: every atom accumulates its own force, so there is no
// scatter and no atomics; each pair's energy, virial and stress enter half per
// direction.
#pragma once

#include "common.hpp"

namespace lpmdx {

enum PotKind {
  kLJ = 1,
  kFastLJ = 2,
  kMorse = 3,
  kBuckingham = 4,
  kSuttonChen = 5,
  kFinnisSinclair = 6,
  kGupta = 7
};

struct DevPot {
  int kind;
  int spc_sum;  // Potential::AppliesTo: negative = all pairs
  double rc2;   // pair filter r2 < rc2 in FP64 (list cutoff and own cutoff)
  float f[12];  // kind-specific FP32 parameters (see PotParams in engine.hip)
  int ni, mi;   // integer exponents (Sutton-Chen), 0 if not integral
  // fastlj table (FP32 copies of the reference's FP64 tables)
  const float* tab_e;
  const float* tab_f;
  double tab_cut; // cutoff, for k = floor(bins * r / cutoff) in FP64
  int bins;
};

__device__ __forceinline__ float ipow(float x, int n) {
  float r = 1.0f;
  float b = x;
  while (n) {
    if (n & 1)
      r *= b;
    b *= b;
    n >>= 1;
  }
  return r;
}

__device__ __forceinline__ float powi_or_powf(float x, int ni, float nf) {
  return ni > 0 ? ipow(x, ni) : powf(x, nf);
}

// Pair potentials: force on i is fscal * rij (rij = x_j - x_i), pair energy e.
template <int KIND>
__device__ __forceinline__ void pair_eval(const DevPot& P, double r2d, float r2, float& fscal,
                                          float& e) {
  if constexpr (KIND == kLJ) {
    // lennardjones.cc:53-67: E = 4 eps (s12 - s6), ff = -48 eps/r2 (s12 - s6/2)
    const float inv = 1.0f / r2;
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
    const float r = sqrtf(r2);
    const float ex = expf(-P.f[0] * (r - P.f[1]));
    const float om = 1.0f - ex;
    fscal = P.f[3] / r * om * ex;
    e = P.f[2] * om * om;
  } else if constexpr (KIND == kBuckingham) {
    // buckingham.cc:58-75: E = B1 e^{-r/Ro} - B2/r^6, ff = 6 B2/r^8 - B1 e^{-r/Ro}/(Ro r)
    const float r = sqrtf(r2);
    const float ir2 = 1.0f / r2;
    const float ir6 = ir2 * ir2 * ir2;
    const float ex = P.f[0] * expf(-r * P.f[3]);
    fscal = P.f[4] * ir6 * ir2 - ex * P.f[3] / r;
    e = ex - P.f[1] * ir6;
  }
}

// EAM pieces in terms of r: pair energy V, dV/dr, rho(r), and G(r) such that
// the many-body force on i along r-hat is (w_i + w_j) G(r), w = rho^{-1/2}.
template <int KIND>
__device__ __forceinline__ float eam_rho(const DevPot& P, float r) {
  if constexpr (KIND == kSuttonChen) { // suttonchen.cc:78-81: a^m r^-m
    return powi_or_powf(P.f[1] / r, P.mi, P.f[3]);
  } else if constexpr (KIND == kGupta) { // gupta.cc:69
    return expf(-2.0f * P.f[4] * (r - P.f[1]) / P.f[1]);
  } else { // finnissinclair.cc:82-87
    const float d = P.f[6];
    if (r <= d) {
      const float t = r - d;
      return (t * t) * (1.0f + P.f[4] * t / d);
    }
    return 0.0f;
  }
}

template <int KIND>
__device__ __forceinline__ void eam_pair(const DevPot& P, float r, float& V, float& dV, float& G) {
  if constexpr (KIND == kSuttonChen) {
    // suttonchen.cc:73-95 with f = {e, a, n, m, c}
    const float t = P.f[1] / r;
    const float tn = powi_or_powf(t, P.ni, P.f[2]);
    const float tm = powi_or_powf(t, P.mi, P.f[3]);
    V = P.f[0] * tn;
    dV = -P.f[2] * V / r;
    G = 0.5f * P.f[3] * P.f[4] * P.f[0] * tm / r;
  } else if constexpr (KIND == kGupta) {
    // gupta.cc:67-82 with f = {A, r0, p, B, q}
    const float ep = expf(-P.f[2] * (r - P.f[1]) / P.f[1]);
    const float eq = expf(-2.0f * P.f[4] * (r - P.f[1]) / P.f[1]);
    V = P.f[0] * ep;
    dV = -(P.f[0] * P.f[2] / P.f[1]) * ep;
    G = (P.f[3] * P.f[4] / P.f[1]) * eq;
  } else {
    // finnissinclair.cc:75-108 with f = {c0, c1, c2, A, B, c, d}
    const float c = P.f[5], d = P.f[6];
    V = 0.0f;
    dV = 0.0f;
    G = 0.0f;
    if (r <= c) {
      const float t = r - c;
      const float poly = P.f[0] + P.f[1] * r + P.f[2] * r * r;
      V = t * t * poly;
      dV = 2.0f * t * poly + t * t * (P.f[1] + 2.0f * P.f[2] * r);
    }
    if (r <= d) {
      const float t = r - d;
      G = -0.5f * P.f[3] * (2.0f * t + 3.0f * P.f[4] * t * t / d);
    }
  }
}

// Integration modes of the force kernels.
enum StepMode {
  kForcesOnly = 0, // a = F/m (or a += with accumulate)
  kInitDrift = 1,  // a = F/m, then next positions xn = drift(x, v, a)
  kStep = 2        // a = F/m, v += dt/2 (a_old + a), xn = drift(x, v, a)
};

struct ForceArgs {
  const double* x;   // [3n] current positions
  double* xn;        // [3n] next positions (modes 1, 2)
  double* v;         // [3n]
  double* a;         // [3n] in: a_old (mode 2); out: new accelerations
  const int* z;      // [n] atomic numbers
  const double* ffm; // [n] forcefactor / m
  const int* nbr;
  const int* cnt;
  const double* w_rho;  // EAM: [n] rho^{-1/2} (FP64)
  double* partials;     // [nblocks * 12] when evaluating energies
  unsigned int* maxd2;  // per-step max displacement^2 (float bits)
  DevPot P;
  Box b;
  double ll[3]; // |cell vector q| as FittedInside reads it
  double dt;
  long n;
  long stride;
  int mode;
  int accumulate; // add to a instead of overwriting (several potentials)
};

// Velocity Verlet tail for atom i given its new acceleration (FP64):
// plugins/velocityverlet.cc:60-75, in the reference's order of operations.
__device__ __forceinline__ void vv_tail(const ForceArgs& A, long i, double ax, double ay,
                                        double az, float& d2) {
  double* a = A.a + 3 * i;
  d2 = 0.0f;
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
  const float ddx = (float)(px - x[0]), ddy = (float)(py - x[1]), ddz = (float)(pz - x[2]);
  d2 = ddx * ddx + ddy * ddy + ddz * ddz;
  double* xn = A.xn + 3 * i;
  xn[0] = fitted_inside(px, A.ll[0]);
  xn[1] = fitted_inside(py, A.ll[1]);
  xn[2] = fitted_inside(pz, A.ll[2]);
}

__device__ __forceinline__ void flush_maxd2(const ForceArgs& A, float d2) {
  float m = wave_max(d2);
  if ((threadIdx.x & (kWave - 1)) == 0 && A.mode != kForcesOnly && m > 0.0f)
    atomicMax(A.maxd2, __float_as_uint(m));
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

template <int LPA> __device__ __forceinline__ float group_sum(float v) {
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
  __shared__ double sh[9 * (kBlock / kWave)];
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

template <int KIND, int LPA, bool EVAL>
__global__ __launch_bounds__(kBlock) void k_pair_force(ForceArgs A) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long i = t / LPA;
  const int l = (int)(t % LPA);
  const bool active = i < A.n;
  CompF fx, fy, fz;
  Tally tal;
  if (active) {
    const double xi = A.x[3 * i], yi = A.x[3 * i + 1], zi = A.x[3 * i + 2];
    const int c = A.cnt[i];
    const int si = A.z[i];
    const bool all = A.P.spc_sum < 0;
    const long rowlen = A.stride * LPA;
    for (int k = l; k < c; k += LPA) {
      const int j = A.nbr[(long)(k / LPA) * rowlen + i * LPA + l];
      double dx = A.x[3 * (long)j] - xi;
      double dy = A.x[3 * (long)j + 1] - yi;
      double dz = A.x[3 * (long)j + 2] - zi;
      min_image(A.b, dx, dy, dz);
      const double r2 = sqmod(dx, dy, dz);
      if (r2 < A.P.rc2 && (all || si + A.z[j] == A.P.spc_sum)) {
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz, fr2 = (float)r2;
        float fs, e;
        pair_eval<KIND>(A.P, r2, fr2, fs, e);
        fx.add(fs * fdx);
        fy.add(fs * fdy);
        fz.add(fs * fdz);
        if constexpr (EVAL)
          tal.add(fs, fdx, fdy, fdz, fr2, e);
      }
    }
  }
  float Fx = group_sum<LPA>(fx.value()), Fy = group_sum<LPA>(fy.value()),
        Fz = group_sum<LPA>(fz.value());
  float d2 = 0.0f;
  if (active && l == 0) {
    const double m = A.ffm[i];
    vv_tail(A, i, (double)Fx * m, (double)Fy * m, (double)Fz * m, d2);
  }
  flush_maxd2(A, d2);
  if constexpr (EVAL)
    write_partials(A, tal, 0.0);
}

// EAM pass 1: rho_i over the full list within the list cutoff (no species
// check, metalpotential.cc:121-132); writes rho_i^{-1/2} and, when evaluating,
// the embedding energy F(rho_i) = -K sqrt(rho_i) into the block partials.
template <int KIND, int LPA, bool EVAL>
__global__ __launch_bounds__(kBlock) void k_eam_rho(ForceArgs A, double* w_out, double* rho_out) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long i = t / LPA;
  const int l = (int)(t % LPA);
  const bool active = i < A.n;
  double rho = 0.0;
  if (active) {
    const double xi = A.x[3 * i], yi = A.x[3 * i + 1], zi = A.x[3 * i + 2];
    const int c = A.cnt[i];
    const long rowlen = A.stride * LPA;
    for (int k = l; k < c; k += LPA) {
      const int j = A.nbr[(long)(k / LPA) * rowlen + i * LPA + l];
      double dx = A.x[3 * (long)j] - xi;
      double dy = A.x[3 * (long)j + 1] - yi;
      double dz = A.x[3 * (long)j + 2] - zi;
      min_image(A.b, dx, dy, dz);
      const double r2 = sqmod(dx, dy, dz);
      if (r2 < A.P.rc2)
        rho += (double)eam_rho<KIND>(A.P, sqrtf((float)r2));
    }
  }
  rho = group_sum<LPA>(rho);
  double emb = 0.0;
  if (active && l == 0) {
    w_out[i] = 1.0 / sqrt(rho);
    if (rho_out)
      rho_out[i] = rho;
    emb = -(double)A.P.f[11] * sqrt(rho);
  }
  if constexpr (EVAL) {
    Tally tal;
    write_partials(A, tal, emb);
  }
}

// EAM pass 2: pair + many-body forces over the full list with AppliesTo
// (metalpotential.cc:140-169), fused with the integrator tail.
template <int KIND, int LPA, bool EVAL>
__global__ __launch_bounds__(kBlock) void k_eam_force(ForceArgs A) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long i = t / LPA;
  const int l = (int)(t % LPA);
  const bool active = i < A.n;
  CompF fx, fy, fz;
  Tally tal;
  if (active) {
    const double xi = A.x[3 * i], yi = A.x[3 * i + 1], zi = A.x[3 * i + 2];
    const int c = A.cnt[i];
    const int si = A.z[i];
    const float wi = (float)A.w_rho[i];
    const bool all = A.P.spc_sum < 0;
    const long rowlen = A.stride * LPA;
    for (int k = l; k < c; k += LPA) {
      const int j = A.nbr[(long)(k / LPA) * rowlen + i * LPA + l];
      double dx = A.x[3 * (long)j] - xi;
      double dy = A.x[3 * (long)j + 1] - yi;
      double dz = A.x[3 * (long)j + 2] - zi;
      min_image(A.b, dx, dy, dz);
      const double r2 = sqmod(dx, dy, dz);
      if (r2 < A.P.rc2 && (all || si + A.z[j] == A.P.spc_sum)) {
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz, fr2 = (float)r2;
        const float r = sqrtf(fr2);
        float V, dV, G;
        eam_pair<KIND>(A.P, r, V, dV, G);
        const float s = (dV + (wi + (float)A.w_rho[j]) * G) / r;
        fx.add(s * fdx);
        fy.add(s * fdy);
        fz.add(s * fdz);
        if constexpr (EVAL)
          tal.add(s, fdx, fdy, fdz, fr2, V);
      }
    }
  }
  float Fx = group_sum<LPA>(fx.value()), Fy = group_sum<LPA>(fy.value()),
        Fz = group_sum<LPA>(fz.value());
  float d2 = 0.0f;
  if (active && l == 0) {
    const double m = A.ffm[i];
    vv_tail(A, i, (double)Fx * m, (double)Fy * m, (double)Fz * m, d2);
  }
  flush_maxd2(A, d2);
  if constexpr (EVAL)
    write_partials(A, tal, 0.0);
}

// Velocity Verlet with precomputed accelerations (several potentials):
// a_new holds the summed accelerations.
__global__ __launch_bounds__(kBlock) void k_integrate(ForceArgs A, const double* __restrict__ anew) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float d2 = 0.0f;
  if (i < A.n)
    vv_tail(A, i, anew[3 * i], anew[3 * i + 1], anew[3 * i + 2], d2);
  flush_maxd2(A, d2);
}

// Drift only: xn = FittedInside((x + v dt) + a/2 dt^2) (velocityverlet.cc:60-67).
__global__ __launch_bounds__(kBlock) void k_drift(ForceArgs A) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= A.n)
    return;
  const double dt = A.dt;
  for (int q = 0; q < 3; ++q) {
    double p = (A.x[3 * i + q] + A.v[3 * i + q] * dt) + A.a[3 * i + q] * 0.5 * dt * dt;
    A.xn[3 * i + q] = fitted_inside(p, A.ll[q]);
  }
}

// Kinetic energy and momentum partials (properties.h:22-81).
__global__ __launch_bounds__(kBlock) void k_kinetic(const double* __restrict__ v,
                                                    const double* __restrict__ mass, long n,
                                                    double* __restrict__ partials) {
  __shared__ double sh[4 * (kBlock / kWave)];
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double t[4] = {0, 0, 0, 0};
  if (i < n) {
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
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ p, long nblk,
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

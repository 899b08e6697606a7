// Shared device helpers for the LPMD engine (gfx950, wave64).
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lpmdx {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kMaxWaves = 16; // wavefronts of the largest workgroup (1024 threads)

// Unit constants as the reference's potentials and properties see them:
// forcefactor re-parsed from its 8-decimal session string
// (liblpmd/runtime/units.cc:6, liblpmd/runtime/session.cc:17-23,31),
// KIN2EV/KBOLTZMANN macros (liblpmd/lpmd/properties.h:10-11).
constexpr double kForceFactor = 0.00964853;
constexpr double kKin2eV = 103.64269;
constexpr double kBoltzmann = 8.6173422E-05;
constexpr double kPressFactor = 160217.65;
constexpr double kUa3ToGrCm3 = 1.6605; // amu/A^3 -> g/cm^3 (units.cc:10), the session's "1.66050000"
// log2 e, for the kernels' base-2 exponentials
constexpr double kLog2e = 1.4426950408889634;

struct Box {
  double L[3];
  double half[3]; // 0.5 * L, as OrthogonalCell::Displacement forms it
  double inv[3];  // 1 / L
};

// OrthogonalCell::Displacement (liblpmd/lpmd/orthogonalcell.h:131-149): one
// wrap per axis, >= on the positive side.  Written with selects instead of
// branches: d + (-L) and d + (+L) round exactly like d - L and d + L, and
// d + 0.0 equals d (a -0.0 component may come out as +0.0, which no squared
// distance or force component can tell apart).  Compiled without contraction.
__device__ __forceinline__ double wrap1(double d, double L, double h) {
  const double w = (d >= h) ? -L : ((d < -h) ? L : 0.0);
  return d + w;
}

// The same wrap for a displacement known to lie within m of 0 or of +-L, with
// m < L/2 (1 - 1e-6): t = rint(d / L) is then 1, 0 or -1 exactly where wrap1
// adds -L, 0 or +L, and fma(-t, L, d) rounds d - t L once, as d + w does
// (t L is exact).  Three instructions per axis instead of two compares and
// four selects.
__device__ __forceinline__ double wrap1_near(double d, double L, double inv) {
  const double t = rint(d * inv);
  return fma(-t, L, d);
}

__device__ __forceinline__ void min_image_near(const Box& b, double& dx, double& dy, double& dz) {
  dx = wrap1_near(dx, b.L[0], b.inv[0]);
  dy = wrap1_near(dy, b.L[1], b.inv[1]);
  dz = wrap1_near(dz, b.L[2], b.inv[2]);
}

__device__ __forceinline__ void min_image(const Box& b, double& dx, double& dy, double& dz) {
  dx = wrap1(dx, b.L[0], b.half[0]);
  dy = wrap1(dy, b.L[1], b.half[1]);
  dz = wrap1(dz, b.L[2], b.half[2]);
}

// GenericVector::SquareModule (vector.h:110): (x*x + y*y) + z*z.
__device__ __forceinline__ double sqmod(double x, double y, double z) {
  return x * x + y * y + z * z;
}

// OrthogonalCell::FittedInside (orthogonalcell.h:102-114); x == L stays.
__device__ __forceinline__ double fitted_inside(double x, double ll) {
  if (x < 0.0)
    x -= ll * floor(x / ll);
  else if (x > ll)
    x -= ll * floor(x / ll);
  return x;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Remap the block index so that XCD x works through one contiguous range of
// logical blocks (a slab of the cell-sorted atoms): the neighbour cells a
// block gathers from are then in flight on the same XCD and hit its L2.
constexpr int kNumXcd = 8;
constexpr int kNumCu = 256;
__device__ __forceinline__ long xcd_block() {
  const long nb = gridDim.x;
  const long b = blockIdx.x;
  const long q = nb / kNumXcd, r = nb % kNumXcd;
  const long x = b % kNumXcd;
  return x * q + (x < r ? x : r) + b / kNumXcd;
}

// A value every lane of the wavefront holds alike (a tile's index, its
// system), made visibly uniform to the compiler: what is loaded through it
// (a system's box, grid, potentials) then lives in scalar registers.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The same for a whole record of 32-bit words (a kernel's per-system
// arguments assembled from memory inside the kernel).
template <typename T> __device__ __forceinline__ T wave_uniform_record(const T& v) {
  static_assert(sizeof(T) % 4 == 0 && std::is_trivially_copyable<T>::value, "a record of 32-bit words");
  T r;
  const int* a = reinterpret_cast<const int*>(&v);
  int* b = reinterpret_cast<int*>(&r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k)
    b[k] = __builtin_amdgcn_readfirstlane(a[k]);
  return r;
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1)
    v += __shfl_xor(v, o, kWave);
  return v;
}

template <typename T> __device__ __forceinline__ T wave_max(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, kWave);
    v = v > w ? v : w;
  }
  return v;
}

// Atomics on one word from many wavefronts serialize: 62 500 per-wavefront
// atomicMax on one address take 0.71 ms on MI355X (11 ns each), against
// 6 us when each is skipped unless a relaxed read shows it can raise the
// word.  The maximum is the same either way: stored values only grow, and a
// skipped value is at most one already stored.
__device__ __forceinline__ void atomic_max_sparse(unsigned int* p, unsigned int v) {
  if (v > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(p, v);
}

// Sums that every wavefront adds to are spread over kSumSlots counters, one
// 128-byte line each (slot = wavefront index mod kSumSlots), and summed by
// k_sum_slots afterwards: integer sums, so the total does not depend on the
// order.
constexpr int kSumSlots = 64;
constexpr int kSlotStride = 16; // 64-bit words per slot

__device__ __forceinline__ void atomic_add_spread(unsigned long long* slots, unsigned long long v) {
  const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  atomicAdd(slots + (w & (kSumSlots - 1)) * kSlotStride, v);
}

// One wavefront: *out = sum of the kSumSlots slots.
static __global__ __launch_bounds__(kWave) void k_sum_slots(const unsigned long long* __restrict__ slots,
                                                     unsigned long long* __restrict__ out) {
  const int t = threadIdx.x;
  const unsigned long long v = wave_sum(t < kSumSlots ? slots[t * kSlotStride] : 0ull);
  if (t == 0)
    *out = v;
}

// Sum of NV doubles over a 256-thread block in a fixed order; the result is
// valid in thread 0.  `sh` needs NV * (kBlock / kWave) doubles.
template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      sh[w * NV + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int q = 0; q < (int)(blockDim.x / kWave); ++q)
        s += sh[q * NV + k];
      v[k] = s;
    }
  }
}

// Per-launch control block in device memory.
constexpr int kMaxChunk = 4096;
struct Ctrl {
  unsigned int maxd2[kMaxChunk];  // per step: max |dx|^2 as float bits
  unsigned int max_count;         // largest list length of the last build
  unsigned int pad0;
  unsigned long long sum_count;   // total list entries of the last build (8-byte aligned)
  unsigned int halo_max;          // largest brick halo of the last build (atoms)
  unsigned int brick_max;         // most atoms in one brick of the last build
  unsigned int cell_max;          // most atoms in one cell (lcbinary)
  unsigned int pad[9];
};
static_assert(offsetof(Ctrl, sum_count) % 8 == 0, "64-bit counter must be 8-byte aligned");
static_assert(offsetof(Ctrl, sum_count) == offsetof(Ctrl, max_count) + 8, "counter block layout");

} // namespace lpmdx

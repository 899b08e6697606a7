// Host-side helpers shared by the single-system engine and the batched engine:
// status/error plumbing of the C ABI, device buffers, the reference's plugin
// keyword parsing and the potential parameter sets.
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
#include "table.hpp"

namespace lpmdx {

// message of the last failed C ABI call on this thread (defined in engine.hip)
extern thread_local std::string g_err;


struct Error {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

inline void hipchk(hipError_t e, const char* what) {
  if (e != hipSuccess)
    fail(e == hipErrorOutOfMemory ? LPMDX_ENOMEM : LPMDX_EDEVICE,
         std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F> int guard(F&& f) {
  try {
    f();
    return LPMDX_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return LPMDX_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return LPMDX_EINVAL;
  }
}

// ---------------------------------------------------------------------------
// device buffers

template <typename T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t m) {
    if (m <= n && p)
      return;
    release();
    if (m == 0)
      m = 1;
    hipchk(hipMalloc(&p, m * sizeof(T)), "hipMalloc");
    n = m;
  }
};

// ---------------------------------------------------------------------------
// plugin keyword strings (Module::ProcessArguments, liblpmd/plugins/module.cc:77-98)

struct Keywords {
  std::map<std::string, std::string> kv;
  Keywords(const char* args, std::initializer_list<std::pair<const char*, const char*>> defaults) {
    for (auto& d : defaults)
      kv[d.first] = d.second;
    std::istringstream in(args ? args : "");
    std::vector<std::string> w;
    std::string s;
    while (in >> s)
      w.push_back(s);
    // words are taken as key value pairs; a dangling key is the reference's
    // "Unexpected end of input while reading module arguments" (module.cc:101-104)
    if (w.size() % 2)
      fail(LPMDX_EINVAL, "Unexpected end of input while reading module arguments");
    for (size_t k = 0; k < w.size(); k += 2)
      kv[w[k]] = w[k + 1];
  }
  // Parameter conversions, liblpmd/lpmd/parameter.h:26-28: atof / atoi / "true";
  // a keyword without default and without value reads as "".
  double d(const char* k) const {
    auto it = kv.find(k);
    return it == kv.end() ? atof("") : atof(it->second.c_str());
  }
  int i(const char* k) const {
    auto it = kv.find(k);
    return it == kv.end() ? atoi("") : atoi(it->second.c_str());
  }
  bool b(const char* k) const {
    auto it = kv.find(k);
    return it != kv.end() && it->second == "true";
  }
  std::string s(const char* k) const {
    auto it = kv.find(k);
    return it == kv.end() ? std::string() : it->second;
  }
  bool has(const char* k) const { return kv.count(k) > 0; }
};

// ---------------------------------------------------------------------------
// potentials

struct PotParams {
  int kind = 0;
  int spc_sum = -2;
  double cutoff = HUGE_VAL; // Potential::rcutoff default (potential.cc:13)
  double p[10] = {0};
  long bins = 0;
  DBuf<float> tab_e, tab_f;
  bool metal() const { return metal_kind(kind); }
};

inline std::unique_ptr<PotParams> make_potential(const std::string& name, const char* args, int s1, int s2) {
  auto P = std::make_unique<PotParams>();
  P->spc_sum = s1 + s2;
  if (name == "lennardjones") { // plugins/lennardjones.cc:11-22
    Keywords k(args, {});
    P->kind = kLJ;
    P->p[0] = k.d("sigma");
    P->p[1] = k.d("epsilon");
    P->cutoff = P->p[2] = k.d("cutoff");
  } else if (name == "fastlj") { // plugins/fastlj.cc:11-26,64-80
    Keywords k(args, {{"bins", "500"}});
    P->kind = kFastLJ;
    P->p[0] = k.d("sigma");
    P->p[1] = k.d("epsilon");
    P->cutoff = P->p[2] = k.d("cutoff");
    P->bins = k.i("bins");
    if (P->bins <= 0)
      fail(LPMDX_EINVAL, "fastlj: bins must be positive");
    std::vector<float> e(P->bins), f(P->bins);
    const double sigma = P->p[0], epsilon = P->p[1], cutoff = P->p[2];
    for (long i = 0; i < P->bins; ++i) {
      double r = cutoff * double(i) / double(P->bins);
      double r6 = pow(sigma / r, 6.0e0);
      double r12 = r6 * r6;
      e[i] = (float)(4.0e0 * epsilon * (r12 - r6));
      f[i] = (float)(-48.0e0 * (epsilon / (r * r)) * (r12 - 0.50e0 * r6));
    }
    P->tab_e.alloc(P->bins);
    P->tab_f.alloc(P->bins);
    hipchk(hipMemcpy(P->tab_e.p, e.data(), P->bins * sizeof(float), hipMemcpyHostToDevice), "table");
    hipchk(hipMemcpy(P->tab_f.p, f.data(), P->bins * sizeof(float), hipMemcpyHostToDevice), "table");
  } else if (name == "morse") { // plugins/morse.cc:11-21 (no cutoff keyword)
    Keywords k(args, {});
    P->kind = kMorse;
    P->p[0] = k.d("depth");
    P->p[1] = k.d("a");
    P->p[2] = k.d("re");
  } else if (name == "buckingham") { // plugins/buckingham.cc:11-24
    Keywords k(args, {});
    P->kind = kBuckingham;
    P->p[0] = k.d("B1");
    P->p[1] = k.d("B2");
    P->p[2] = k.d("Ro");
    P->cutoff = P->p[3] = k.d("cutoff");
  } else if (name == "suttonchen") { // plugins/suttonchen.cc:11-32
    Keywords k(args, {{"corrections", "false"}});
    P->kind = kSuttonChen;
    P->p[0] = k.d("e");
    P->p[1] = k.d("a");
    P->p[2] = k.d("n");
    P->p[3] = k.d("m");
    P->p[4] = k.d("c");
    P->cutoff = P->p[5] = k.d("cutoff");
  } else if (name == "finnissinclair") { // plugins/finnissinclair.cc:10-31
    Keywords k(args, {});
    P->kind = kFinnisSinclair;
    const char* keys[7] = {"c0", "c1", "c2", "A", "B", "c", "d"};
    for (int q = 0; q < 7; ++q)
      P->p[q] = k.d(keys[q]);
    P->cutoff = (P->p[6] >= P->p[5]) ? P->p[6] : P->p[5];
  } else if (name == "gupta") { // plugins/gupta.cc:10-27
    Keywords k(args, {});
    P->kind = kGupta;
    const char* keys[5] = {"A", "r0", "p", "B", "q"};
    for (int q = 0; q < 5; ++q)
      P->p[q] = k.d(keys[q]);
    P->cutoff = P->p[5] = k.d("cutoff");
  } else if (name == "tabulatedpair") { // plugins/tabulatedpair.cc:12-96
    // the fastlj kernels' table lookup, k = floor(bins r / cutoff), on the file's table
    Keywords k(args, {{"file", ""}});
    const std::string file = k.s("file");
    if (file.empty())
      fail(LPMDX_EINVAL, "tabulatedpair: no table file given (keyword file)");
    const PairTable T = read_pair_table(file);
    if (!T.ok)
      fail(LPMDX_EINVAL, "tabulatedpair: cannot open the table file " + file);
    if (T.e.empty() || !(T.cutoff > 0.0))
      fail(LPMDX_EINVAL, "tabulatedpair: no table rows (r E ff) with a positive r in " + file);
    P->kind = kFastLJ;
    P->bins = (long)T.e.size();
    P->cutoff = P->p[2] = T.cutoff;
    std::vector<float> e(P->bins), f(P->bins);
    for (long i = 0; i < P->bins; ++i) {
      e[i] = (float)T.e[i];
      f[i] = (float)T.f[i];
    }
    P->tab_e.alloc(P->bins);
    P->tab_f.alloc(P->bins);
    hipchk(hipMemcpy(P->tab_e.p, e.data(), P->bins * sizeof(float), hipMemcpyHostToDevice), "table");
    hipchk(hipMemcpy(P->tab_f.p, f.data(), P->bins * sizeof(float), hipMemcpyHostToDevice), "table");
  } else if (name == "harmonic") { // plugins/harmonic.cc:11-19 (no cutoff keyword)
    Keywords k(args, {});
    P->kind = kHarmonic;
    P->p[0] = k.d("k");
    P->p[1] = k.d("a");
  } else if (name == "lennardjonesMod") { // plugins/lennardjonesMod.cc:11-28
    Keywords k(args, {{"m", "12"}, {"n", "6"}, {"cohesive", "1"}});
    P->kind = kLJMod;
    P->p[0] = k.d("sigma");
    P->p[1] = k.d("epsilon");
    P->p[2] = k.d("m");
    P->p[3] = k.d("n");
    P->p[4] = k.d("cohesive");
    P->cutoff = P->p[5] = k.d("cutoff");
  } else if (name == "finnissinclair-ext") { // plugins/finnissinclair-ext.cc:11-36
    Keywords k(args, {});
    P->kind = kFinnisSinclairExt;
    const char* keys[9] = {"c0", "c1", "c2", "c3", "c4", "A", "B", "c", "d"};
    for (int q = 0; q < 9; ++q)
      P->p[q] = k.d(keys[q]);
    P->cutoff = (P->p[8] >= P->p[7]) ? P->p[8] : P->p[7];
  } else {
    fail(LPMDX_EINVAL, "unknown potential plugin '" + name + "'");
  }
  for (int q = 0; q < 10; ++q)
    if (!std::isfinite(P->p[q]))
      fail(LPMDX_ENONFINITE, name + ": non-finite parameter");
  return P;
}

// v when it is an integer that powu takes (1 .. 31), else 0 (powf at run time)
inline int int_exponent(double v) {
  double r = std::nearbyint(v);
  return (r == v && r >= 1 && r <= 31) ? (int)r : 0;
}

inline DevPot device_view(const PotParams& P, double rc2_force, double rc2_rho) {
  DevPot d{};
  d.kind = P.kind;
  d.spc_sum = P.spc_sum;
  d.rc2 = rc2_force;
  const double* p = P.p;
  switch (P.kind) {
  case kLJ:
    d.f[0] = (float)(p[0] * p[0]);
    d.f[1] = (float)(48.0 * p[1]);
    d.f[2] = (float)(4.0 * p[1]);
    break;
  case kFastLJ:
    d.tab_e = P.tab_e.p;
    d.tab_f = P.tab_f.p;
    d.tab_cut = p[2];
    d.bins = (int)P.bins;
    break;
  case kMorse:
    // {a, re, D, 2 a D, a log2 e}: the kernels' exponentials are base 2
    d.f[0] = (float)p[1];
    d.f[1] = (float)p[2];
    d.f[2] = (float)p[0];
    d.f[3] = (float)(2.0 * p[1] * p[0]);
    d.f[4] = (float)(p[1] * kLog2e);
    break;
  case kBuckingham:
    // {B1, B2, Ro, 1/Ro, 6 B2, log2 e / Ro}
    d.f[0] = (float)p[0];
    d.f[1] = (float)p[1];
    d.f[2] = (float)p[2];
    d.f[3] = (float)(1.0 / p[2]);
    d.f[4] = (float)(6.0 * p[1]);
    d.f[5] = (float)(kLog2e / p[2]);
    break;
  case kSuttonChen:
    // {e, a, n, m, c, -, -, -, -n e, m c e / 2, -, K = c e}
    for (int q = 0; q < 5; ++q)
      d.f[q] = (float)p[q];
    d.f[8] = (float)(-p[2] * p[0]);
    d.f[9] = (float)(0.5 * p[3] * p[4] * p[0]);
    d.ni = int_exponent(p[2]);
    d.mi = int_exponent(p[3]);
    d.f[11] = (float)(p[4] * p[0]);
    break;
  case kGupta:
    // {A, r0, p, B, q, p L, 2q L, 2q/r0 L, p/r0 L, -A p/r0, B q/r0, K = B}, L = log2 e:
    // e^{-p(r-r0)/r0} = 2^{(p - (p/r0) r) L}, e^{-2q(r-r0)/r0} = 2^{(2q - (2q/r0) r) L}
    for (int q = 0; q < 5; ++q)
      d.f[q] = (float)p[q];
    d.f[5] = (float)(p[2] * kLog2e);
    d.f[6] = (float)(2.0 * p[4] * kLog2e);
    d.f[7] = (float)(2.0 * p[4] / p[1] * kLog2e);
    d.f[8] = (float)(p[2] / p[1] * kLog2e);
    d.f[9] = (float)(-p[0] * p[2] / p[1]);
    d.f[10] = (float)(p[3] * p[4] / p[1]);
    d.f[11] = (float)p[3];
    break;
  case kFinnisSinclair:
    // {c0, c1, c2, A, B, c, d, -, B/d, 3B/d, -A/2, K = A}
    for (int q = 0; q < 7; ++q)
      d.f[q] = (float)p[q];
    d.f[8] = (float)(p[4] / p[6]);
    d.f[9] = (float)(3.0 * p[4] / p[6]);
    d.f[10] = (float)(-0.5 * p[3]);
    d.f[11] = (float)p[3];
    break;
  case kHarmonic:
    // {k, a, k/2}
    d.f[0] = (float)p[0];
    d.f[1] = (float)p[1];
    d.f[2] = (float)(0.5 * p[0]);
    break;
  case kLJMod:
    // {sigma^2, 4 eps, m, c n, c, m/2, n/2}; mi, ni: m/2, n/2 when integral
    d.f[0] = (float)(p[0] * p[0]);
    d.f[1] = (float)(4.0 * p[1]);
    d.f[2] = (float)p[2];
    d.f[3] = (float)(p[4] * p[3]);
    d.f[4] = (float)p[4];
    d.f[5] = (float)(0.5 * p[2]);
    d.f[6] = (float)(0.5 * p[3]);
    d.mi = int_exponent(0.5 * p[2]);
    d.ni = int_exponent(0.5 * p[3]);
    break;
  case kFinnisSinclairExt:
    // p = {c0, c1, c2, c3, c4, A, B, c, d} -> f = {c0, c1, c2, A, c3, c, d, c4, B^2, 4 B^2, -A/2, K = A}
    d.f[0] = (float)p[0];
    d.f[1] = (float)p[1];
    d.f[2] = (float)p[2];
    d.f[3] = (float)p[5];
    d.f[4] = (float)p[3];
    d.f[5] = (float)p[7];
    d.f[6] = (float)p[8];
    d.f[7] = (float)p[4];
    d.f[8] = (float)(p[6] * p[6]);
    d.f[9] = (float)(4.0 * p[6] * p[6]);
    d.f[10] = (float)(-0.5 * p[5]);
    d.f[11] = (float)p[5];
    for (int q = 0; q < 5; ++q)
      d.g[q] = p[q];
    d.g[5] = p[7];
    break;
  }
  d.tab_cut = P.kind == kFastLJ ? p[2] : 0.0;
  // rho pass filter travels in the (otherwise unused) FP64 slot below
  (void)rc2_rho;
  return d;
}

// ---------------------------------------------------------------------------
// cell managers

enum CmKind { kNone = 0, kCmLinkedCell = 1, kCmVerlet = 2, kCmMinImage = 3, kCmLcBinary = 4 };

struct CellMgr {
  int kind = kNone;
  double cutoff = 0.0;
  // linkedcell (plugins/linkedcell.cc:17-41), lcbinary (plugins/lcbinary.cc:9-33)
  int nx = 0, ny = 0, nz = 0, side = 0;
  bool mode_auto = false, warn_outside = true, grid_fixed = false;
  // verletlist (plugins/verletlist.cc:12-29)
  double extcut = 0.0, skin = -1.0;
  long maxneighbors = 80, each = 5;
  long ucalls = 0;
};

// linkedcell.cc:146-165
inline int lc_auto_n(double edge, double cutoff) {
  double fnn = (cutoff / 2.5);
  int n = 1;
  while (true) {
    double previo = fabs(edge / n - fnn);
    n++;
    double actual = fabs(edge / n - fnn);
    if (actual < 1E-3)
      break;
    else if (actual > previo) {
      n--;
      break;
    }
  }
  if ((n % 2) != 0)
    n--;
  return n;
}

// lcbinary.cc:88-145: the same stepping toward cutoff/5, an odd n rounded up
inline int lcb_auto_n(double edge, double cutoff) {
  double fnn = cutoff / 5.0;
  int n = 1;
  while (true) {
    double previo = fabs(edge / n - fnn);
    n++;
    double actual = fabs(edge / n - fnn);
    if (actual < 1E-3)
      break;
    else if (actual > previo) {
      n--;
      break;
    }
  }
  if ((n % 2) != 0)
    n++;
  return n;
}

inline double edge_module(double L) { return sqrt(L * L + 0.0 * 0.0 + 0.0 * 0.0); }

// ElemMass[Z] of liblpmd/support/elements.cc:20-35 (the mass BasicAtom takes,
// basicatom.h:57-67); Z outside 0..118, or an entry the table leaves 0, has no mass.
inline double ElemMass(int z) {
  static const double kMass[119] = {
      5.4859e-4, 1.00794, 4.002602, 6.941, 9.01218, 10.811, 12.011, 14.00674,
      15.9994, 18.9984303, 20.1797, 22.989768, 24.305, 26.981539, 28.0855, 30.973762,
      32.066, 35.4527, 39.948, 39.0983, 40.078, 44.95591, 47.88, 50.9415,
      51.9961, 54.93805, 55.847, 58.9332, 58.6934, 63.546, 65.39, 69.723,
      72.61, 74.92159, 78.96, 79.904, 83.8, 85.4678, 87.62, 88.90585,
      91.224, 92.90638, 95.94, 97.9072, 102.07, 102.9055, 106.42, 107.8682,
      112.411, 114.818, 118.71, 121.760, 127.6, 126.90447, 131.29, 132.90543,
      137.327, 138.9055, 140.115, 140.90765, 144.24, 144.9127, 150.36, 151.965,
      157.25, 158.92534, 162.50, 164.93032, 167.26, 168.93421, 173.04, 174.967,
      178.49, 180.9479, 183.84, 186.207, 190.23, 192.22, 195.08, 196.96654,
      200.59, 204.3833, 207.2, 208.98037, 208.9824, 209.9871, 222.0176, 223.0197,
      226.0254, 227.0278, 232.0381, 231.03588, 238.0289, 237.048, 244.0642, 243.0614,
      247.0703, 247.0703, 251.0796, 252.083, 257.0951, 258.1, 259.1009, 262.11,
      261, 262, 266, 264, 269, 268, 269, 272,
      277, 0, 289, 0, 0, 0, 0,
  };
  return (z >= 0 && z < 119) ? kMass[z] : -1.0;
}

inline long blocks_for(long threads) { return (threads + kBlock - 1) / kBlock; }

// The 48-bit linear congruential generator behind drand48 as POSIX specifies
// it: X' = (0x5DEECE66D X + 0xB) mod 2^48, drand48() = X' / 2^48;
// srand48(s) sets X = (low 32 bits of s) * 2^16 + 0x330E; a program that
// never seeds starts from X = 0x1234ABCD330E.
struct Drand48 {
  uint64_t x = 0x1234ABCD330EULL;
  void seed(long s) { x = ((uint64_t)(uint32_t)s << 16) | 0x330EULL; }
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    return std::ldexp((double)x, -48);
  }
};

// SimulationEngine::InitVelocities (liblpmd/simulation/simulationbuilder.cc:167-185):
// each atom gets a random direction (RandomVector, liblpmd/lpmd/vector.h:161-167:
// three 2 drand48() - 1 draws scaled to length sqrt(2 K / m), K = (3/2) k_B
// 1 K with k_B the unit system's 8.6173422e-5 eV/K), then the mean momentum
// per atom is taken out, v -= (sum m v / N) / m, in the reference's order of
// operations.  v[3n] in caller order.
inline void init_velocities(long n, const double* mass, Drand48& rng, double* v) {
  const double kboltzmann = 8.6173422e-05; // liblpmd/runtime/units.cc:7
  const double targettemp = 1.0;
  const double Kpart = (3.0 / 2.0) * kboltzmann * targettemp;
  double tp[3] = {0.0, 0.0, 0.0};
  for (long i = 0; i < n; ++i) {
    const double v2 = 2.0 * Kpart / mass[i];
    const double m = std::sqrt(v2);
    double w[3];
    for (int q = 0; q < 3; ++q)
      w[q] = 2.0 * rng.next() - 1.0;
    const double r = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int q = 0; q < 3; ++q) {
      v[3 * i + q] = m * w[q] / r;
      tp[q] += v[3 * i + q] * mass[i];
    }
  }
  for (int q = 0; q < 3; ++q)
    tp[q] = tp[q] / (double)n;
  for (long i = 0; i < n; ++i)
    for (int q = 0; q < 3; ++q)
      v[3 * i + q] -= tp[q] / mass[i];
}

} // namespace lpmdx

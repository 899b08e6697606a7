// The per-pair functions of the drop-in potentials in FP64 on the host: the
// virtuals a caller reaches through lpmd::PairPotential (pairEnergy,
// pairForce, liblpmd/lpmd/pairpotential.h:23-25) and lpmd::MetalPotential
// (pairEnergy, rhoij, F, PairForce, ManyBodies, UpdateCorrections,
// liblpmd/lpmd/metalpotential.h:26-37).  Each form follows its plugin's
// formula and order of operations (plugins/<name>.cc, cited per case), so a
// caller that evaluates single pairs on the host (plugins/localpressure.cc,
// plugins/atomenergy.cc) gets the CPU plugin's numbers.  The device kernels
// do not use these: forces come from the engine.
#pragma once

#include <lpmd/vector.h>

#include <cmath>
#include <string>
#include <vector>

namespace lpmdx_lpmd {

struct PairForm {
  std::string kind;
  // keyword values in the plugin's keyword order (see MakeForm)
  double p[10] = {0};
  double cutoff = HUGE_VAL;
  // fastlj (Tabulate, fastlj.cc:64-80) and tabulatedpair (the file's rows)
  long bins = 0;
  std::vector<double> etable, fftable;
  double tcut = 0.0; // the table's cutoff in k = floor(bins r / cutoff)
  // suttonchen: a^n and a^m (suttonchen.cc:29-30)
  double an = 0.0, am = 0.0;

  // the table entry of r; the reference reads past the table when r rounds
  // up to the cutoff, the entry is clamped here as in the engine
  long bin(double r) const {
    long k = long(floor(bins * r / tcut));
    return k > bins - 1 ? bins - 1 : k;
  }

  // ---- pair potentials
  double pairEnergy(double r) const {
    if (kind == "lennardjones") { // lennardjones.cc:53-58
      double rtmp = p[0] / r;
      double r6 = rtmp * rtmp * rtmp * rtmp * rtmp * rtmp;
      double r12 = r6 * r6;
      return 4.0e0 * p[1] * (r12 - r6);
    }
    if (kind == "fastlj" || kind == "tabulatedpair") // fastlj.cc:81-85, tabulatedpair.cc:81-85
      return etable[bin(r)];
    if (kind == "morse") // morse.cc:53-55
      return p[0] * (1.0 - exp(-p[1] * (r - p[2]))) * (1.0 - exp(-p[1] * (r - p[2])));
    if (kind == "buckingham") { // buckingham.cc:58-63
      double T1 = p[0] * exp(-r / p[2]);
      double r6 = r * r * r * r * r * r;
      double T2 = p[1] / r6;
      return T1 - T2;
    }
    if (kind == "harmonic") // harmonic.cc:50-52
      return 0.5e0 * p[0] * (fabs(r) - p[1]) * (fabs(r) - p[1]);
    if (kind == "lennardjonesMod") { // lennardjonesMod.cc:65-70
      double rtmp = p[0] / r;
      double rn = pow(rtmp, p[3]);
      double rm = pow(rtmp, p[2]);
      return 4.0e0 * p[1] * (rm - p[4] * rn);
    }
    // ---- metals
    if (kind == "suttonchen") { // suttonchen.cc:73-76
      double ir = 1 / r;
      return p[0] * an * pow(ir, p[2]);
    }
    if (kind == "finnissinclair") { // finnissinclair.cc:75-80
      if (r <= p[5])
        return (r - p[5]) * (r - p[5]) * (p[0] + p[1] * r + p[2] * r * r);
      return 0.0e0;
    }
    if (kind == "gupta") // gupta.cc:67
      return p[0] * exp(-p[2] * (r - p[1]) / p[1]);
    if (kind == "finnissinclair-ext") { // finnissinclair-ext.cc:85-91
      if (r <= p[7]) {
        double r2 = r * r;
        return (r - p[7]) * (r - p[7]) * (p[0] + p[1] * r + p[2] * r2 + p[3] * r2 * r + p[4] * r2 * r2);
      }
      return 0.0e0;
    }
    return NAN;
  }

  lpmd::Vector pairForce(const lpmd::Vector& r) const {
    double ff = NAN;
    if (kind == "lennardjones") { // lennardjones.cc:60-67
      double rr2 = r.SquareModule();
      double r6 = pow(p[0] * p[0] / rr2, 3.0e0);
      double r12 = r6 * r6;
      ff = -48.0e0 * (p[1] / rr2) * (r12 - 0.50e0 * r6);
    } else if (kind == "fastlj" || kind == "tabulatedpair") { // fastlj.cc:87-95, tabulatedpair.cc:87-96
      double rr2 = r.SquareModule();
      double rr = sqrt(rr2);
      ff = fftable[bin(rr)];
    } else if (kind == "morse") { // morse.cc:57-62
      double rr = sqrt(r.SquareModule());
      ff = 2.0 * p[1] * (p[0] / rr) * (1.0 - exp(-p[1] * (rr - p[2]))) * exp(-p[1] * (rr - p[2]));
    } else if (kind == "buckingham") { // buckingham.cc:65-75
      double rr2 = r.SquareModule();
      double rr = sqrt(rr2);
      double r8 = rr * rr * rr * rr * rr * rr * rr * rr;
      double f6 = 6 * p[1] / r8;
      double pe = p[0] * exp(-rr / p[2]) / (p[2] * rr);
      ff = f6 - pe;
    } else if (kind == "harmonic") { // harmonic.cc:54-60
      double rr = r.Module();
      ff = p[0] * (rr - p[1]) / rr;
    } else if (kind == "lennardjonesMod") { // lennardjonesMod.cc:72-79
      double rr2 = r.SquareModule();
      double rn = pow(p[0] * p[0] / rr2, p[3] / 2.0);
      double rm = pow(p[0] * p[0] / rr2, p[2] / 2.0);
      ff = -4.0e0 * (p[1] / rr2) * (p[2] * rm - p[4] * p[3] * rn);
    }
    return r * ff;
  }

  // ---- metals
  double rhoij(double r) const {
    if (kind == "suttonchen") { // suttonchen.cc:78-81
      double ir = 1 / r;
      return am * pow(ir, p[3]);
    }
    if (kind == "finnissinclair") { // finnissinclair.cc:82-87
      if (r <= p[6])
        return ((r - p[6]) * (r - p[6])) * (1.0e0 + p[4] * (r - p[6]) / p[6]);
      return 0.0e0;
    }
    if (kind == "gupta") // gupta.cc:69
      return exp(-2.0e0 * p[4] * (r - p[1]) / p[1]);
    if (kind == "finnissinclair-ext") { // finnissinclair-ext.cc:93-99
      if (r <= p[8]) {
        double rd = r - p[8];
        return (rd * rd) * (1.0e0 + p[6] * p[6] * rd * rd);
      }
      return 0.0e0;
    }
    return NAN;
  }

  double F(double rho) const {
    if (kind == "suttonchen") // suttonchen.cc:83
      return -p[4] * p[0] * sqrt(rho);
    if (kind == "finnissinclair") // finnissinclair.cc:89
      return -p[3] * sqrt(rho);
    if (kind == "gupta") // gupta.cc:71
      return -p[3] * sqrt(rho);
    if (kind == "finnissinclair-ext") // finnissinclair-ext.cc:101
      return -p[5] * sqrt(rho);
    return NAN;
  }

  lpmd::Vector PairForce(const lpmd::Vector& normrij, double r) const {
    if (kind == "suttonchen") { // suttonchen.cc:85-88
      double ir = 1.0e0 / r;
      return -p[2] * p[0] * an * pow(ir, p[2] + 1) * normrij;
    }
    if (kind == "finnissinclair") { // finnissinclair.cc:91-98
      if (r <= p[5]) {
        double t1 = 2.0e0 * (r - p[5]) * (p[0] + p[1] * r + p[2] * r * r);
        double t2 = (r - p[5]) * (r - p[5]) * (p[1] + 2.0e0 * p[2] * r);
        return (t1 + t2) * normrij;
      }
      return lpmd::Vector(0.0, 0.0, 0.0);
    }
    if (kind == "gupta") // gupta.cc:73-75
      return -(p[0] * p[2] / p[1]) * exp(-p[2] * (r - p[1]) / p[1]) * normrij;
    if (kind == "finnissinclair-ext") { // finnissinclair-ext.cc:103-112
      if (r <= p[7]) {
        double r2 = r * r;
        double rc = r - p[7];
        double t1 = 2.0e0 * rc * (p[0] + p[1] * r + p[2] * r2 + p[3] * r2 * r + p[4] * r2 * r2);
        double t2 = (p[1] + 2.0e0 * p[2] * r + 3.0e0 * p[3] * r2 + 4.0e0 * p[4] * r2 * r) * rc * rc;
        return (t1 + t2) * normrij;
      }
      return lpmd::Vector(0.0e0, 0.0e0, 0.0e0);
    }
    return lpmd::Vector(NAN, NAN, NAN);
  }

  lpmd::Vector ManyBodies(const lpmd::Vector& normrij, double rhoi, double rhoj, double r) const {
    if (kind == "suttonchen") { // suttonchen.cc:90-95
      double ir = 1.0e0 / r;
      double tmp = 0.5 * p[3] * p[4] * p[0] * ((1 / sqrt(rhoi)) + 1 / (sqrt(rhoj))) * am * pow(ir, p[3] + 1);
      return tmp * normrij;
    }
    if (kind == "finnissinclair") { // finnissinclair.cc:100-108
      if (r <= p[6]) {
        double t1 = ((1.0e0 / sqrt(rhoi)) + (1.0e0 / sqrt(rhoj))) * p[3] / 2.0e0;
        double t2 = 2.0e0 * (r - p[6]) + 3.0e0 * p[4] * (r - p[6]) * (r - p[6]) / p[6];
        return -t1 * t2 * normrij;
      }
      return lpmd::Vector(0.0e0, 0.0e0, 0.0e0);
    }
    if (kind == "gupta") { // gupta.cc:77-82
      double tmp = (p[3] * p[4] / p[1]) * ((1.0e0 / sqrt(rhoi)) + (1.0e0 / sqrt(rhoj))) *
                   exp(-2.0e0 * p[4] * (r - p[1]) / p[1]);
      return tmp * normrij;
    }
    if (kind == "finnissinclair-ext") { // finnissinclair-ext.cc:114-123
      if (r <= p[8]) {
        double rd = r - p[8];
        double t1 = p[5] * 0.5e0 * ((1.0e0 / sqrt(rhoi)) + (1.0e0 / sqrt(rhoj)));
        double t2 = 2.0e0 * rd + 4.0e0 * p[6] * p[6] * rd * rd * rd;
        return -t1 * t2 * normrij;
      }
      return lpmd::Vector(0.0e0, 0.0e0, 0.0e0);
    }
    return lpmd::Vector(NAN, NAN, NAN);
  }

  // MetalPotential::UpdateCorrections: the long-range corrections (drho, du,
  // dvir) of suttonchen (suttonchen.cc:97-121, without its diagnostic prints);
  // zero for the others (finnissinclair.cc:110-113, gupta.cc:84-86,
  // finnissinclair-ext.cc:125-128)
  lpmd::Vector UpdateCorrections(double rho, int N, double sinv) const {
    if (kind != "suttonchen")
      return lpmd::Vector(0, 0, 0);
    const double e = p[0], a = p[1], n = p[2], m = p[3], c = p[4], rcut = p[5];
    double drho = 0.0e0, du = 0.0e0, dvir = 0.0e0;
    double fac = -c * e * 0.5e0 * sinv;
    drho = (4 * M_PI * rho * rcut / (m - 1)) * pow((a / (rcut)), m);
    int pot = n - 3;
    double aor = a / rcut;
    du = (2 * M_PI * N * rho * e / (n - 3)) * a * a * a * pow(aor, pot);
    double v1 = -2.0e0 * M_PI * rho * N * e * n * a * a * a * pow(aor, pot) / (n - 3);
    pot = m - 3;
    double v2 = -4.0e0 * M_PI * fac * rho * m * a * a * a * pow(aor, pot) / (m - 3);
    dvir = v1 + v2;
    return lpmd::Vector(drho, du, dvir);
  }
};

// The form of a plugin from its keyword values, read through `value(key)` in
// each plugin's keyword order (plugins/<name>.cc constructors); the table
// kinds also get their tables (fastlj's Tabulate, or the rows of a
// tabulatedpair file).
template <typename Get>
PairForm MakeForm(const std::string& kind, Get value, const std::vector<double>* te = nullptr,
                  const std::vector<double>* tf = nullptr, double tcut = 0.0) {
  PairForm F;
  F.kind = kind;
  auto put = [&](std::initializer_list<const char*> keys) {
    int q = 0;
    for (const char* k : keys)
      F.p[q++] = value(k);
  };
  if (kind == "lennardjones") {
    put({"sigma", "epsilon", "cutoff"});
    F.cutoff = F.p[2];
  } else if (kind == "fastlj") {
    put({"sigma", "epsilon", "cutoff", "bins"});
    F.cutoff = F.p[2];
    F.bins = long(F.p[3]);
    F.tcut = F.p[2];
    const double sigma = F.p[0], epsilon = F.p[1], cutoff = F.p[2];
    F.etable.resize(F.bins > 0 ? F.bins : 0);
    F.fftable.resize(F.bins > 0 ? F.bins : 0);
    for (long i = 0; i < F.bins; ++i) { // fastlj.cc:64-80
      double r = cutoff * double(i) / double(F.bins);
      double r6 = pow(sigma / r, 6.0e0);
      double r12 = r6 * r6;
      F.etable[i] = 4.0e0 * epsilon * (r12 - r6);
      F.fftable[i] = -48.0e0 * (epsilon / (r * r)) * (r12 - 0.50e0 * r6);
    }
  } else if (kind == "tabulatedpair") {
    F.etable = *te;
    F.fftable = *tf;
    F.bins = long(te->size());
    F.tcut = tcut;
    F.cutoff = tcut;
  } else if (kind == "morse") {
    put({"depth", "a", "re"});
  } else if (kind == "buckingham") {
    put({"B1", "B2", "Ro", "cutoff"});
    F.cutoff = F.p[3];
  } else if (kind == "harmonic") {
    put({"k", "a"});
  } else if (kind == "lennardjonesMod") {
    put({"sigma", "epsilon", "m", "n", "cohesive", "cutoff"});
    F.cutoff = F.p[5];
  } else if (kind == "suttonchen") {
    put({"e", "a", "n", "m", "c", "cutoff"});
    F.an = pow(F.p[1], F.p[2]);
    F.am = pow(F.p[1], F.p[3]);
    F.cutoff = F.p[5];
  } else if (kind == "finnissinclair") {
    put({"c0", "c1", "c2", "A", "B", "c", "d"});
    F.cutoff = F.p[6] >= F.p[5] ? F.p[6] : F.p[5];
  } else if (kind == "gupta") {
    put({"A", "r0", "p", "B", "q", "cutoff"});
    F.cutoff = F.p[5];
  } else if (kind == "finnissinclair-ext") {
    put({"c0", "c1", "c2", "c3", "c4", "A", "B", "c", "d"});
    F.cutoff = F.p[8] >= F.p[7] ? F.p[8] : F.p[7];
  }
  return F;
}

} // namespace lpmdx_lpmd

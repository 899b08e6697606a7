// LPMD plugins that run the reference's force path on the MI355X engine.
//
// These classes implement the reference's plugin interfaces and forward to the
// engine's C ABI (include/lpmdx.h):
//   DeviceCellManager     lpmd::CellManager (liblpmd/lpmd/cellmanager.h:15-27) as
//                         "linkedcell" (plugins/linkedcell.cc:17-41),
//                         "verletlist" (plugins/verletlist.cc:12-29),
//                         "minimumimage" (plugins/minimumimage.cc:12-19) and
//                         "lcbinary" (plugins/lcbinary.cc:9-33)
//   DevicePotential       lpmd::Potential (liblpmd/lpmd/potential.h:11-36) as
//                         lennardjones fastlj morse buckingham suttonchen
//                         finnissinclair gupta harmonic lennardjonesMod
//                         finnissinclair-ext tabulatedpair (plugins/<name>.cc)
//   DeviceVelocityVerlet  lpmd::Integrator (liblpmd/lpmd/integrator.h:15-40) as
//                         "velocityverlet" (plugins/velocityverlet.cc:13-20)
//   DeviceLeapfrog        lpmd::Integrator as "leapfrog" (plugins/leapfrog.cc:14-23)
//   DeviceBeeman          lpmd::Integrator as "beeman" (plugins/beeman.cc)
// They take the same plugin names and keyword strings, so a directory holding
// their <name>.so files placed first on LPMD_PATH replaces the CPU plugins.
// The classes live in one shared library (liblpmdx_lpmd.so) so that every
// plugin object sees one type identity: a potential finds the device copy of
// its Configuration through its DeviceCellManager.
#pragma once

#include <lpmd/cellmanager.h>
#include <lpmd/configuration.h>
#include <lpmd/integrator.h>
#include <lpmd/plugin.h>
#include <lpmd/potential.h>

#include <map>
#include <string>
#include <vector>

#include "lpmdx.h"

namespace lpmdx_lpmd {

class DevicePotential;

// The device copy of one Configuration lives in its cell manager: the
// reference attaches one CellManager per Configuration
// (Configuration::SetCellManager, liblpmd/lpmd/configuration.h:54-57).
class DeviceCellManager : public lpmd::CellManager, public lpmd::Plugin {
public:
  DeviceCellManager(const std::string& name, std::string args);
  ~DeviceCellManager() override;
  void ShowHelp() const override;

  double Cutoff() const override;
  void Reset() override;
  void UpdateCell(lpmd::Configuration& conf) override;
  void UpdateAtom(lpmd::Configuration& conf, long i) override;
  void BuildNeighborList(lpmd::Configuration& conf, long i, lpmd::NeighborList& nlist, bool full,
                         double rcut) override;

  // The device system of conf, created on first use; upload=true sends the
  // host positions.
  lpmdx_system* Bind(lpmd::Configuration& conf, bool upload);
  // Device index of a potential of this configuration (appended on first use).
  int IndexOf(const DevicePotential& p);
  // Device potentials in exactly this order (re-appended when they differ).
  void UsePotentials(const std::vector<const DevicePotential*>& pots);
  lpmdx_system* System() const { return sys_; }
  // Records x as the positions the device already holds (after a device step).
  void Mirror(const std::vector<double>& x) {
    xbuf_ = x;
    lists_ok_ = false;
  }

private:
  void Create(lpmd::Configuration& conf);
  std::string kind_, args_;
  double cutoff_ = 0.0;
  lpmdx_system* sys_ = nullptr;
  long n_ = -1;
  double box_[3] = {0, 0, 0};
  std::vector<const DevicePotential*> order_;
  std::vector<double> xbuf_;
  // neighbour lists exported from the device for BuildNeighborList callers
  bool lists_ok_ = false;
  bool lists_full_ = true;
  double lists_rcut_ = 0.0;
  std::vector<long> counts_, offs_, js_;
  std::vector<double> r2_;
};

class DevicePotential : public lpmd::Potential, public lpmd::Plugin {
public:
  DevicePotential(const std::string& name, std::string args);
  void ShowHelp() const override;

  // Potential::UpdateForces of this one potential: adds its accelerations,
  // virial and stress to conf, as PairPotential::UpdateForces
  // (liblpmd/simulation/pairpotential.cc:35-84) and
  // MetalPotential::UpdateForces (metalpotential.cc:62-180) do.
  void UpdateForces(lpmd::Configuration& conf) override;
  double energy(lpmd::Configuration& conf) override;
  double AtomEnergy(lpmd::Configuration& conf, long i) override;

  const std::string& Kind() const { return kind_; }
  const std::string& Args() const { return args_; }
  // Species rule as (s1, s2) for the engine, read back through AppliesTo
  // (potential.cc:23-27: spc_sum < 0 or s1 + s2 == spc_sum).
  void Species(int& s1, int& s2) const;

private:
  std::string kind_, args_;
  std::vector<double> acc_;
};

class DeviceVelocityVerlet : public lpmd::Integrator, public lpmd::Plugin {
public:
  explicit DeviceVelocityVerlet(std::string args);
  void ShowHelp() const override;
  void Initialize(lpmd::Simulation& sim, lpmd::Potential& p) override;
  // One velocity-Verlet step.  When every potential is a DevicePotential and
  // the cell manager a DeviceCellManager the step runs on the device (one
  // force evaluation, state kept in HBM, x, v, a copied back for the host's
  // tags and monitors); otherwise the reference's two-step scheme runs on the
  // host (twostepintegrator.cc:12-68, velocityverlet.cc:55-75).
  void Advance(lpmd::Simulation& sim, lpmd::Potential& p) override;

private:
  void HostStep(lpmd::Simulation& sim, lpmd::Potential& p);
  long start_ = 1;
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* last_ = nullptr;
};

class DeviceLeapfrog : public lpmd::Integrator, public lpmd::Plugin {
public:
  explicit DeviceLeapfrog(std::string args);
  void ShowHelp() const override;
  // Leapfrog::Initialize (leapfrog.cc:58-67): half-step velocities from the
  // forces at the old configuration, once per integrator object.  On the
  // device (lpmdx_leapfrog_init) under the same condition as
  // DeviceVelocityVerlet; the half-step velocities then stay in HBM.
  void Initialize(lpmd::Simulation& sim, lpmd::Potential& p) override;
  // One leapfrog step (onestepintegrator.cc:13-46, leapfrog.cc:69-77), on the
  // device or, when Initialize ran on the host, on the host.
  void Advance(lpmd::Simulation& sim, lpmd::Potential& p) override;

private:
  enum Mode { kUnset, kDevice, kHost };
  Mode mode_ = kUnset;
  long start_ = 1;
  std::vector<lpmd::Vector> vh_; // host path: the old configuration's velocities
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* sys_ = nullptr;
};

class DeviceBeeman : public lpmd::Integrator, public lpmd::Plugin {
public:
  explicit DeviceBeeman(std::string args);
  void ShowHelp() const override;
  // Beeman::Initialize (beeman.cc:55-61): the previous accelerations from the
  // forces at the old configuration, once per integrator object; on the
  // device (lpmdx_beeman_init) under the same condition as
  // DeviceVelocityVerlet, the history then staying in HBM.
  void Initialize(lpmd::Simulation& sim, lpmd::Potential& p) override;
  // One Beeman step (twostepintegrator.cc:12-68, beeman.cc:63-84), on the
  // device or, when Initialize ran on the host, on the host.
  void Advance(lpmd::Simulation& sim, lpmd::Potential& p) override;

private:
  enum Mode { kUnset, kDevice, kHost };
  Mode mode_ = kUnset;
  long start_ = 1;
  std::vector<lpmd::Vector> a_old_; // host path: the old configuration's accelerations
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* sys_ = nullptr;
};

// Plugin factory behind every <name>.so (create/destroy, liblpmd/lpmd/plugin.h:42-46).
lpmd::Plugin* Create(const std::string& name, const std::string& args);

} // namespace lpmdx_lpmd

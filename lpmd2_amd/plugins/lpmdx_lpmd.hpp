// LPMD plugins that run the reference's force path on the MI355X engine.
//
// These classes implement the reference's plugin interfaces and forward to the
// engine's C ABI (include/lpmdx.h):
//   DeviceCellManager     lpmd::CellManager (liblpmd/lpmd/cellmanager.h:15-27) as
//                         "linkedcell" (plugins/linkedcell.cc:17-41),
//                         "verletlist" (plugins/verletlist.cc:12-29),
//                         "minimumimage" (plugins/minimumimage.cc:12-19) and
//                         "lcbinary" (plugins/lcbinary.cc:9-33)
//   DevicePairPotential   lpmd::PairPotential (liblpmd/lpmd/pairpotential.h:15-26)
//                         as lennardjones fastlj morse buckingham harmonic
//                         lennardjonesMod tabulatedpair (plugins/<name>.cc)
//   DeviceMetalPotential  lpmd::MetalPotential (liblpmd/lpmd/metalpotential.h:13-41)
//                         as suttonchen finnissinclair gupta finnissinclair-ext
//   DeviceVelocityVerlet  lpmd::Integrator (liblpmd/lpmd/integrator.h:15-40) as
//                         "velocityverlet" (plugins/velocityverlet.cc:13-20)
//   DeviceLeapfrog        lpmd::Integrator as "leapfrog" (plugins/leapfrog.cc:14-23)
//   DeviceBeeman          lpmd::Integrator as "beeman" (plugins/beeman.cc)
//   DeviceNoseHoover      lpmd::Integrator as "nosehoover" (plugins/nosehoover.cc)
//   DeviceSimulation      lpmd::Simulation (liblpmd/lpmd/simulation.h:19-46) whose
//                         steps stay on the device, built by DeviceSimulationBuilder
//                         with SimulationBuilder's signatures (simulationbuilder.h:16-29)
// They take the same plugin names and keyword strings, so a directory holding
// their <name>.so files placed first on LPMD_PATH replaces the CPU plugins.
// The classes live in one shared library (liblpmdx_lpmd.so) so that every
// plugin object sees one type identity: a potential finds the device copy of
// its Configuration through its DeviceCellManager.
#pragma once

#include <lpmd/cellmanager.h>
#include <lpmd/combinedpotential.h>
#include <lpmd/configuration.h>
#include <lpmd/fixedsizeparticleset.h>
#include <lpmd/integrator.h>
#include <lpmd/metalpotential.h>
#include <lpmd/orthogonalcell.h>
#include <lpmd/pairpotential.h>
#include <lpmd/plugin.h>
#include <lpmd/potential.h>
#include <lpmd/simulation.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "lpmdx.h"
#include "pair_forms.hpp"

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
    ++version_;
  }
  // Changes whenever the device positions, the system or its potentials
  // change (what per-atom results computed before no longer describe).
  long Version() const { return version_; }
  // The fixed-atom flags of conf's fixedvel / fixedpos tags sent to the
  // device (only when they differ from the last ones sent).
  void SyncFixed(lpmd::Configuration& conf);

private:
  void Create(lpmd::Configuration& conf);
  std::string kind_, args_;
  double cutoff_ = 0.0;
  lpmdx_system* sys_ = nullptr;
  long n_ = -1;
  double box_[3] = {0, 0, 0};
  std::vector<const DevicePotential*> order_;
  std::vector<double> xbuf_;
  long version_ = 0;
  std::vector<int> fixed_; // flags last sent (empty: none)
  // neighbour lists exported from the device for BuildNeighborList callers
  bool lists_ok_ = false;
  bool lists_full_ = true;
  double lists_rcut_ = 0.0;
  std::vector<long> counts_, offs_, js_;
  std::vector<double> r2_;
};

// What the pair and EAM drop-ins share: the device term of the potential, its
// forces and per-atom energies through the cell manager's device system.
class DevicePotential {
public:
  virtual ~DevicePotential() = default;
  const std::string& Kind() const { return kind_; }
  const std::string& Args() const { return args_; }
  // Species rule as (s1, s2) for the engine, read back through AppliesTo
  // (potential.cc:23-27: spc_sum < 0 or s1 + s2 == spc_sum).
  void Species(int& s1, int& s2) const;
  virtual const lpmd::Potential& Pot() const = 0;

protected:
  DevicePotential(const std::string& kind, const std::string& args) : kind_(kind), args_(args) {}
  // Potential::UpdateForces of this one potential: adds its accelerations,
  // virial and stress to conf, as PairPotential::UpdateForces
  // (liblpmd/simulation/pairpotential.cc:35-84) and
  // MetalPotential::UpdateForces (metalpotential.cc:62-180) do; returns its energy.
  double DeviceForces(lpmd::Configuration& conf);
  // Potential::AtomEnergy (pairpotential.cc:21-33, metalpotential.cc:47-60)
  // from the device's per-atom energies, kept until the positions change.
  double DeviceAtomEnergy(lpmd::Configuration& conf, long i);
  std::string kind_, args_;
  std::vector<double> acc_, eatom_;
  const void* eatom_sys_ = nullptr;
  long eatom_version_ = -1;
};

// lpmd::PairPotential drop-in: lennardjones fastlj morse buckingham harmonic
// lennardjonesMod tabulatedpair.  UpdateForces, energy and AtomEnergy run on
// the device; pairEnergy / pairForce (pairpotential.h:23-25) are the
// plugin's FP64 forms on the host, for callers that evaluate single pairs.
class DevicePairPotential : public lpmd::PairPotential, public lpmd::Plugin, public DevicePotential {
public:
  DevicePairPotential(const std::string& name, std::string args);
  void ShowHelp() const override;
  void UpdateForces(lpmd::Configuration& conf) override { energycache = DeviceForces(conf); }
  double energy(lpmd::Configuration&) override { return energycache; }
  double AtomEnergy(lpmd::Configuration& conf, long i) override { return DeviceAtomEnergy(conf, i); }
  double pairEnergy(const double& r) const override { return form_.pairEnergy(r); }
  lpmd::Vector pairForce(const lpmd::Vector& r) const override { return form_.pairForce(r); }
  const lpmd::Potential& Pot() const override { return *this; }

private:
  PairForm form_;
};

// lpmd::MetalPotential drop-in: suttonchen finnissinclair gupta
// finnissinclair-ext.  UpdateForces, energy and AtomEnergy run on the
// device; pairEnergy, rhoij, F, PairForce, ManyBodies and UpdateCorrections
// (metalpotential.h:26-37) are the plugin's FP64 forms on the host.
class DeviceMetalPotential : public lpmd::MetalPotential, public lpmd::Plugin, public DevicePotential {
public:
  DeviceMetalPotential(const std::string& name, std::string args);
  void ShowHelp() const override;
  void UpdateForces(lpmd::Configuration& conf) override { energycache = DeviceForces(conf); }
  double energy(lpmd::Configuration&) override { return energycache; }
  double AtomEnergy(lpmd::Configuration& conf, long i) override { return DeviceAtomEnergy(conf, i); }
  double pairEnergy(const double& r) const override { return form_.pairEnergy(r); }
  double rhoij(const double& r) const override { return form_.rhoij(r); }
  double F(const double& rho) const override { return form_.F(rho); }
  lpmd::Vector PairForce(const lpmd::Vector& normr, const double& mod) const override {
    return form_.PairForce(normr, mod);
  }
  lpmd::Vector ManyBodies(const lpmd::Vector& normrij, const double& rhoi, const double& rhoj,
                          const double& mod) const override {
    return form_.ManyBodies(normrij, rhoi, rhoj, mod);
  }
  lpmd::Vector UpdateCorrections(const double& rho, const int& N, const double& sinv) const override {
    return form_.UpdateCorrections(rho, N, sinv);
  }
  const lpmd::Potential& Pot() const override { return *this; }

private:
  PairForm form_;
};

// The device runs the steps when every potential is a DevicePotential and the
// cell manager a DeviceCellManager; returns that cell manager (and the
// potentials in dev), else null.
DeviceCellManager* DevicePath(lpmd::Simulation& sim, lpmd::Potential& p, std::vector<const DevicePotential*>& dev);

// What the drop-in integrators share with DeviceSimulation: several steps on
// the device in one call, with nothing copied back.
class DeviceIntegrator {
public:
  virtual ~DeviceIntegrator() = default;
  // Runs `steps` steps of the scheme on cm's device system (DevicePath holds;
  // the schemes with an old configuration have run Initialize on the
  // device).  upload: the host holds the current state, which is sent first
  // as Advance sends a state the host changed; otherwise the device holds it.
  // Returns the potential energy of the last step's force evaluation.
  virtual double DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm, const std::vector<const DevicePotential*>& dev,
                           long steps, bool upload) = 0;
  // Whether the scheme can run on the device for this simulation now (its
  // Initialize ran on the device, or it needs none).
  virtual bool DeviceReady() const = 0;
};

class DeviceVelocityVerlet : public lpmd::Integrator, public lpmd::Plugin, public DeviceIntegrator {
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
  double DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm, const std::vector<const DevicePotential*>& dev,
                   long steps, bool upload) override;
  bool DeviceReady() const override { return true; }

private:
  void HostStep(lpmd::Simulation& sim, lpmd::Potential& p);
  long start_ = 1;
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* last_ = nullptr;
};

class DeviceLeapfrog : public lpmd::Integrator, public lpmd::Plugin, public DeviceIntegrator {
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
  double DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm, const std::vector<const DevicePotential*>& dev,
                   long steps, bool upload) override;
  bool DeviceReady() const override { return mode_ == kDevice; }

private:
  enum Mode { kUnset, kDevice, kHost };
  Mode mode_ = kUnset;
  long start_ = 1;
  std::vector<lpmd::Vector> vh_; // host path: the old configuration's velocities
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* sys_ = nullptr;
};

class DeviceBeeman : public lpmd::Integrator, public lpmd::Plugin, public DeviceIntegrator {
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
  double DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm, const std::vector<const DevicePotential*>& dev,
                   long steps, bool upload) override;
  bool DeviceReady() const override { return mode_ == kDevice; }

private:
  enum Mode { kUnset, kDevice, kHost };
  Mode mode_ = kUnset;
  long start_ = 1;
  std::vector<lpmd::Vector> a_old_; // host path: the old configuration's accelerations
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* sys_ = nullptr;
};

class DeviceNoseHoover : public lpmd::Integrator, public lpmd::Plugin, public DeviceIntegrator {
public:
  explicit DeviceNoseHoover(std::string args);
  void ShowHelp() const override;
  // NoseHoover::Initialize (nosehoover.cc:62-68): the forces at the old
  // configuration give the first old accelerations, the friction starts at
  // zero; once per integrator object, on the device (lpmdx_nosehoover_init,
  // with the session's kboltzmann) under the same condition as
  // DeviceVelocityVerlet, the history and the friction then staying in HBM.
  void Initialize(lpmd::Simulation& sim, lpmd::Potential& p) override;
  // One Nose-Hoover step (twostepintegrator.cc:12-68, nosehoover.cc:70-96),
  // on the device or, when Initialize ran on the host, on the host.
  void Advance(lpmd::Simulation& sim, lpmd::Potential& p) override;
  double DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm, const std::vector<const DevicePotential*>& dev,
                   long steps, bool upload) override;
  bool DeviceReady() const override { return mode_ == kDevice; }

private:
  enum Mode { kUnset, kDevice, kHost };
  Mode mode_ = kUnset;
  long start_ = 1;
  double q_ = 0.0, temp_ = 300.0;
  // host path: the friction, old.a and NoseHoover's auxlist
  double friction_ = 0.0;
  std::vector<lpmd::Vector> a_old_, aux_;
  std::vector<double> hx_, hv_, ha_;
  lpmdx_system* sys_ = nullptr;
};

// lpmd::Simulation whose steps stay on the device: SimulationEngine
// (liblpmd/simulation/simulationbuilder.cc:23-293) over the reference's own
// FixedSizeParticleSet and OrthogonalCell, with DoStep / DoSteps split in
// two paths.
//  * Device path (the integrator is a DeviceIntegrator ready on the device,
//    every potential a DevicePotential, the cell manager a
//    DeviceCellManager): DoSteps(n) runs the n steps in one call of the
//    engine and sets the step's tags (SetSimulationTags,
//    simulationbuilder.cc:84-124) and the potential-energy tag, virial and
//    stress from the engine's device reductions.  Nothing is copied back:
//    the host atoms are refreshed from the device when they are next read
//    (OriginalAtoms, and so Atoms()), and a host state that changed since
//    the device last saw it (positions or velocities written through
//    Atoms()) is sent before the next device step.  References to atoms
//    taken before a DoSteps see the host copy of that time; read Atoms()
//    again after a DoSteps.
//  * Otherwise each step is SimulationEngine::DoStep on the host (Virial and
//    stress zeroed, Initialize once, Integrator::Advance, step + 1,
//    SetSimulationTags).
// Without a runtime context the session values the tags read (pressfactor,
// ua3togrcm3, kboltzmann) are those of the default unit system
// (runtime/units.cc:5-13 as runtime/session.cc:17-38 formats them).
class DeviceSimulation : public lpmd::Simulation {
public:
  DeviceSimulation(long natoms, const lpmd::BasicAtom& at);
  ~DeviceSimulation() override = default;

  lpmd::BasicParticleSet& OriginalAtoms() override;
  const lpmd::BasicParticleSet& OriginalAtoms() const override;
  // the cell handed out writable: the atoms are refreshed first, so a device
  // system made again for a changed cell starts from the current state
  lpmd::BasicCell& OriginalCell() override;
  const lpmd::BasicCell& OriginalCell() const override { return cell_; }

  void SetTemperature(double temp, bool tag) override;
  lpmd::CombinedPotential& Potentials() override { return potarray_; }
  const lpmd::CombinedPotential& Potentials() const override { return potarray_; }
  void DoStep() override { DoSteps(1); }
  void DoSteps(long n) override;
  long CurrentStep() const override { return step_; }
  void AdjustCurrentStep(long s) override;
  void SetIntegrator(lpmd::Integrator& itg) override { integ_ = &itg; }
  lpmd::Integrator& Integrator() override;
  void RescalePositions(const lpmd::BasicCell& old_cell) override;
  void Dump(const std::string& path) const override;
  void Restore(const std::string& path) override;

  // Counters: host refreshes from the device, device runs (one per DoSteps
  // on the device path), states sent to the device by a DoSteps, host steps.
  long HostRefreshes() const { return refreshes_; }
  long DeviceRuns() const { return runs_; }
  long Uploads() const { return uploads_; }
  long HostSteps() const { return host_steps_; }

private:
  void Initialize();
  void InitVelocities();
  void HostStep();
  void SetSimulationTags();
  void DeviceTags(lpmdx_system* sys, double epot);
  void Refresh() const;
  double Session(const char* key, double fallback) const;
  bool HostChanged() const;

  lpmd::FixedSizeParticleSet atoms_;
  lpmd::OrthogonalCell cell_;
  lpmd::CombinedPotential potarray_;
  lpmd::Integrator* integ_ = nullptr;
  bool velocities_set_ = false, initialized_ = false;
  long step_ = 0, initial_step_ = 0;
  // device view: the device holds a state newer than atoms_ (of dev_'s system)
  mutable bool device_ahead_ = false;
  // atoms_ were handed out writable since host and device last agreed
  mutable bool handed_out_ = true;
  // during a device run: OriginalAtoms hands out atoms_ as they are (their
  // tags, sizes and masses are current; positions are not read)
  bool raw_view_ = false;
  // Restore replaced the host state: the next device run sends it, whatever
  // the integrator's Initialize did
  bool restored_ = false;
  mutable DeviceCellManager* dev_ = nullptr;
  mutable std::vector<double> agreed_x_, agreed_v_; // the state host and device last agreed on
  mutable long refreshes_ = 0;
  long runs_ = 0, uploads_ = 0, host_steps_ = 0;
};

// SimulationBuilder (liblpmd/lpmd/simulationbuilder.h:16-29) for DeviceSimulation:
// the same creation calls; the builder owns the simulations it makes.
class DeviceSimulationBuilder {
public:
  // SimulationBuilder::CreateFixedOrthogonal (simulationbuilder.cc:321-326)
  static lpmd::Simulation& CreateFixedOrthogonal(lpmd::RuntimeContext& context, long atoms, const lpmd::BasicAtom& at);
  // the same without a runtime context (the default unit system's session values)
  static lpmd::Simulation& CreateFixedOrthogonal(long atoms, const lpmd::BasicAtom& at);
  // SimulationBuilder::CloneOptimized (simulationbuilder.cc:342-375) for an
  // orthogonal cell: atoms, their tags, the cell and the step tag copied
  static lpmd::Simulation& CloneOptimized(lpmd::RuntimeContext& context, const lpmd::Simulation& sim);
  // Frees a simulation this builder made (and its device memory, through
  // its cell manager's system when that is released).
  static void Destroy(lpmd::Simulation& sim);
};

// Plugin factory behind every <name>.so (create/destroy, liblpmd/lpmd/plugin.h:42-46).
lpmd::Plugin* Create(const std::string& name, const std::string& args);
// Create for the create() entry point: an error (a bad keyword string, an
// unknown name, a table file that does not open) ends the program with the
// message, as lpmd::EndWithError does (liblpmd/lpmd/util.h:19-22), instead of
// unwinding out of the dlopen'd plugin: the loader dlcloses the plugin while
// the exception is in flight, and its type information goes with it
// (liblpmd/plugins/plugin.cc:32-60).
lpmd::Plugin* CreateOrExit(const std::string& name, const std::string& args);

} // namespace lpmdx_lpmd

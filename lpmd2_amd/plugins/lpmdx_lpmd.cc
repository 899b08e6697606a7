// Implementation of the LPMD plugin classes over the engine's C ABI.
#include "lpmdx_lpmd.hpp"

#include "../csrc/table.hpp"

#include <lpmd/basicparticleset.h>
#include <lpmd/combinedpotential.h>
#include <lpmd/error.h>
#include <lpmd/properties.h>
#include <lpmd/simulation.h>
#include <runtime/runtime_context.h>

#include <lpmd/colorhandler.h>

#include <cmath>
#include <fstream>
#include <iostream>

namespace lpmdx_lpmd {

namespace {

void check(int rc, const std::string& what) {
  if (rc != LPMDX_OK)
    throw lpmd::RuntimeError(what + ": " + lpmdx_last_error());
}

struct Keyword {
  const char* name;
  const char* def; // nullptr: the reference defines it without a default
};

// Plugin keywords with the reference's names and defaults
// (plugins/<name>.cc constructors), so control files carry over unchanged;
// "skin" is this engine's verletlist extension.
const std::map<std::string, std::vector<Keyword>>& Keywords() {
  static const std::map<std::string, std::vector<Keyword>> m = {
      {"lennardjones", {{"sigma", nullptr}, {"epsilon", nullptr}, {"cutoff", nullptr}}},
      {"fastlj", {{"sigma", nullptr}, {"epsilon", nullptr}, {"cutoff", nullptr}, {"bins", "500"}}},
      {"morse", {{"depth", nullptr}, {"a", nullptr}, {"re", nullptr}}},
      {"buckingham", {{"B1", nullptr}, {"B2", nullptr}, {"Ro", nullptr}, {"cutoff", nullptr}}},
      {"suttonchen",
       {{"e", nullptr}, {"a", nullptr}, {"n", nullptr}, {"m", nullptr}, {"c", nullptr}, {"cutoff", nullptr},
        {"corrections", "false"}}},
      {"finnissinclair",
       {{"c0", nullptr}, {"c1", nullptr}, {"c2", nullptr}, {"A", nullptr}, {"B", nullptr}, {"c", nullptr},
        {"d", nullptr}}},
      {"gupta", {{"A", nullptr}, {"r0", nullptr}, {"p", nullptr}, {"B", nullptr}, {"q", nullptr}, {"cutoff", nullptr}}},
      {"harmonic", {{"k", nullptr}, {"a", nullptr}}},
      {"tabulatedpair", {{"file", ""}}},
      {"lennardjonesMod",
       {{"sigma", nullptr}, {"epsilon", nullptr}, {"m", "12"}, {"n", "6"}, {"cohesive", "1"}, {"cutoff", nullptr}}},
      {"finnissinclair-ext",
       {{"c0", nullptr}, {"c1", nullptr}, {"c2", nullptr}, {"c3", nullptr}, {"c4", nullptr}, {"A", nullptr},
        {"B", nullptr}, {"c", nullptr}, {"d", nullptr}}},
      {"linkedcell",
       {{"cutoff", "7.0"}, {"nx", "0"}, {"ny", "0"}, {"nz", "0"}, {"mode", "noauto"}, {"warn-outside", "true"}}},
      {"verletlist", {{"cutoff", "0.0"}, {"extcutoff", "0.0"}, {"maxneighbors", "80"}, {"each", "5"}, {"skin", "0"}}},
      {"minimumimage", {{"cutoff", "0.0"}, {"skin", "0"}}},
      {"lcbinary", {{"cutoff", "7.0"}, {"nx", "0"}, {"ny", "0"}, {"nz", "0"}, {"mode", "noauto"}}},
      {"velocityverlet", {{"dt", "1.0"}, {"start", "1"}}},
      {"leapfrog", {{"dt", "1.0"}, {"start", "1"}}},
      {"beeman", {{"dt", "1.0"}, {"start", "1"}}},
      {"nosehoover", {{"fmass", nullptr}, {"dt", "1.0"}, {"start", "1"}, {"t", "300.0"}}},
  };
  return m;
}

void Define(lpmd::Module& m, const std::string& name) {
  for (const Keyword& k : Keywords().at(name)) {
    if (k.def)
      m.DefineKeyword(k.name, k.def);
    else
      m.DefineKeyword(k.name);
  }
}

const char* kVersion = "2.1-mi355x";

void Positions(lpmd::Configuration& conf, std::vector<double>& x) {
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const long n = atoms.Size();
  x.resize(3 * n);
  for (long i = 0; i < n; ++i) {
    const lpmd::Vector& p = atoms[i].Position();
    x[3 * i] = p[0];
    x[3 * i + 1] = p[1];
    x[3 * i + 2] = p[2];
  }
}

void Velocities(lpmd::Configuration& conf, std::vector<double>& v) {
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const long n = atoms.Size();
  v.resize(3 * n);
  for (long i = 0; i < n; ++i) {
    const lpmd::Vector& p = atoms[i].Velocity();
    v[3 * i] = p[0];
    v[3 * i + 1] = p[1];
    v[3 * i + 2] = p[2];
  }
}

// Whether the host still holds the positions (and, with hv, the velocities)
// the last device step handed back: one pass over the atoms, nothing copied.
bool HostUnchanged(lpmd::Configuration& conf, const std::vector<double>& hx, const std::vector<double>* hv) {
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const long n = atoms.Size();
  if ((long)hx.size() != 3 * n || (hv && (long)hv->size() != 3 * n))
    return false;
  for (long i = 0; i < n; ++i) {
    const lpmd::BasicAtom& at = atoms[i];
    const lpmd::Vector& p = at.Position();
    if (p[0] != hx[3 * i] || p[1] != hx[3 * i + 1] || p[2] != hx[3 * i + 2])
      return false;
    if (hv) {
      const lpmd::Vector& u = at.Velocity();
      const double* w = hv->data() + 3 * i;
      if (u[0] != w[0] || u[1] != w[1] || u[2] != w[2])
        return false;
    }
  }
  return true;
}

// The fixedvel / fixedpos tags as TwoStepIntegrator::Advance and
// OneStepIntegrator::Advance read them (twostepintegrator.cc:17-65,
// onestepintegrator.cc:18-38): zero[i] for a fixedvel atom (its acceleration
// is set to zero after each force evaluation), hold[i] for a fixedpos atom
// whose tag reads "true" (not advanced).
struct HostTags {
  bool any = false;
  std::vector<char> zero, hold;
};

HostTags ReadTags(lpmd::Simulation& sim) {
  HostTags t;
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const bool hv = atoms.HaveAny(lpmd::Tag("fixedvel")), hp = atoms.HaveAny(lpmd::Tag("fixedpos"));
  if (!hv && !hp)
    return t;
  const long n = atoms.Size();
  t.any = true;
  t.zero.assign(n, 0);
  t.hold.assign(n, 0);
  for (long i = 0; i < n; ++i) {
    const lpmd::BasicAtom& at = atoms[i];
    t.zero[i] = hv && atoms.Have(at, lpmd::Tag("fixedvel"));
    t.hold[i] = hp && atoms.Have(at, lpmd::Tag("fixedpos")) && atoms.GetTag(at, lpmd::Tag("fixedpos")) == "true";
  }
  return t;
}

void ZeroFixedVel(lpmd::Simulation& sim, const HostTags& t) {
  if (!t.any)
    return;
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  for (long i = 0; i < atoms.Size(); ++i)
    if (t.zero[i])
      atoms[i].Acceleration() = lpmd::Vector(0.0, 0.0, 0.0);
}

bool Held(const HostTags& t, long i) { return t.any && t.hold[i]; }

// The session's kboltzmann as NoseHoover reads it (nosehoover.cc:69): a
// string parameter, "0.00008617" for the default unit system
// (runtime/units.cc:7 formatted by runtime/session.cc:17-23,32).  A
// configuration without a runtime context (a harness driving the plugin
// directly, where the reference plugin would throw) gets that string's value.
double SessionKBoltzmann(lpmd::Simulation& sim) {
  if (!sim.HasRuntimeContext())
    return 0.00008617;
  return double(sim.Context().session()["kboltzmann"]);
}

} // namespace

// ------------------------------------------------------------------ cell manager

DeviceCellManager::DeviceCellManager(const std::string& name, std::string args)
    : lpmd::Plugin(name, kVersion), kind_(name), args_(args) {
  Define(*this, name);
  ProcessArguments(args);
  cutoff_ = double((*this)["cutoff"]);
}

DeviceCellManager::~DeviceCellManager() {
  if (sys_)
    lpmdx_system_destroy(sys_);
}

void DeviceCellManager::ShowHelp() const {
  std::cout << " Module Name        = " << kind_ << " (MI355X engine)\n"
            << "      Cell binning and neighbour lists on the GPU; keywords as the CPU plugin"
            << ((kind_ == "verletlist" || kind_ == "minimumimage")
                    ? ", plus skin (list radius cutoff+skin, rebuilt when exceeded)"
                    : "")
            << ".\n";
}

double DeviceCellManager::Cutoff() const { return cutoff_; }

void DeviceCellManager::Reset() {
  if (sys_)
    lpmdx_system_destroy(sys_);
  sys_ = nullptr;
  n_ = -1;
  order_.clear();
  lists_ok_ = false;
}

void DeviceCellManager::Create(lpmd::Configuration& conf) {
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const lpmd::BasicCell& cell = conf.Cell();
  if (!cell.IsOrthogonal())
    throw lpmd::RuntimeError(kind_ + " (MI355X): the cell must be orthogonal");
  for (int q = 0; q < 3; ++q)
    if (!cell.Periodicity(q))
      throw lpmd::RuntimeError(kind_ + " (MI355X): the cell must be periodic along every axis");
  const long n = atoms.Size();
  std::vector<int> z(n);
  std::vector<double> m(n);
  for (long i = 0; i < n; ++i) {
    z[i] = atoms[i].Z();
    m[i] = atoms[i].Mass();
  }
  double box[3];
  for (int q = 0; q < 3; ++q)
    box[q] = cell[q][q];
  Reset();
  check(lpmdx_system_create(&sys_, n, box, z.data(), m.data()), kind_);
  ++version_;
  fixed_.clear();
  n_ = n;
  for (int q = 0; q < 3; ++q)
    box_[q] = box[q];
  Positions(conf, xbuf_);
  check(lpmdx_system_set_state(sys_, xbuf_.data(), nullptr, nullptr), kind_);
  check(lpmdx_set_cellmanager(sys_, kind_.c_str(), args_.c_str()), kind_); // runs one UpdateCell
}

lpmdx_system* DeviceCellManager::Bind(lpmd::Configuration& conf, bool upload) {
  const lpmd::BasicCell& cell = conf.Cell();
  bool same = sys_ && n_ == conf.Atoms().Size();
  for (int q = 0; q < 3 && same; ++q)
    same = box_[q] == cell[q][q];
  if (!same) {
    Create(conf);
    return sys_;
  }
  if (upload && !HostUnchanged(conf, xbuf_, nullptr)) {
    std::vector<double> x;
    Positions(conf, x);
    xbuf_.swap(x);
    check(lpmdx_system_set_state(sys_, xbuf_.data(), nullptr, nullptr), kind_);
    lists_ok_ = false;
    ++version_;
  }
  return sys_;
}

void DeviceCellManager::UpdateCell(lpmd::Configuration& conf) {
  const bool existed = sys_ != nullptr;
  Bind(conf, true);
  if (existed)
    check(lpmdx_update_cell(sys_), kind_);
  lists_ok_ = false;
  if (kind_ == "minimumimage" && fabs(cutoff_) < 1E-10) {
    // minimumimage.cc:63-69: without a cutoff, the largest half cell edge
    const lpmd::BasicCell& cell = conf.Cell();
    for (int q = 0; q < 3; ++q)
      if (0.5 * cell[q].Module() > cutoff_)
        cutoff_ = 0.5 * cell[q].Module();
  }
}

void DeviceCellManager::UpdateAtom(lpmd::Configuration&, long) {
  // positions are sent with the next UpdateCell or force call
  lists_ok_ = false;
}

void DeviceCellManager::BuildNeighborList(lpmd::Configuration& conf, long i, lpmd::NeighborList& nlist, bool full,
                                          double rcut) {
  if (!sys_)
    UpdateCell(conf);
  if (!lists_ok_ || full != lists_full_ || rcut != lists_rcut_) {
    const long n = n_;
    counts_.assign(n, 0);
    long total = 0;
    check(lpmdx_neighbor_lists(sys_, full ? 1 : 0, rcut, counts_.data(), 0, nullptr, nullptr, &total), kind_);
    js_.assign(total > 0 ? total : 1, 0);
    r2_.assign(total > 0 ? total : 1, 0.0);
    check(lpmdx_neighbor_lists(sys_, full ? 1 : 0, rcut, counts_.data(), total, js_.data(), r2_.data(), &total),
          kind_);
    offs_.assign(n + 1, 0);
    for (long k = 0; k < n; ++k)
      offs_[k + 1] = offs_[k] + counts_[k];
    lists_ok_ = true;
    lists_full_ = full;
    lists_rcut_ = rcut;
  }
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const lpmd::BasicCell& cell = conf.Cell();
  nlist.Clear();
  for (long k = offs_[i]; k < offs_[i + 1]; ++k) {
    const long j = js_[k];
    lpmd::AtomPair nn;
    nn.i = &atoms[i];
    nn.j = &atoms[j];
    nn.i_index = i;
    nn.j_index = j;
    nn.rij = cell.Displacement(atoms[i].Position(), atoms[j].Position());
    nn.r2 = r2_[k];
    nlist.Append(nn);
  }
}

int DeviceCellManager::IndexOf(const DevicePotential& p) {
  for (size_t k = 0; k < order_.size(); ++k)
    if (order_[k] == &p)
      return (int)k;
  int s1, s2;
  p.Species(s1, s2);
  check(lpmdx_potential_append(sys_, p.Kind().c_str(), p.Args().c_str(), s1, s2), p.Kind());
  order_.push_back(&p);
  ++version_;
  return (int)order_.size() - 1;
}

void DeviceCellManager::UsePotentials(const std::vector<const DevicePotential*>& pots) {
  if (pots == order_)
    return;
  check(lpmdx_potential_clear(sys_), kind_);
  order_.clear();
  ++version_;
  for (const DevicePotential* p : pots)
    IndexOf(*p);
}

void DeviceCellManager::SyncFixed(lpmd::Configuration& conf) {
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const bool hv = atoms.HaveAny(lpmd::Tag("fixedvel")), hp = atoms.HaveAny(lpmd::Tag("fixedpos"));
  std::vector<int> f;
  if (hv || hp) {
    // TwoStepIntegrator::Advance (twostepintegrator.cc:17-65): fixedvel by
    // presence, fixedpos when it reads "true"; either tag, whatever its
    // value, leaves the atom out of the tag-aware sums (properties.h:22-81)
    const long n = atoms.Size();
    f.assign(n, 0);
    for (long i = 0; i < n; ++i) {
      const lpmd::BasicAtom& at = atoms[i];
      if (hv && atoms.Have(at, lpmd::Tag("fixedvel")))
        f[i] |= LPMDX_FIX_VEL | LPMDX_FIX_TAGGED;
      if (hp && atoms.Have(at, lpmd::Tag("fixedpos"))) {
        f[i] |= LPMDX_FIX_TAGGED;
        if (atoms.GetTag(at, lpmd::Tag("fixedpos")) == "true")
          f[i] |= LPMDX_FIX_POS;
      }
    }
  }
  if (f == fixed_)
    return;
  check(lpmdx_system_set_fixed(sys_, f.empty() ? nullptr : f.data()), kind_);
  fixed_.swap(f);
}

// ------------------------------------------------------------------ potentials

namespace {

// Potential::SetCutoff as each plugin's constructor calls it: the cutoff
// keyword; max(c, d) for finnissinclair and finnissinclair-ext
// (finnissinclair.cc:10-31, finnissinclair-ext.cc:11-36); the table's last
// increasing r for tabulatedpair (tabulatedpair.cc:34-79); morse and harmonic
// keep Potential's HUGE_VAL (morse.cc:11-21, harmonic.cc:11-19).
template <typename P> PairForm Configure(P& pot, const std::string& name) {
  if (name == "tabulatedpair") {
    // the reference ends the program when the file does not open
    const lpmdx::PairTable T = lpmdx::read_pair_table(std::string(pot["file"]));
    if (!T.ok)
      throw lpmd::RuntimeError("Impossible open the potential data file.");
    pot.SetCutoff(T.cutoff);
    return MakeForm(name, [](const char*) { return 0.0; }, &T.e, &T.f, T.cutoff);
  }
  PairForm F = MakeForm(name, [&](const char* k) { return double(pot[k]); });
  if (name != "morse" && name != "harmonic")
    pot.SetCutoff(F.cutoff);
  return F;
}

DeviceCellManager* DeviceOf(lpmd::Configuration& conf, const std::string& who) {
  auto* cm = dynamic_cast<DeviceCellManager*>(&conf.GetCellManager());
  if (!cm)
    throw lpmd::RuntimeError(who + " (MI355X): the cell manager must be an MI355X cell manager plugin "
                                   "(linkedcell, verletlist, minimumimage, lcbinary)");
  return cm;
}

} // namespace

DevicePairPotential::DevicePairPotential(const std::string& name, std::string args)
    : lpmd::Plugin(name, kVersion), DevicePotential(name, args) {
  Define(*this, name);
  ProcessArguments(args);
  form_ = Configure(*this, name);
}

void DevicePairPotential::ShowHelp() const {
  std::cout << " Module Name        = " << kind_ << " (MI355X engine)\n"
            << "      Forces on the GPU through the lpmdx engine; keywords as the CPU plugin.\n";
}

DeviceMetalPotential::DeviceMetalPotential(const std::string& name, std::string args)
    : lpmd::Plugin(name, kVersion), DevicePotential(name, args) {
  Define(*this, name);
  ProcessArguments(args);
  form_ = Configure(*this, name);
}

void DeviceMetalPotential::ShowHelp() const {
  std::cout << " Module Name        = " << kind_ << " (MI355X engine)\n"
            << "      Two-pass EAM forces on the GPU through the lpmdx engine; keywords as the CPU plugin.\n";
}

void DevicePotential::Species(int& s1, int& s2) const {
  const lpmd::Potential& P = Pot();
  const int far = 1 << 20;
  if (P.AppliesTo(0, far) && P.AppliesTo(0, far + 1)) {
    s1 = s2 = -1; // wildcard
    return;
  }
  s1 = 0;
  s2 = far;
  for (int v = 0; v < 4096; ++v)
    if (P.AppliesTo(0, v)) {
      s2 = v;
      return;
    }
}

double DevicePotential::DeviceForces(lpmd::Configuration& conf) {
  DeviceCellManager* cm = DeviceOf(conf, kind_);
  lpmdx_system* sys = cm->Bind(conf, true);
  const int k = cm->IndexOf(*this);
  lpmd::BasicParticleSet& atoms = conf.Atoms();
  const long n = atoms.Size();
  acc_.resize(3 * (n > 0 ? n : 1));
  double e = 0.0, w = 0.0, s[9];
  check(lpmdx_potential_forces(sys, k, acc_.data(), &e, &w, s), kind_);
  for (long i = 0; i < n; ++i)
    atoms[i].Acceleration() += lpmd::Vector(acc_[3 * i], acc_[3 * i + 1], acc_[3 * i + 2]);
  conf.Virial() += w;
  lpmd::Matrix& st = conf.StressTensor();
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      st.Set(a, b, st.Get(a, b) + s[3 * a + b]);
  return e;
}

double DevicePotential::DeviceAtomEnergy(lpmd::Configuration& conf, long i) {
  DeviceCellManager* cm = DeviceOf(conf, kind_);
  lpmdx_system* sys = cm->Bind(conf, true);
  const int k = cm->IndexOf(*this);
  const long n = conf.Atoms().Size();
  if (i < 0 || i >= n)
    throw lpmd::RuntimeError(kind_ + " (MI355X): AtomEnergy of an atom index out of range");
  if (eatom_sys_ != sys || eatom_version_ != cm->Version() || (long)eatom_.size() != n) {
    eatom_.assign(n > 0 ? n : 1, 0.0);
    check(lpmdx_atom_energies(sys, k, eatom_.data()), kind_);
    eatom_sys_ = sys;
    eatom_version_ = cm->Version();
  }
  return eatom_[i];
}

// ------------------------------------------------------------------ integrator

DeviceVelocityVerlet::DeviceVelocityVerlet(std::string args) : lpmd::Plugin("velocityverlet", kVersion) {
  Define(*this, "velocityverlet");
  ProcessArguments(args);
  dt = double((*this)["dt"]);
  start_ = int((*this)["start"]);
}

void DeviceVelocityVerlet::ShowHelp() const {
  std::cout << " Module Name        = velocityverlet (MI355X engine)\n"
            << "      Velocity Verlet with dt (fs); on the GPU when every potential and the cell\n"
            << "      manager are MI355X plugins.\n";
}

void DeviceVelocityVerlet::Initialize(lpmd::Simulation&, lpmd::Potential&) {
  last_ = nullptr; // the next step sends the host state
}

void DeviceVelocityVerlet::HostStep(lpmd::Simulation& sim, lpmd::Potential& p) {
  // TwoStepIntegrator::Advance with VelocityVerlet's position and velocity
  // updates, in the reference's order of operations
  const HostTags tags = ReadTags(sim);
  p.UpdateForces(sim);
  ZeroFixedVel(sim, tags);
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  lpmd::BasicCell& cell = sim.Cell();
  const long n = atoms.Size();
  std::vector<lpmd::Vector> aold(n);
  for (long i = 0; i < n; ++i) {
    if (Held(tags, i))
      continue;
    const lpmd::Vector np =
        atoms[i].Position() + atoms[i].Velocity() * dt + 0.5 * atoms[i].Acceleration() * dt * dt;
    aold[i] = atoms[i].Acceleration();
    atoms[i].Position() = cell.FittedInside(np);
  }
  p.UpdateForces(sim);
  ZeroFixedVel(sim, tags);
  for (long i = 0; i < n; ++i)
    if (!Held(tags, i))
      atoms[i].Velocity() = atoms[i].Velocity() + 0.5 * dt * (aold[i] + atoms[i].Acceleration());
  last_ = nullptr;
}

DeviceCellManager* DevicePath(lpmd::Simulation& sim, lpmd::Potential& p, std::vector<const DevicePotential*>& dev) {
  auto* cp = dynamic_cast<lpmd::CombinedPotential*>(&p);
  auto* cm = dynamic_cast<DeviceCellManager*>(&sim.GetCellManager());
  dev.clear();
  if (!cm || !cp || cp->Size() == 0)
    return nullptr;
  for (long k = 0; k < cp->Size(); ++k) {
    auto* d = dynamic_cast<DevicePotential*>(&(*cp)[k]);
    if (!d)
      return nullptr;
    dev.push_back(d);
  }
  return cm;
}

namespace {

// After a device step: x, v, a into the host atoms (and hx, hv, ha), the
// observables CombinedPotential::UpdateForces leaves behind
// (combinedpotential.h:39-52) and the potential-energy tag.
void ReadBack(lpmd::Simulation& sim, lpmdx_system* sys, DeviceCellManager* cm, double ep, std::vector<double>& hx,
              std::vector<double>& hv, std::vector<double>& ha, const char* who) {
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  hx.resize(3 * n);
  hv.resize(3 * n);
  ha.resize(3 * n);
  check(lpmdx_system_get_state(sys, hx.data(), hv.data(), ha.data()), who);
  for (long i = 0; i < n; ++i) {
    atoms[i].Position() = lpmd::Vector(hx[3 * i], hx[3 * i + 1], hx[3 * i + 2]);
    atoms[i].Velocity() = lpmd::Vector(hv[3 * i], hv[3 * i + 1], hv[3 * i + 2]);
    atoms[i].Acceleration() = lpmd::Vector(ha[3 * i], ha[3 * i + 1], ha[3 * i + 2]);
  }
  cm->Mirror(hx); // host and device hold the same positions
  double obs[LPMDX_NOBS];
  check(lpmdx_observables(sys, obs), who);
  sim.Virial() = obs[LPMDX_OBS_VIRIAL];
  const lpmd::BasicCell& cell = sim.Cell();
  const double vol = cell[0][0] * cell[1][1] * cell[2][2];
  lpmd::Matrix& st = sim.StressTensor();
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      st.Set(a, b, obs[LPMDX_OBS_STRESS + 3 * a + b] * vol / 160217.65); // units.cc:9 pressfactor
  lpmd::Tag tag("potential-energy");
  sim.SetTag(sim, tag, ep);
}

} // namespace

void DeviceVelocityVerlet::Advance(lpmd::Simulation& sim, lpmd::Potential& p) {
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  if (!cm) {
    HostStep(sim, p);
    return;
  }
  const bool upload = cm->System() != last_ || !HostUnchanged(sim, hx_, &hv_);
  const double ep = DeviceRun(sim, cm, dev, 1, upload);
  ReadBack(sim, cm->System(), cm, ep, hx_, hv_, ha_, "velocityverlet");
}

double DeviceVelocityVerlet::DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm,
                                       const std::vector<const DevicePotential*>& dev, long steps, bool upload) {
  lpmdx_system* sys = cm->Bind(sim, false);
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (upload || sys != last_) {
    // the host changed the state (or this is the first step): send it; the
    // engine then starts from a fresh force evaluation
    std::vector<double> x, v;
    Positions(sim, x);
    Velocities(sim, v);
    check(lpmdx_system_set_state(sys, x.data(), v.data(), nullptr), "velocityverlet");
    hx_.swap(x);
    hv_.swap(v);
    last_ = sys;
  }
  double ep = 0.0, ek = 0.0;
  check(lpmdx_velocityverlet(sys, dt, steps, steps, &ep, &ek), "velocityverlet");
  return ep;
}

DeviceLeapfrog::DeviceLeapfrog(std::string args) : lpmd::Plugin("leapfrog", kVersion) {
  Define(*this, "leapfrog");
  ProcessArguments(args);
  dt = double((*this)["dt"]);
  start_ = int((*this)["start"]);
}

void DeviceLeapfrog::ShowHelp() const {
  std::cout << " Module Name        = leapfrog (MI355X engine)\n"
            << "      Leapfrog with dt (fs); on the GPU when every potential and the cell\n"
            << "      manager are MI355X plugins.\n";
}

void DeviceLeapfrog::Initialize(lpmd::Simulation& sim, lpmd::Potential& p) {
  // Integrator::UseOldConfig sets the old configuration up once per
  // integrator object (integrator.cc:16-22)
  if (mode_ != kUnset)
    return;
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  if (cm) {
    lpmdx_system* sys = cm->Bind(sim, false);
    cm->UsePotentials(dev);
    std::vector<double> x, v, a(3 * n);
    Positions(sim, x);
    Velocities(sim, v);
    for (long i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q)
        a[3 * i + q] = atoms[i].Acceleration()[q];
    check(lpmdx_system_set_state(sys, x.data(), v.data(), a.data()), "leapfrog");
    check(lpmdx_leapfrog_init(sys, dt), "leapfrog");
    hx_.swap(x);
    sys_ = sys;
    mode_ = kDevice;
    return;
  }
  // Leapfrog::Initialize (leapfrog.cc:58-67) on the host: the old
  // configuration of Integrator::GoBack (integrator.cc:31-42) is set in
  // place, its forces taken, and x, a restored
  lpmd::BasicCell& cell = sim.Cell();
  std::vector<lpmd::Vector> keep_x(n), keep_a(n);
  vh_.resize(n);
  for (long i = 0; i < n; ++i) {
    keep_x[i] = atoms[i].Position();
    keep_a[i] = atoms[i].Acceleration();
    const lpmd::Vector np = atoms[i].Position() - atoms[i].Velocity() * dt;
    vh_[i] = atoms[i].Velocity() - atoms[i].Acceleration() * dt;
    atoms[i].Position() = cell.FittedInside(np);
  }
  p.UpdateForces(sim);
  for (long i = 0; i < n; ++i) {
    vh_[i] += atoms[i].Acceleration() * 0.5 * dt;
    atoms[i].Position() = keep_x[i];
    atoms[i].Acceleration() = keep_a[i];
  }
  mode_ = kHost;
}

void DeviceLeapfrog::Advance(lpmd::Simulation& sim, lpmd::Potential& p) {
  if (mode_ == kUnset)
    throw lpmd::RuntimeError("leapfrog (MI355X): Advance before Initialize");
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  if ((cm != nullptr) != (mode_ == kDevice))
    throw lpmd::RuntimeError("leapfrog (MI355X): the potentials or the cell manager changed between Initialize and "
                             "Advance");
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  if (!cm) {
    // OneStepIntegrator::Advance with Leapfrog::Advance (onestepintegrator.cc:13-46,
    // leapfrog.cc:69-77); vh_ plays the old configuration's velocities
    const HostTags tags = ReadTags(sim);
    p.UpdateForces(sim);
    ZeroFixedVel(sim, tags);
    lpmd::BasicCell& cell = sim.Cell();
    for (long i = 0; i < n; ++i) {
      if (Held(tags, i))
        continue;
      const lpmd::Vector vhalf = vh_[i] + atoms[i].Acceleration() * dt;
      atoms[i].Position() = cell.FittedInside(atoms[i].Position() + vhalf * dt);
      atoms[i].Velocity() = 0.5 * (vhalf + vh_[i]);
      vh_[i] = vhalf;
    }
    return;
  }
  const double ep = DeviceRun(sim, cm, dev, 1, !HostUnchanged(sim, hx_, nullptr));
  ReadBack(sim, sys_, cm, ep, hx_, hv_, ha_, "leapfrog");
}

double DeviceLeapfrog::DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm,
                                 const std::vector<const DevicePotential*>& dev, long steps, bool upload) {
  lpmdx_system* sys = cm->Bind(sim, false);
  if (mode_ != kDevice || sys != sys_)
    throw lpmd::RuntimeError("leapfrog (MI355X): the cell manager's system changed since Initialize");
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (upload) {
    // the host moved atoms: Leapfrog::Advance reads the positions and the
    // half-step velocities only, so the positions alone are sent
    std::vector<double> x;
    Positions(sim, x);
    check(lpmdx_system_set_state(sys, x.data(), nullptr, nullptr), "leapfrog");
    hx_.swap(x);
  }
  double ep = 0.0, ek = 0.0;
  check(lpmdx_leapfrog(sys, dt, steps, steps, &ep, &ek), "leapfrog");
  return ep;
}

DeviceBeeman::DeviceBeeman(std::string args) : lpmd::Plugin("beeman", kVersion) {
  Define(*this, "beeman");
  ProcessArguments(args);
  dt = double((*this)["dt"]);
  start_ = int((*this)["start"]);
}

void DeviceBeeman::ShowHelp() const {
  std::cout << " Module Name        = beeman (MI355X engine)\n"
            << "      Beeman with dt (fs); on the GPU when every potential and the cell\n"
            << "      manager are MI355X plugins.\n";
}

void DeviceBeeman::Initialize(lpmd::Simulation& sim, lpmd::Potential& p) {
  if (mode_ != kUnset) // UseOldConfig sets the old configuration up once
    return;
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  if (cm) {
    lpmdx_system* sys = cm->Bind(sim, false);
    cm->UsePotentials(dev);
    std::vector<double> x, v, a(3 * n);
    Positions(sim, x);
    Velocities(sim, v);
    for (long i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q)
        a[3 * i + q] = atoms[i].Acceleration()[q];
    check(lpmdx_system_set_state(sys, x.data(), v.data(), a.data()), "beeman");
    check(lpmdx_beeman_init(sys, dt), "beeman");
    hx_.swap(x);
    hv_.swap(v);
    sys_ = sys;
    mode_ = kDevice;
    return;
  }
  // Beeman::Initialize (beeman.cc:55-61) on the host: forces at the old
  // configuration of Integrator::GoBack (integrator.cc:31-42), set in place
  // and then undone, give the previous accelerations
  lpmd::BasicCell& cell = sim.Cell();
  std::vector<lpmd::Vector> keep_x(n), keep_a(n);
  for (long i = 0; i < n; ++i) {
    keep_x[i] = atoms[i].Position();
    keep_a[i] = atoms[i].Acceleration();
    atoms[i].Position() = cell.FittedInside(atoms[i].Position() - atoms[i].Velocity() * dt);
  }
  p.UpdateForces(sim);
  a_old_.resize(n);
  for (long i = 0; i < n; ++i) {
    a_old_[i] = atoms[i].Acceleration();
    atoms[i].Position() = keep_x[i];
    atoms[i].Acceleration() = keep_a[i];
  }
  mode_ = kHost;
}

void DeviceBeeman::Advance(lpmd::Simulation& sim, lpmd::Potential& p) {
  if (mode_ == kUnset)
    throw lpmd::RuntimeError("beeman (MI355X): Advance before Initialize");
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  if ((cm != nullptr) != (mode_ == kDevice))
    throw lpmd::RuntimeError("beeman (MI355X): the potentials or the cell manager changed between Initialize and "
                             "Advance");
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  if (!cm) {
    // TwoStepIntegrator::Advance with Beeman's AdvancePosition and
    // AdvanceVelocity (twostepintegrator.cc:12-68, beeman.cc:63-84)
    lpmd::BasicCell& cell = sim.Cell();
    const HostTags tags = ReadTags(sim);
    p.UpdateForces(sim);
    ZeroFixedVel(sim, tags);
    std::vector<lpmd::Vector> aux(n);
    for (long i = 0; i < n; ++i) {
      if (Held(tags, i))
        continue;
      aux[i] = a_old_[i];
      const lpmd::Vector np = atoms[i].Position() + atoms[i].Velocity() * dt +
                              (2.0 / 3.0) * atoms[i].Acceleration() * dt * dt -
                              (1.0 / 6.0) * a_old_[i] * dt * dt;
      a_old_[i] = atoms[i].Acceleration();
      atoms[i].Position() = cell.FittedInside(np);
    }
    p.UpdateForces(sim);
    ZeroFixedVel(sim, tags);
    for (long i = 0; i < n; ++i)
      if (!Held(tags, i))
        atoms[i].Velocity() = atoms[i].Velocity() + (1.0 / 3.0) * atoms[i].Acceleration() * dt +
                              (5.0 / 6.0) * a_old_[i] * dt - (1.0 / 6.0) * aux[i] * dt;
    return;
  }
  const double ep = DeviceRun(sim, cm, dev, 1, !HostUnchanged(sim, hx_, &hv_));
  ReadBack(sim, sys_, cm, ep, hx_, hv_, ha_, "beeman");
}

double DeviceBeeman::DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm,
                               const std::vector<const DevicePotential*>& dev, long steps, bool upload) {
  lpmdx_system* sys = cm->Bind(sim, false);
  if (mode_ != kDevice || sys != sys_)
    throw lpmd::RuntimeError("beeman (MI355X): the cell manager's system changed since Initialize");
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (upload) {
    std::vector<double> x, v;
    Positions(sim, x);
    Velocities(sim, v);
    // the host changed the state: send it (the previous accelerations stay)
    check(lpmdx_system_set_state(sys, x.data(), v.data(), nullptr), "beeman");
    hx_.swap(x);
    hv_.swap(v);
  }
  double ep = 0.0, ek = 0.0;
  check(lpmdx_beeman(sys, dt, steps, steps, &ep, &ek), "beeman");
  return ep;
}

// ------------------------------------------------------------------ nosehoover

DeviceNoseHoover::DeviceNoseHoover(std::string args) : lpmd::Plugin("nosehoover", kVersion) {
  Define(*this, "nosehoover");
  ProcessArguments(args);
  dt = double((*this)["dt"]);
  start_ = int((*this)["start"]);
  q_ = double((*this)["fmass"]);
  temp_ = double((*this)["t"]);
}

void DeviceNoseHoover::ShowHelp() const {
  std::cout << " Module Name        = nosehoover (MI355X engine)\n"
            << "      Nose-Hoover thermostat (NVT) with dt (fs), fmass (the fictitious mass Q)\n"
            << "      and t (the reservoir temperature, K); on the GPU when every potential\n"
            << "      and the cell manager are MI355X plugins.\n";
}

void DeviceNoseHoover::Initialize(lpmd::Simulation& sim, lpmd::Potential& p) {
  if (mode_ != kUnset) // UseOldConfig sets the old configuration up once
    return;
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  const double kb = SessionKBoltzmann(sim);
  if (cm) {
    lpmdx_system* sys = cm->Bind(sim, false);
    cm->UsePotentials(dev);
    std::vector<double> x, v, a(3 * n);
    Positions(sim, x);
    Velocities(sim, v);
    for (long i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q)
        a[3 * i + q] = atoms[i].Acceleration()[q];
    check(lpmdx_system_set_state(sys, x.data(), v.data(), a.data()), "nosehoover");
    check(lpmdx_nosehoover_init(sys, dt, q_, temp_, kb), "nosehoover");
    hx_.swap(x);
    hv_.swap(v);
    sys_ = sys;
    mode_ = kDevice;
    return;
  }
  // on the host: forces at the old configuration of Integrator::GoBack
  // (integrator.cc:31-42), set in place and then undone, give old.a
  lpmd::BasicCell& cell = sim.Cell();
  std::vector<lpmd::Vector> keep_x(n), keep_a(n);
  for (long i = 0; i < n; ++i) {
    keep_x[i] = atoms[i].Position();
    keep_a[i] = atoms[i].Acceleration();
    atoms[i].Position() = cell.FittedInside(atoms[i].Position() - atoms[i].Velocity() * dt);
  }
  p.UpdateForces(sim);
  a_old_.resize(n);
  aux_.assign(n, lpmd::Vector());
  for (long i = 0; i < n; ++i) {
    a_old_[i] = atoms[i].Acceleration();
    atoms[i].Position() = keep_x[i];
    atoms[i].Acceleration() = keep_a[i];
  }
  friction_ = 0.0;
  mode_ = kHost;
}

void DeviceNoseHoover::Advance(lpmd::Simulation& sim, lpmd::Potential& p) {
  if (mode_ == kUnset)
    throw lpmd::RuntimeError("nosehoover (MI355X): Advance before Initialize");
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = DevicePath(sim, p, dev);
  if ((cm != nullptr) != (mode_ == kDevice))
    throw lpmd::RuntimeError("nosehoover (MI355X): the potentials or the cell manager changed between Initialize "
                             "and Advance");
  lpmd::BasicParticleSet& atoms = sim.Atoms();
  const long n = atoms.Size();
  if (!cm) {
    // TwoStepIntegrator::Advance with NoseHoover's AdvancePosition and
    // AdvanceVelocity (twostepintegrator.cc:12-68, nosehoover.cc:70-96).  The
    // temperature the friction reads is that of the velocities, which the
    // position loop does not change: one evaluation serves every atom.
    const double kb = SessionKBoltzmann(sim);
    lpmd::BasicCell& cell = sim.Cell();
    const HostTags tags = ReadTags(sim);
    p.UpdateForces(sim);
    ZeroFixedVel(sim, tags);
    const double T = lpmd::Temperature(atoms);
    for (long i = 0; i < n; ++i) {
      if (Held(tags, i))
        continue;
      aux_[i] = a_old_[i];
      friction_ += ((3.0 * n / q_) * kb * (T - temp_) * dt);
      const lpmd::Vector newacc = atoms[i].Acceleration() - atoms[i].Velocity() * friction_;
      atoms[i].Acceleration() = newacc;
      const lpmd::Vector np = atoms[i].Position() + atoms[i].Velocity() * dt + (2.0 / 3.0) * newacc * dt * dt -
                              (1.0 / 6.0) * a_old_[i] * dt * dt;
      a_old_[i] = newacc;
      atoms[i].Position() = cell.FittedInside(np);
    }
    p.UpdateForces(sim);
    ZeroFixedVel(sim, tags);
    for (long i = 0; i < n; ++i)
      if (!Held(tags, i))
        atoms[i].Velocity() = atoms[i].Velocity() + (1.0 / 3.0) * atoms[i].Acceleration() * dt +
                              (5.0 / 6.0) * a_old_[i] * dt - (1.0 / 6.0) * aux_[i] * dt;
    return;
  }
  const double ep = DeviceRun(sim, cm, dev, 1, !HostUnchanged(sim, hx_, &hv_));
  ReadBack(sim, sys_, cm, ep, hx_, hv_, ha_, "nosehoover");
}

double DeviceNoseHoover::DeviceRun(lpmd::Simulation& sim, DeviceCellManager* cm,
                                   const std::vector<const DevicePotential*>& dev, long steps, bool upload) {
  lpmdx_system* sys = cm->Bind(sim, false);
  if (mode_ != kDevice || sys != sys_)
    throw lpmd::RuntimeError("nosehoover (MI355X): the cell manager's system changed since Initialize");
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (upload) {
    std::vector<double> x, v;
    Positions(sim, x);
    Velocities(sim, v);
    // the host changed the state (a modifier, a tag): send it; the history
    // and the friction stay
    check(lpmdx_system_set_state(sys, x.data(), v.data(), nullptr), "nosehoover");
    hx_.swap(x);
    hv_.swap(v);
  }
  double ep = 0.0, ek = 0.0;
  check(lpmdx_nosehoover(sys, dt, steps, steps, &ep, &ek), "nosehoover");
  return ep;
}

// ------------------------------------------------------------------ simulation

namespace {

// Default unit system's session values (runtime/units.cc:5-13 through the
// session's 8-decimal strings, runtime/session.cc:17-38), for a simulation
// without a runtime context.
constexpr double kPressFactorDefault = 160217.65;
constexpr double kUa3ToGrCm3Default = 1.6605;
constexpr double kBoltzmannDefault = 8.6173422E-05;

std::vector<std::unique_ptr<DeviceSimulation>>& Owned() {
  static std::vector<std::unique_ptr<DeviceSimulation>> owned;
  return owned;
}

} // namespace

DeviceSimulation::DeviceSimulation(long natoms, const lpmd::BasicAtom& at) : atoms_(natoms, at) {
  SetTag(*this, lpmd::Tag("step"), step_);
}

void DeviceSimulation::Refresh() const {
  // the device holds the newer state: x, v and a into the host atoms
  if (!device_ahead_ || raw_view_)
    return;
  lpmdx_system* sys = dev_ ? dev_->System() : nullptr;
  device_ahead_ = false;
  if (!sys)
    return;
  auto& atoms = const_cast<lpmd::FixedSizeParticleSet&>(atoms_);
  const long n = atoms.Size();
  std::vector<double> a(3 * n);
  agreed_x_.resize(3 * n);
  agreed_v_.resize(3 * n);
  check(lpmdx_system_get_state(sys, agreed_x_.data(), agreed_v_.data(), a.data()), "simulation");
  for (long i = 0; i < n; ++i) {
    const double* x = agreed_x_.data() + 3 * i;
    const double* v = agreed_v_.data() + 3 * i;
    atoms[i].Position() = lpmd::Vector(x[0], x[1], x[2]);
    atoms[i].Velocity() = lpmd::Vector(v[0], v[1], v[2]);
    atoms[i].Acceleration() = lpmd::Vector(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
  }
  dev_->Mirror(agreed_x_); // the device holds these positions
  ++refreshes_;
}

lpmd::BasicParticleSet& DeviceSimulation::OriginalAtoms() {
  if (raw_view_)
    return atoms_;
  Refresh();
  handed_out_ = true;
  return atoms_;
}

const lpmd::BasicParticleSet& DeviceSimulation::OriginalAtoms() const {
  Refresh();
  return atoms_;
}

lpmd::BasicCell& DeviceSimulation::OriginalCell() {
  if (!raw_view_) {
    Refresh();
    handed_out_ = true;
  }
  return cell_;
}

bool DeviceSimulation::HostChanged() const {
  // positions or velocities written since host and device last agreed
  const long n = atoms_.Size();
  if ((long)agreed_x_.size() != 3 * n || (long)agreed_v_.size() != 3 * n)
    return true;
  for (long i = 0; i < n; ++i) {
    const lpmd::BasicAtom& at = atoms_[i];
    const lpmd::Vector& p = at.Position();
    const lpmd::Vector& u = at.Velocity();
    const double* x = agreed_x_.data() + 3 * i;
    const double* v = agreed_v_.data() + 3 * i;
    if (p[0] != x[0] || p[1] != x[1] || p[2] != x[2] || u[0] != v[0] || u[1] != v[1] || u[2] != v[2])
      return true;
  }
  return false;
}

double DeviceSimulation::Session(const char* key, double fallback) const {
  if (!HasRuntimeContext())
    return fallback;
  return double(const_cast<DeviceSimulation*>(this)->Context().session()[key]);
}

lpmd::Integrator& DeviceSimulation::Integrator() {
  if (integ_ == nullptr)
    throw lpmd::MissingComponent("integrator");
  return *integ_;
}

void DeviceSimulation::AdjustCurrentStep(long s) {
  step_ = s;
  SetTag(*this, lpmd::Tag("step"), step_);
}

// SimulationEngine::Initialize (simulationbuilder.cc:187-193)
void DeviceSimulation::Initialize() {
  initial_step_ = int(lpmd::Parameter(GetTag(*this, lpmd::Tag("step"))));
  Integrator().Initialize(*this, potarray_);
  step_ = initial_step_;
  SetTag(*this, lpmd::Tag("step"), step_);
  initialized_ = true;
}

// SimulationEngine::InitVelocities (simulationbuilder.cc:167-185): a random
// direction of fixed speed per atom (drand48 through RandomVector), the mean
// momentum over the mass taken out.
void DeviceSimulation::InitVelocities() {
  lpmd::BasicParticleSet& atoms = OriginalAtoms();
  const long nparts = atoms.Size();
  const double kb = HasRuntimeContext() ? Context().units().kboltzmann : kBoltzmannDefault;
  const double Kpart = (3.0 / 2.0) * kb * 1.0;
  lpmd::Vector vel, totalp;
  for (long i = 0; i < nparts; ++i) {
    const double v2 = 2.0 * Kpart / atoms[i].Mass();
    vel = lpmd::RandomVector(sqrt(v2));
    totalp += vel * atoms[i].Mass();
    atoms[i].Velocity() = vel;
  }
  totalp = totalp / nparts;
  for (long i = 0; i < nparts; ++i)
    atoms[i].Velocity() -= (totalp / atoms[i].Mass());
  velocities_set_ = true;
}

// SimulationEngine::SetTemperature (simulationbuilder.cc:53-71), on the host
// atoms; a changed state is sent before the next device step.
void DeviceSimulation::SetTemperature(double temp, bool tag) {
  if (!velocities_set_)
    InitVelocities();
  lpmd::BasicParticleSet& atoms = Atoms();
  const double ti = lpmd::Temperature(atoms, tag);
  for (long i = 0; i < atoms.Size(); ++i) {
    if (tag && (atoms.Have(atoms[i], lpmd::Tag("fixedvel")) || atoms.Have(atoms[i], lpmd::Tag("fixedpos"))))
      continue;
    atoms[i].Velocity() *= sqrt(temp / ti);
  }
}

// SimulationEngine::SetSimulationTags (simulationbuilder.cc:84-124) from the host atoms
void DeviceSimulation::SetSimulationTags() {
  const double pressfactor = Session("pressfactor", kPressFactorDefault);
  const double densityfact = Session("ua3togrcm3", kUa3ToGrCm3Default);
  lpmd::BasicParticleSet& atoms = Atoms();
  const lpmd::BasicCell& cell = Cell();
  const double v = cell.Volume();
  const bool tagged = atoms.HaveAny(lpmd::Tag("fixedvel")) || atoms.HaveAny(lpmd::Tag("fixedpos"));
  const lpmd::Vector P = lpmd::Momentum(atoms, tagged);
  SetTag(*this, lpmd::Tag("step"), step_);
  SetTag(*this, lpmd::Tag("temperature"), lpmd::Temperature(atoms, tagged));
  SetTag(*this, lpmd::Tag("volume"), v);
  SetTag(*this, lpmd::Tag("volume-per-atom"), v / double(atoms.Size()));
  SetTag(*this, lpmd::Tag("cell-a"), cell[0].Module());
  SetTag(*this, lpmd::Tag("cell-b"), cell[1].Module());
  SetTag(*this, lpmd::Tag("cell-c"), cell[2].Module());
  SetTag(*this, lpmd::Tag("particle-density"), double(atoms.Size()) / v);
  SetTag(*this, lpmd::Tag("density"), densityfact * lpmd::Density(atoms, cell));
  SetTag(*this, lpmd::Tag("momentum"), P.Module());
  SetTag(*this, lpmd::Tag("px"), P[0]);
  SetTag(*this, lpmd::Tag("py"), P[1]);
  SetTag(*this, lpmd::Tag("pz"), P[2]);
  const double potenerg = double(lpmd::Parameter(GetTag(*this, lpmd::Tag("potential-energy"))));
  const double kinenerg = lpmd::KineticEnergy(atoms, tagged);
  SetTag(*this, lpmd::Tag("kinetic-energy"), kinenerg);
  SetTag(*this, lpmd::Tag("total-energy"), potenerg + kinenerg);
  const lpmd::Matrix& stress = StressTensor();
  SetTag(*this, lpmd::Tag("virial-pressure"), (pressfactor / v) * (1.0 / 3.0) * Virial());
  SetTag(*this, lpmd::Tag("kinetic-pressure"), (pressfactor / v) * (2.0 / 3.0) * kinenerg);
  SetTag(*this, lpmd::Tag("pressure"), (pressfactor / v) * ((2.0 / 3.0) * kinenerg + (1.0 / 3.0) * Virial()));
  const char letter[3] = {'x', 'y', 'z'};
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      const char label[4] = {'s', letter[q], letter[p], '\0'};
      SetTag(*this, lpmd::Tag(std::string(label)), (pressfactor / v) * stress.Get(q, p));
    }
}

// The same tags from the engine's device reductions after a device run
// (lpmdx_observables: the kinetic sums leave tagged atoms out when any atom
// carries fixedvel or fixedpos, as SetSimulationTags' tag argument does), the
// potential-energy tag CombinedPotential::UpdateForces sets
// (combinedpotential.h:39-52), and the virial and stress of the last force
// evaluation.
void DeviceSimulation::DeviceTags(lpmdx_system* sys, double epot) {
  double obs[LPMDX_NOBS];
  check(lpmdx_observables(sys, obs), "simulation");
  const double pressfactor = Session("pressfactor", kPressFactorDefault);
  const double densityfact = Session("ua3togrcm3", kUa3ToGrCm3Default);
  const double v = cell_.Volume();
  Virial() = obs[LPMDX_OBS_VIRIAL];
  lpmd::Matrix& st = StressTensor();
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      st.Set(a, b, obs[LPMDX_OBS_STRESS + 3 * a + b] * v / kPressFactorDefault);
  SetTag(*this, lpmd::Tag("potential-energy"), epot);
  const double px = obs[LPMDX_OBS_PX], py = obs[LPMDX_OBS_PX + 1], pz = obs[LPMDX_OBS_PX + 2];
  const double kinenerg = obs[LPMDX_OBS_KINETIC];
  SetTag(*this, lpmd::Tag("step"), step_);
  SetTag(*this, lpmd::Tag("temperature"), obs[LPMDX_OBS_TEMP]);
  SetTag(*this, lpmd::Tag("volume"), v);
  SetTag(*this, lpmd::Tag("volume-per-atom"), v / double(atoms_.Size()));
  SetTag(*this, lpmd::Tag("cell-a"), cell_[0].Module());
  SetTag(*this, lpmd::Tag("cell-b"), cell_[1].Module());
  SetTag(*this, lpmd::Tag("cell-c"), cell_[2].Module());
  SetTag(*this, lpmd::Tag("particle-density"), double(atoms_.Size()) / v);
  SetTag(*this, lpmd::Tag("density"), densityfact * obs[LPMDX_OBS_DENSITY] / kUa3ToGrCm3Default);
  SetTag(*this, lpmd::Tag("momentum"), obs[LPMDX_OBS_MOMENTUM]);
  SetTag(*this, lpmd::Tag("px"), px);
  SetTag(*this, lpmd::Tag("py"), py);
  SetTag(*this, lpmd::Tag("pz"), pz);
  SetTag(*this, lpmd::Tag("kinetic-energy"), kinenerg);
  SetTag(*this, lpmd::Tag("total-energy"), epot + kinenerg);
  SetTag(*this, lpmd::Tag("virial-pressure"), (pressfactor / v) * (1.0 / 3.0) * Virial());
  SetTag(*this, lpmd::Tag("kinetic-pressure"), (pressfactor / v) * (2.0 / 3.0) * kinenerg);
  SetTag(*this, lpmd::Tag("pressure"), (pressfactor / v) * ((2.0 / 3.0) * kinenerg + (1.0 / 3.0) * Virial()));
  const char letter[3] = {'x', 'y', 'z'};
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      const char label[4] = {'s', letter[q], letter[p], '\0'};
      SetTag(*this, lpmd::Tag(std::string(label)), (pressfactor / v) * st.Get(q, p));
    }
}

// SimulationEngine::DoStep (simulationbuilder.cc:126-136) on the host
void DeviceSimulation::HostStep() {
  Virial() = 0.0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q)
      StressTensor().Set(q, p, 0.0);
  if (!initialized_)
    Initialize();
  Integrator().Advance(*this, potarray_);
  step_++;
  SetSimulationTags();
  ++host_steps_;
}

void DeviceSimulation::DoSteps(long n) {
  if (n <= 0)
    return;
  const bool first = !initialized_;
  if (first)
    Initialize(); // as the first DoStep does, before anything else
  auto* di = dynamic_cast<DeviceIntegrator*>(&Integrator());
  std::vector<const DevicePotential*> dev;
  DeviceCellManager* cm = di ? DevicePath(*this, potarray_, dev) : nullptr;
  if (!cm || !di->DeviceReady()) {
    for (long k = 0; k < n; ++k)
      HostStep();
    return;
  }
  bool send = false;
  if (first && restored_) {
    // a restored state: sent to the device before the run (an integrator with
    // an old configuration initialises once per object, as the reference's
    // do, and keeps its history)
    send = true;
    restored_ = false;
  } else if (first) {
    // Initialize has just handed the host state to the integrator (the
    // schemes with an old configuration sent it to the device there;
    // velocity Verlet sends it with its first run): host and device agree
    const long na = atoms_.Size();
    agreed_x_.resize(3 * na);
    agreed_v_.resize(3 * na);
    for (long i = 0; i < na; ++i)
      for (int q = 0; q < 3; ++q) {
        agreed_x_[3 * i + q] = atoms_[i].Position()[q];
        agreed_v_[3 * i + q] = atoms_[i].Velocity()[q];
      }
    dev_ = cm;
    device_ahead_ = false;
    handed_out_ = false;
  }
  // the host state goes to the device when the host holds the current state
  // and it differs from what host and device last agreed on (or the device
  // never held this simulation's state)
  const bool upload = send || (!device_ahead_ && (dev_ != cm || (handed_out_ && HostChanged())));
  Virial() = 0.0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q)
      StressTensor().Set(q, p, 0.0);
  double ep = 0.0;
  raw_view_ = !upload; // the integrator reads tags and sizes only, unless it sends the host state
  try {
    ep = di->DeviceRun(*this, cm, dev, n, upload);
  } catch (...) {
    raw_view_ = false;
    throw;
  }
  raw_view_ = false;
  if (upload) {
    // what was sent is what host and device agree on until the run moves it
    ++uploads_;
  }
  dev_ = cm;
  device_ahead_ = true;
  handed_out_ = false;
  ++runs_;
  step_ += n;
  DeviceTags(cm->System(), ep);
}

// SimulationEngine::RescalePositions (simulationbuilder.cc:155-162)
void DeviceSimulation::RescalePositions(const lpmd::BasicCell& old_cell) {
  lpmd::BasicParticleSet& t_atoms = Atoms();
  lpmd::BasicCell& t_cell = Cell();
  for (long i = 0; i < t_atoms.Size(); ++i)
    t_atoms[i].Position() = t_cell.FittedInside(t_cell.Cartesian(old_cell.Fractional(t_atoms[i].Position())));
}

// SimulationEngine::Dump (simulationbuilder.cc:195-236): step, cell vectors,
// atoms (Z, x, v, a), colour flags, atom tags
void DeviceSimulation::Dump(const std::string& path) const {
  const lpmd::BasicCell& cell = Cell();
  const lpmd::BasicParticleSet& atoms = Atoms();
  std::ofstream f(path.c_str());
  f << step_ << '\n';
  for (int i = 0; i < 3; ++i) {
    for (int q = 0; q < 3; ++q)
      f << cell[i][q] << " ";
    f << '\n';
  }
  f << atoms.Size() << '\n';
  for (long i = 0; i < atoms.Size(); ++i) {
    const lpmd::Vector& pos = atoms[i].Position();
    const lpmd::Vector& vel = atoms[i].Velocity();
    const lpmd::Vector& acc = atoms[i].Acceleration();
    f << atoms[i].Z() << " ";
    for (int q = 0; q < 3; ++q)
      f << pos[q] << " ";
    for (int q = 0; q < 3; ++q)
      f << vel[q] << " ";
    for (int q = 0; q < 3; ++q)
      f << acc[q] << " ";
    f << '\n';
  }
  for (long i = 0; i < atoms.Size(); ++i) {
    const bool has = lpmd::ColorHandler::HaveColor(atoms[i]);
    f << std::boolalpha << has << " ";
    if (has) {
      const lpmd::Color& color = lpmd::ColorHandler::ColorOfAtom(atoms[i]);
      for (int q = 0; q < 3; ++q)
        f << color[q] << " ";
    }
    f << '\n';
  }
  for (long i = 0; i < atoms.Size(); ++i) {
    const lpmd::Array<lpmd::Parameter> ap = atoms.GetTags(atoms[i]);
    f << ap.Size() << " ";
    for (int q = 0; q < ap.Size(); ++q)
      f << ap[q] << " " << atoms.GetTag(atoms[i], lpmd::Tag(ap[q])) << " ";
    f << '\n';
  }
}

// SimulationEngine::Restore (simulationbuilder.cc:238-290): the inverse of
// Dump; the velocities count as set and the integrator is initialised again
// at the next step
void DeviceSimulation::Restore(const std::string& path) {
  lpmd::BasicCell& cell = Cell();
  lpmd::BasicParticleSet& atoms = Atoms();
  std::ifstream f(path.c_str());
  f >> initial_step_;
  SetTag(*this, lpmd::Tag("step"), initial_step_);
  for (int i = 0; i < 3; ++i)
    for (int q = 0; q < 3; ++q)
      f >> cell[i][q];
  long natoms;
  f >> natoms;
  atoms.Clear();
  for (long i = 0; i < natoms; ++i) {
    int z;
    f >> z;
    atoms.Append(lpmd::Atom(z));
    lpmd::Vector& pos = atoms[i].Position();
    lpmd::Vector& vel = atoms[i].Velocity();
    lpmd::Vector& acc = atoms[i].Acceleration();
    for (int q = 0; q < 3; ++q)
      f >> pos[q];
    for (int q = 0; q < 3; ++q)
      f >> vel[q];
    for (int q = 0; q < 3; ++q)
      f >> acc[q];
  }
  std::string key, value;
  for (long i = 0; i < natoms; ++i) {
    f >> key;
    if (key == "true") {
      lpmd::Color color;
      for (int q = 0; q < 3; ++q)
        f >> color[q];
      lpmd::ColorHandler::ColorOfAtom(atoms[i]) = color;
    }
  }
  for (long i = 0; i < natoms; ++i) {
    long tag_number;
    f >> tag_number;
    for (long q = 0; q < tag_number; ++q) {
      f >> key;
      f >> value;
      atoms.SetTag(atoms[i], lpmd::Tag(key), value);
    }
  }
  velocities_set_ = true;
  initialized_ = false;
  restored_ = true;
  step_ = initial_step_;
}

lpmd::Simulation& DeviceSimulationBuilder::CreateFixedOrthogonal(lpmd::RuntimeContext& context, long atoms,
                                                                 const lpmd::BasicAtom& at) {
  lpmd::Simulation& s = CreateFixedOrthogonal(atoms, at);
  s.SetRuntimeContext(context);
  return s;
}

lpmd::Simulation& DeviceSimulationBuilder::CreateFixedOrthogonal(long atoms, const lpmd::BasicAtom& at) {
  Owned().push_back(std::make_unique<DeviceSimulation>(atoms, at));
  return *Owned().back();
}

lpmd::Simulation& DeviceSimulationBuilder::CloneOptimized(lpmd::RuntimeContext& context, const lpmd::Simulation& sim) {
  const lpmd::BasicParticleSet& atoms = sim.Atoms();
  const lpmd::BasicCell& cell = sim.Cell();
  if (!cell.IsOrthogonal())
    throw lpmd::RuntimeError("DeviceSimulationBuilder (MI355X): the cell must be orthogonal");
  lpmd::Simulation& s = CreateFixedOrthogonal(context, atoms.Size(), atoms[0]);
  s.AdjustCurrentStep(int(lpmd::Parameter(sim.GetTag(sim, lpmd::Tag("step")))));
  lpmd::BasicParticleSet& newatoms = s.Atoms();
  for (long i = 0; i < atoms.Size(); ++i) {
    newatoms[i] = atoms[i];
    const lpmd::Array<lpmd::Parameter> ap = atoms.GetTags(atoms[i]);
    for (int q = 0; q < ap.Size(); ++q)
      newatoms.SetTag(newatoms[i], lpmd::Tag(ap[q]), atoms.GetTag(atoms[i], lpmd::Tag(ap[q])));
  }
  for (int q = 0; q < 3; ++q)
    s.Cell()[q] = cell[q];
  return s;
}

void DeviceSimulationBuilder::Destroy(lpmd::Simulation& sim) {
  auto& v = Owned();
  for (auto it = v.begin(); it != v.end(); ++it)
    if (it->get() == &sim) {
      v.erase(it);
      return;
    }
}

// ------------------------------------------------------------------ factory

lpmd::Plugin* Create(const std::string& name, const std::string& args) {
  static const char* pair[] = {"lennardjones", "fastlj", "morse", "buckingham", "harmonic", "lennardjonesMod",
                               "tabulatedpair"};
  static const char* metal[] = {"suttonchen", "finnissinclair", "gupta", "finnissinclair-ext"};
  for (const char* p : pair)
    if (name == p)
      return new DevicePairPotential(name, args);
  for (const char* p : metal)
    if (name == p)
      return new DeviceMetalPotential(name, args);
  if (name == "linkedcell" || name == "verletlist" || name == "minimumimage" || name == "lcbinary")
    return new DeviceCellManager(name, args);
  if (name == "velocityverlet")
    return new DeviceVelocityVerlet(args);
  if (name == "leapfrog")
    return new DeviceLeapfrog(args);
  if (name == "beeman")
    return new DeviceBeeman(args);
  if (name == "nosehoover")
    return new DeviceNoseHoover(args);
  throw lpmd::RuntimeError("no MI355X plugin named " + name);
}

lpmd::Plugin* CreateOrExit(const std::string& name, const std::string& args) {
  try {
    return Create(name, args);
  } catch (const lpmd::Error& e) {
    std::cerr << "[Error] " << e.what() << '\n';
  } catch (const std::exception& e) {
    std::cerr << "[Error] " << name << ": " << e.what() << '\n';
  } catch (...) {
    std::cerr << "[Error] " << name << ": unknown error while creating the plugin\n";
  }
  std::exit(1);
}

} // namespace lpmdx_lpmd

// Implementation of the LPMD plugin classes over the engine's C ABI.
#include "lpmdx_lpmd.hpp"

#include "../csrc/table.hpp"

#include <lpmd/basicparticleset.h>
#include <lpmd/combinedpotential.h>
#include <lpmd/error.h>
#include <lpmd/properties.h>
#include <lpmd/simulation.h>
#include <runtime/runtime_context.h>

#include <cmath>
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

namespace {

// The device runs the step when every potential is a DevicePotential and the
// cell manager a DeviceCellManager; returns that cell manager, else null.
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
  lpmdx_system* sys = cm->Bind(sim, false);
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (sys != last_ || !HostUnchanged(sim, hx_, &hv_)) {
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
  check(lpmdx_velocityverlet(sys, dt, 1, 1, &ep, &ek), "velocityverlet");
  ReadBack(sim, sys, cm, ep, hx_, hv_, ha_, "velocityverlet");
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
  lpmdx_system* sys = cm->Bind(sim, false);
  if (sys != sys_)
    throw lpmd::RuntimeError("leapfrog (MI355X): the cell manager's system changed since Initialize");
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (!HostUnchanged(sim, hx_, nullptr)) {
    // the host moved atoms: Leapfrog::Advance reads the positions and the
    // half-step velocities only, so the positions alone are sent
    std::vector<double> x;
    Positions(sim, x);
    check(lpmdx_system_set_state(sys, x.data(), nullptr, nullptr), "leapfrog");
    hx_.swap(x);
  }
  double ep = 0.0, ek = 0.0;
  check(lpmdx_leapfrog(sys, dt, 1, 1, &ep, &ek), "leapfrog");
  ReadBack(sim, sys, cm, ep, hx_, hv_, ha_, "leapfrog");
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
  lpmdx_system* sys = cm->Bind(sim, false);
  if (sys != sys_)
    throw lpmd::RuntimeError("beeman (MI355X): the cell manager's system changed since Initialize");
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (!HostUnchanged(sim, hx_, &hv_)) {
    std::vector<double> x, v;
    Positions(sim, x);
    Velocities(sim, v);
    // the host changed the state: send it (the previous accelerations stay)
    check(lpmdx_system_set_state(sys, x.data(), v.data(), nullptr), "beeman");
    hx_.swap(x);
    hv_.swap(v);
  }
  double ep = 0.0, ek = 0.0;
  check(lpmdx_beeman(sys, dt, 1, 1, &ep, &ek), "beeman");
  ReadBack(sim, sys, cm, ep, hx_, hv_, ha_, "beeman");
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
  lpmdx_system* sys = cm->Bind(sim, false);
  if (sys != sys_)
    throw lpmd::RuntimeError("nosehoover (MI355X): the cell manager's system changed since Initialize");
  cm->UsePotentials(dev);
  cm->SyncFixed(sim);
  if (!HostUnchanged(sim, hx_, &hv_)) {
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
  check(lpmdx_nosehoover(sys, dt, 1, 1, &ep, &ek), "nosehoover");
  ReadBack(sim, sys, cm, ep, hx_, hv_, ha_, "nosehoover");
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

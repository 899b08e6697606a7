// Checker-only harness around a partial build of the LPMD reference.
//
// Test infrastructure: this file is compiled by oracle/ref.mk together with the
// reference's own plugin sources (plugins/linkedcell.cc, plugins/verletlist.cc,
// the seven potential plugins, plugins/crystal3d.cc, plugins/velocityverlet.cc)
// and the core sources they link against, into oracle/_ref/.  Only tests/ and
// tests/golden/gen_golden.py call it, to pin the FP64 oracle to the reference's
// real behaviour.  Nothing in the product path loads it.
//
// What it cannot cover: the reference's PairPotential/MetalPotential
// UpdateForces and CombinedPotential need a RuntimeContext, whose Session lives
// in liblpmd/runtime/session.cc, which includes the CMake-generated config.h.
// That generated header is not written here, so those compositions are pinned
// by invariants instead.  The integrator is driven with a harness potential
// (ForceTable below) that replays forces the caller supplies.
//
// Every entry point is extern "C", returns 0 on success and -1 on a caught
// exception (message via ref_last_error()).

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <lpmd/cellgenerator.h>
#include <lpmd/cellmanager.h>
#include <lpmd/combinedpotential.h>
#include <lpmd/configuration.h>
#include <lpmd/elements.h>
#include <lpmd/fixedsizeparticleset.h>
#include <lpmd/integrator.h>
#include <lpmd/metalpotential.h>
#include <lpmd/orthogonalcell.h>
#include <lpmd/pairpotential.h>
#include <lpmd/plugin.h>
#include <lpmd/simulation.h>
#include <lpmd/twostepintegrator.h>

#include "verletlist.h"

using namespace lpmd;

namespace {

std::string g_err;

// A plain Configuration over the reference's own storage classes
// (FixedSizeParticleSet + OrthogonalCell, as SimulationBuilder::CreateFixedOrthogonal
// uses, liblpmd/simulation/simulationbuilder.cc:321-326).
class HarnessConf : public Simulation {
public:
  HarnessConf(long n, int z) : atoms_(n, BasicAtomOf(z)) {}

  BasicParticleSet& OriginalAtoms() override { return atoms_; }
  const BasicParticleSet& OriginalAtoms() const override { return atoms_; }
  BasicCell& OriginalCell() override { return cell_; }
  const BasicCell& OriginalCell() const override { return cell_; }

  // Simulation interface: only what TwoStepIntegrator touches is meaningful.
  void SetTemperature(double, bool) override { Unused(); }
  CombinedPotential& Potentials() override { return pots_; }
  const CombinedPotential& Potentials() const override { return pots_; }
  void DoStep() override { Unused(); }
  void DoSteps(long) override { Unused(); }
  long CurrentStep() const override { return 0; }
  void AdjustCurrentStep(long) override {}
  void SetIntegrator(lpmd::Integrator&) override { Unused(); }
  lpmd::Integrator& Integrator() override { throw std::runtime_error("no integrator"); }
  void RescalePositions(const BasicCell&) override { Unused(); }
  void Dump(const std::string&) const override { Unused(); }
  void Restore(const std::string&) override { Unused(); }

private:
  static const BasicAtom& BasicAtomOf(int z) {
    static std::vector<Vector> store(3);
    static std::vector<BasicAtom*> keep;
    BasicAtom* a = new BasicAtom(z, &store[0], &store[1], &store[2]);
    keep.push_back(a);
    return *a;
  }
  [[noreturn]] static void Unused() { throw std::runtime_error("not used by the harness"); }
  FixedSizeParticleSet atoms_;
  OrthogonalCell cell_;
  CombinedPotential pots_;
};

struct Loaded {
  void* dl = nullptr;
  Plugin* plugin = nullptr;
};

Loaded Load(const char* so_path, const char* args) {
  Loaded L;
  L.dl = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
  if (!L.dl)
    throw std::runtime_error(std::string("dlopen: ") + dlerror());
  auto* create = reinterpret_cast<create_t*>(dlsym(L.dl, "create"));
  if (!create)
    throw std::runtime_error("no create symbol");
  L.plugin = create(std::string(args));
  return L;
}

// Replays per-atom accelerations supplied by the caller, one table per
// UpdateForces call, so that the reference's integrator can be driven exactly.
class ForceTable : public Potential {
public:
  std::vector<std::vector<double>> table; // per call: 3n accelerations
  size_t calls = 0;
  double energy(Configuration&) override { return 0.0; }
  double AtomEnergy(Configuration&, long) override { return 0.0; }
  void UpdateForces(Configuration& conf) override {
    BasicParticleSet& atoms = conf.Atoms();
    const std::vector<double>& a = table.at(calls < table.size() ? calls : table.size() - 1);
    for (long i = 0; i < atoms.Size(); ++i)
      atoms[i].Acceleration() = Vector(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    ++calls;
  }
};

template <typename F> int Guard(F&& f) {
  try {
    f();
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
  } catch (lpmd::Error& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown exception";
  }
  return -1;
}

void SetBox(HarnessConf& c, const double* L) {
  for (int q = 0; q < 3; ++q) {
    Vector v(0.0, 0.0, 0.0);
    v[q] = L[q];
    c.Cell()[q] = v;
  }
}

} // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// ---- configurations -------------------------------------------------------
void* ref_conf_new(long n, const double* L, const double* pos, const int* z) {
  HarnessConf* c = nullptr;
  if (Guard([&] {
        c = new HarnessConf(n, z ? z[0] : 18);
        SetBox(*c, L);
        for (long i = 0; i < n; ++i) {
          c->Atoms()[i].Position() = Vector(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
          if (z) {
            c->Atoms()[i].Z() = z[i];
            c->Atoms()[i].Mass() = ElemMass[z[i]];
          }
        }
      }) != 0)
    return nullptr;
  return c;
}

int ref_conf_set_positions(void* conf, const double* pos) {
  return Guard([&] {
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    for (long i = 0; i < c->Atoms().Size(); ++i)
      c->Atoms()[i].Position() = Vector(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  });
}

int ref_conf_get(void* conf, double* pos, double* vel, double* acc) {
  return Guard([&] {
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    for (long i = 0; i < c->Atoms().Size(); ++i)
      for (int q = 0; q < 3; ++q) {
        if (pos)
          pos[3 * i + q] = c->Atoms()[i].Position()[q];
        if (vel)
          vel[3 * i + q] = c->Atoms()[i].Velocity()[q];
        if (acc)
          acc[3 * i + q] = c->Atoms()[i].Acceleration()[q];
      }
  });
}

int ref_conf_set_velocities(void* conf, const double* vel) {
  return Guard([&] {
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    for (long i = 0; i < c->Atoms().Size(); ++i)
      c->Atoms()[i].Velocity() = Vector(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
  });
}

void ref_conf_free(void* conf) { delete static_cast<HarnessConf*>(conf); }

// ---- cell managers (plugins/linkedcell.cc, plugins/verletlist.cc) ---------
void* ref_cm_open(const char* so_path, const char* args) {
  Loaded* L = new Loaded();
  if (Guard([&] {
        *L = Load(so_path, args);
        if (!dynamic_cast<CellManager*>(L->plugin))
          throw std::runtime_error("not a CellManager");
      }) != 0) {
    delete L;
    return nullptr;
  }
  return L;
}

// Configuration::SetCellManager (configuration.h:54-57) or a later UpdateCell.
int ref_cm_update(void* cm, void* conf, int attach) {
  return Guard([&] {
    CellManager* m = dynamic_cast<CellManager*>(static_cast<Loaded*>(cm)->plugin);
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    if (attach)
      c->SetCellManager(*m);
    else
      m->UpdateCell(*c);
  });
}

// BuildNeighborList for every atom; writes counts[n] and, while room remains
// (cap entries), j indices, r2 and rij in the list's own order.
int ref_cm_lists(void* cm, void* conf, int full, double rcut, long cap, long* counts, long* js,
                 double* r2, double* rij, long* total) {
  return Guard([&] {
    CellManager* m = dynamic_cast<CellManager*>(static_cast<Loaded*>(cm)->plugin);
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    long k = 0;
    NeighborList nl;
    for (long i = 0; i < c->Atoms().Size(); ++i) {
      m->BuildNeighborList(*c, i, nl, full != 0, rcut);
      counts[i] = nl.Size();
      for (long p = 0; p < nl.Size(); ++p, ++k) {
        if (k < cap) {
          js[k] = nl[p].j_index;
          r2[k] = nl[p].r2;
          for (int q = 0; q < 3; ++q)
            rij[3 * k + q] = nl[p].rij[q];
        }
      }
    }
    *total = k;
  });
}

void ref_cm_close(void* cm) { delete static_cast<Loaded*>(cm); }

// verletlist reads members its constructor never sets (extcut, step_count,
// old_size; plugins/verletlist.h:27-28).  Built in zero-filled memory they take
// the values a fresh heap gives them, which is the behaviour the reference
// shows when it runs at all.
void* ref_verlet_open(const char* args) {
  Loaded* L = new Loaded();
  if (Guard([&] {
        void* mem = calloc(1, sizeof(VerletListCellManager));
        L->plugin = new (mem) VerletListCellManager(std::string(args));
      }) != 0) {
    delete L;
    return nullptr;
  }
  return L;
}

// ---- potentials (plugins/{lennardjones,fastlj,morse,buckingham}.cc, -------
// ---- plugins/{suttonchen,finnissinclair,gupta}.cc) -----------------------
void* ref_pot_open(const char* so_path, const char* args) {
  Loaded* L = new Loaded();
  if (Guard([&] {
        *L = Load(so_path, args);
        if (!dynamic_cast<Potential*>(L->plugin))
          throw std::runtime_error("not a Potential");
      }) != 0) {
    delete L;
    return nullptr;
  }
  return L;
}

double ref_pot_cutoff(void* pot) {
  return dynamic_cast<Potential*>(static_cast<Loaded*>(pot)->plugin)->GetCutoff();
}

// PairPotential::pairEnergy(sqrt(r2)) and pairForce(rij) for m displacement vectors.
int ref_pair_eval(void* pot, long m, const double* rij, double* e, double* f) {
  return Guard([&] {
    PairPotential* p = dynamic_cast<PairPotential*>(static_cast<Loaded*>(pot)->plugin);
    if (!p)
      throw std::runtime_error("not a PairPotential");
    for (long k = 0; k < m; ++k) {
      Vector v(rij[3 * k], rij[3 * k + 1], rij[3 * k + 2]);
      e[k] = p->pairEnergy(sqrt(v.SquareModule()));
      Vector ff = p->pairForce(v);
      for (int q = 0; q < 3; ++q)
        f[3 * k + q] = ff[q];
    }
  });
}

// MetalPotential virtuals for m pairs: rij (3m), rho_i, rho_j (m each).
// Outputs: pair energy, rhoij, F(rho_i), PairForce (3m), ManyBodies (3m), with
// r = sqrt(rij.SquareModule()) and normr = rij/|rij| as metalpotential.cc:146-149 form them.
int ref_metal_eval(void* pot, long m, const double* rij, const double* rhoi, const double* rhoj,
                   double* epair, double* rho, double* frho, double* pf, double* mb) {
  return Guard([&] {
    MetalPotential* p = dynamic_cast<MetalPotential*>(static_cast<Loaded*>(pot)->plugin);
    if (!p)
      throw std::runtime_error("not a MetalPotential");
    for (long k = 0; k < m; ++k) {
      Vector v(rij[3 * k], rij[3 * k + 1], rij[3 * k + 2]);
      double r = sqrt(v.SquareModule());
      Vector norm = v;
      norm.Normalize();
      epair[k] = p->pairEnergy(r);
      rho[k] = p->rhoij(r);
      frho[k] = p->F(rhoi[k]);
      Vector a = p->PairForce(norm, r);
      Vector b = p->ManyBodies(norm, rhoi[k], rhoj[k], r);
      for (int q = 0; q < 3; ++q) {
        pf[3 * k + q] = a[q];
        mb[3 * k + q] = b[q];
      }
    }
  });
}

void ref_pot_close(void* pot) { delete static_cast<Loaded*>(pot); }

// ---- crystal3d (plugins/crystal3d.cc:62-130) ------------------------------
int ref_crystal(const char* so_path, const char* args, long n, const double* L, double* pos) {
  return Guard([&] {
    Loaded g = Load(so_path, args);
    CellGenerator* gen = dynamic_cast<CellGenerator*>(g.plugin);
    if (!gen)
      throw std::runtime_error("not a CellGenerator");
    HarnessConf c(n, 18);
    SetBox(c, L);
    gen->Generate(c);
    for (long i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q)
        pos[3 * i + q] = c.Atoms()[i].Position()[q];
  });
}

// ---- an integrator plugin (plugins/velocityverlet.cc, plugins/leapfrog.cc)
// ---- over liblpmd/simulation/{two,one}stepintegrator.cc, integrator.cc ----
// Runs Integrator::Initialize, then `steps` calls of Integrator::Advance on
// conf.  The potential replays acc_tables[call] (3n doubles each, ncalls
// tables) on its UpdateForces calls, in order; the cell manager `cm` is
// attached first because Integrator::UseOldConfig needs one.
int ref_vv_run(const char* so_path, const char* args, void* conf, void* cm, long steps,
               long ncalls, const double* acc_tables) {
  return Guard([&] {
    Loaded g = Load(so_path, args);
    lpmd::Integrator* vv = dynamic_cast<lpmd::Integrator*>(g.plugin);
    if (!vv)
      throw std::runtime_error("not an Integrator");
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    CellManager* m = dynamic_cast<CellManager*>(static_cast<Loaded*>(cm)->plugin);
    c->SetCellManager(*m);
    const long n = c->Atoms().Size();
    ForceTable ft;
    for (long t = 0; t < ncalls; ++t)
      ft.table.emplace_back(acc_tables + 3 * n * t, acc_tables + 3 * n * (t + 1));
    vv->Initialize(*c, ft);
    for (long s = 0; s < steps; ++s)
      vv->Advance(*c, ft);
    // keep the plugin object alive with its old-config copy until here
    delete g.plugin;
  });
}

// ---- any plugin, and the reference's composition over loaded plugins -----
// These drive plugins that are not the reference's own (the MI355X drop-in
// plugins) through the reference's interfaces: CombinedPotential
// (combinedpotential.h:39-52), Configuration::SetCellManager and
// Integrator::Advance.
void* ref_plugin_open(const char* so_path, const char* args) {
  Loaded* L = new Loaded();
  if (Guard([&] { *L = Load(so_path, args); }) != 0) {
    delete L;
    return nullptr;
  }
  return L;
}

void ref_plugin_close(void* p) {
  Loaded* L = static_cast<Loaded*>(p);
  if (L)
    delete L->plugin;
  delete L;
}

// CombinedPotential::UpdateForces on conf (its cell manager attached before)
// with the given potentials and species pairs (s = -1: any); acc[3n] and
// out[11] = energy, virial, stress[a][b] (row-major, StressTensor().Get(a, b)).
int ref_combined_forces(void* conf, int npot, void** pots, const int* s1, const int* s2, double* acc,
                        double* out) {
  return Guard([&] {
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    CombinedPotential cp;
    for (int k = 0; k < npot; ++k) {
      Potential* p = dynamic_cast<Potential*>(static_cast<Loaded*>(pots[k])->plugin);
      if (!p)
        throw std::runtime_error("not a Potential");
      p->SetValidSpecies(s1 ? s1[k] : -1, s2 ? s2[k] : -1);
      cp.Append(*p);
    }
    cp.UpdateForces(*c);
    const long n = c->Atoms().Size();
    for (long i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q)
        acc[3 * i + q] = c->Atoms()[i].Acceleration()[q];
    out[0] = cp.energy(*c);
    out[1] = c->Virial();
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        out[2 + 3 * a + b] = c->StressTensor().Get(a, b);
  });
}

// Potential::AtomEnergy(conf, i) of a loaded potential, its species rule set
// to (s1, s2) first as the application sets it (SetValidSpecies), for the
// atoms i < n (its cell manager attached to conf before).
int ref_atom_energies(void* pot, void* conf, int s1, int s2, long n, double* e) {
  return Guard([&] {
    Potential* p = dynamic_cast<Potential*>(static_cast<Loaded*>(pot)->plugin);
    if (!p)
      throw std::runtime_error("not a Potential");
    p->SetValidSpecies(s1, s2);
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    for (long i = 0; i < n; ++i)
      e[i] = p->AtomEnergy(*c, i);
  });
}

// BasicParticleSet::SetTag (liblpmd/lpmd/taghandler.h:124-131) on atom i.
int ref_set_atom_tag(void* conf, long i, const char* key, const char* value) {
  return Guard([&] {
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    BasicParticleSet& atoms = c->Atoms();
    atoms.SetTag(atoms[i], Tag(key), std::string(value));
  });
}

// Integrator::Initialize then `steps` Integrator::Advance calls on conf with
// the given potentials (appended to the configuration's CombinedPotential);
// epot[s] = the potential-energy tag after step s, when the integrator sets it.
int ref_integrate(void* integrator, void* conf, int npot, void** pots, const int* s1, const int* s2, long steps,
                  double* epot) {
  return Guard([&] {
    lpmd::Integrator* itg = dynamic_cast<lpmd::Integrator*>(static_cast<Loaded*>(integrator)->plugin);
    if (!itg)
      throw std::runtime_error("not an Integrator");
    HarnessConf* c = static_cast<HarnessConf*>(conf);
    CombinedPotential& cp = c->Potentials();
    cp.Clear();
    for (int k = 0; k < npot; ++k) {
      Potential* p = dynamic_cast<Potential*>(static_cast<Loaded*>(pots[k])->plugin);
      if (!p)
        throw std::runtime_error("not a Potential");
      p->SetValidSpecies(s1 ? s1[k] : -1, s2 ? s2[k] : -1);
      cp.Append(*p);
    }
    itg->Initialize(*c, cp);
    Tag tag("potential-energy");
    for (long s = 0; s < steps; ++s) {
      itg->Advance(*c, cp);
      if (epot)
        epot[s] = c->HaveAny(tag) ? atof(c->GetTag(*c, tag).c_str()) : 0.0;
    }
  });
}

} // extern "C"

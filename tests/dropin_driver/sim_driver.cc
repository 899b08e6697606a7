// Test driver of the drop-in DeviceSimulation (lpmd2_amd/plugins/lpmdx_lpmd.hpp):
// C entry points that build a simulation with DeviceSimulationBuilder, load
// plugin objects through their create(args) (liblpmd/lpmd/plugin.h:42-46),
// run DoStep / DoSteps and read the state and tags back, so that the Python
// tests can drive it.  Test infrastructure only; built by
// lpmd2_amd/plugins/plugins.mk next to the drop-in plugins.
#include <dlfcn.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../lpmd2_amd/plugins/lpmdx_lpmd.hpp"

namespace {

std::string g_err;

struct Loaded {
  void* dl = nullptr;
  lpmd::Plugin* plugin = nullptr;
  ~Loaded() {
    if (plugin && dl) {
      auto* destroy = reinterpret_cast<void (*)(lpmd::Plugin*)>(dlsym(dl, "destroy"));
      if (destroy)
        destroy(plugin);
    }
  }
};

struct Driver {
  lpmd::Simulation* sim = nullptr;
  lpmdx_lpmd::DeviceSimulation* dsim = nullptr;
  std::vector<std::unique_ptr<Loaded>> plugins; // the simulation refers to them: freed after it
};

std::unique_ptr<Loaded> Load(const char* path, const char* args) {
  auto L = std::make_unique<Loaded>();
  L->dl = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!L->dl)
    throw std::runtime_error(std::string("dlopen: ") + dlerror());
  using create_fn = lpmd::Plugin* (*)(std::string);
  auto* create = reinterpret_cast<create_fn>(dlsym(L->dl, "create"));
  if (!create)
    throw std::runtime_error("no create symbol");
  L->plugin = create(std::string(args));
  return L;
}

template <typename F> int Guard(F&& f) {
  try {
    f();
    return 0;
  } catch (lpmd::Error& e) {
    g_err = e.what();
  } catch (std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown exception";
  }
  return -1;
}

} // namespace

extern "C" {

const char* sim_last_error() { return g_err.c_str(); }

// n atoms of species z[i] in an orthogonal periodic box L at positions x and
// velocities v (AoS [3n]); v may be NULL
void* sim_new(long n, const double* L, const int* z, const double* x, const double* v) {
  Driver* d = nullptr;
  const int rc = Guard([&] {
    auto D = std::make_unique<Driver>();
    D->sim = &lpmdx_lpmd::DeviceSimulationBuilder::CreateFixedOrthogonal(n, lpmd::Atom(z[0]));
    D->dsim = dynamic_cast<lpmdx_lpmd::DeviceSimulation*>(D->sim);
    lpmd::BasicParticleSet& atoms = D->sim->Atoms();
    for (long i = 0; i < n; ++i) {
      atoms[i] = lpmd::Atom(z[i]);
      atoms[i].Position() = lpmd::Vector(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
      if (v)
        atoms[i].Velocity() = lpmd::Vector(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
    for (int q = 0; q < 3; ++q) {
      lpmd::Vector e(0.0, 0.0, 0.0);
      e[q] = L[q];
      D->sim->Cell()[q] = e;
    }
    d = D.release();
  });
  return rc == 0 ? d : nullptr;
}

void sim_free(void* h) {
  auto* d = static_cast<Driver*>(h);
  if (!d)
    return;
  lpmdx_lpmd::DeviceSimulationBuilder::Destroy(*d->sim);
  delete d;
}

// Configuration::SetCellManager with the plugin made by path's create(args)
int sim_set_cellmanager(void* h, const char* path, const char* args) {
  return Guard([&] {
    auto* d = static_cast<Driver*>(h);
    auto L = Load(path, args);
    auto* cm = dynamic_cast<lpmd::CellManager*>(L->plugin);
    if (!cm)
      throw std::runtime_error("not a CellManager");
    d->plugins.push_back(std::move(L));
    d->sim->SetCellManager(*cm);
  });
}

// Potentials().Append with the species rule (s1, s2)
int sim_add_potential(void* h, const char* path, const char* args, int s1, int s2) {
  return Guard([&] {
    auto* d = static_cast<Driver*>(h);
    auto L = Load(path, args);
    auto* p = dynamic_cast<lpmd::Potential*>(L->plugin);
    if (!p)
      throw std::runtime_error("not a Potential");
    p->SetValidSpecies(s1, s2);
    d->plugins.push_back(std::move(L));
    d->sim->Potentials().Append(*p);
  });
}

int sim_set_integrator(void* h, const char* path, const char* args) {
  return Guard([&] {
    auto* d = static_cast<Driver*>(h);
    auto L = Load(path, args);
    auto* itg = dynamic_cast<lpmd::Integrator*>(L->plugin);
    if (!itg)
      throw std::runtime_error("not an Integrator");
    d->plugins.push_back(std::move(L));
    d->sim->SetIntegrator(*itg);
  });
}

int sim_do_steps(void* h, long n) {
  return Guard([&] { static_cast<Driver*>(h)->sim->DoSteps(n); });
}

int sim_do_step(void* h) {
  return Guard([&] { static_cast<Driver*>(h)->sim->DoStep(); });
}

// x, v, a through Atoms() (any may be NULL)
int sim_get(void* h, double* x, double* v, double* a) {
  return Guard([&] {
    lpmd::BasicParticleSet& atoms = static_cast<Driver*>(h)->sim->Atoms();
    for (long i = 0; i < atoms.Size(); ++i)
      for (int q = 0; q < 3; ++q) {
        if (x)
          x[3 * i + q] = atoms[i].Position()[q];
        if (v)
          v[3 * i + q] = atoms[i].Velocity()[q];
        if (a)
          a[3 * i + q] = atoms[i].Acceleration()[q];
      }
  });
}

// Atoms() called once, as a loop that reads or tags the atoms each step does
int sim_touch(void* h) {
  return Guard([&] { (void)static_cast<Driver*>(h)->sim->Atoms().Size(); });
}

// the velocity of atom i written through Atoms()
int sim_set_velocity(void* h, long i, const double* v) {
  return Guard([&] {
    lpmd::BasicParticleSet& atoms = static_cast<Driver*>(h)->sim->Atoms();
    atoms[i].Velocity() = lpmd::Vector(v[0], v[1], v[2]);
  });
}

// BasicParticleSet::SetTag(atoms[i], key, value)
int sim_set_atom_tag(void* h, long i, const char* key, const char* value) {
  return Guard([&] {
    lpmd::BasicParticleSet& atoms = static_cast<Driver*>(h)->sim->Atoms();
    atoms.SetTag(atoms[i], lpmd::Tag(key), std::string(value));
  });
}

int sim_set_temperature(void* h, double t, int tag) {
  return Guard([&] { static_cast<Driver*>(h)->sim->SetTemperature(t, tag != 0); });
}

// the simulation's tag `name` as a number
int sim_tag(void* h, const char* name, double* out) {
  return Guard([&] {
    lpmd::Simulation* s = static_cast<Driver*>(h)->sim;
    *out = double(lpmd::Parameter(s->GetTag(*s, lpmd::Tag(name))));
  });
}

// [host refreshes, device runs, uploads, host steps, current step]
int sim_counters(void* h, long* out) {
  return Guard([&] {
    auto* d = static_cast<Driver*>(h);
    out[0] = d->dsim->HostRefreshes();
    out[1] = d->dsim->DeviceRuns();
    out[2] = d->dsim->Uploads();
    out[3] = d->dsim->HostSteps();
    out[4] = d->sim->CurrentStep();
  });
}

int sim_dump(void* h, const char* path) {
  return Guard([&] { static_cast<Driver*>(h)->sim->Dump(path); });
}

int sim_restore(void* h, const char* path) {
  return Guard([&] { static_cast<Driver*>(h)->sim->Restore(path); });
}

} // extern "C"

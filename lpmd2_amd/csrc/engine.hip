// C ABI of the MI355X LPMD force-and-integration path (include/lpmdx.h): one
// lpmdx_system is one Configuration resident in HBM (engine_system.hpp).  The
// entry points are declared in include/lpmdx.h, each next to the reference
// interface it serves.
#include "engine_system.hpp"

#include <mutex>

namespace lpmdx {
thread_local std::string g_err;
}

// ---------------------------------------------------------------------------
// C ABI

namespace {

void need(const void* p, const char* what) {
  if (!p)
    fail(LPMDX_EINVAL, std::string(what) + " is NULL");
}

void finite(const double* p, long m, const char* what) {
  for (long k = 0; k < m; ++k)
    if (!std::isfinite(p[k]))
      fail(LPMDX_ENONFINITE, std::string(what) + " holds a non-finite value");
}

void ensure_device() {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0)
    fail(LPMDX_EDEVICE, "no HIP device available");
}

} // namespace

extern "C" {

const char* lpmdx_last_error(void) { return g_err.c_str(); }

int lpmdx_abi_version(void) { return LPMDX_ABI_VERSION; }

int lpmdx_device_count(int* count) {
  return guard([&] {
    need(count, "count");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess)
      c = 0;
    *count = c;
  });
}

int lpmdx_system_create(lpmdx_system** out, long n, const double box[3], const int* z,
                        const double* mass) {
  return guard([&] {
    need(out, "out");
    need(box, "box");
    *out = nullptr;
    if (n < 0 || n > (1L << 30))
      fail(LPMDX_EINVAL, "atom count out of range");
    if (n > 0)
      need(z, "z");
    finite(box, 3, "box");
    for (int q = 0; q < 3; ++q)
      if (!(box[q] > 0.0))
        fail(LPMDX_EINVAL, "box edges must be positive");
    ensure_device();
    auto S = std::make_unique<lpmdx_system>();
    hipchk(hipGetDevice(&S->device), "hipGetDevice");
    S->n = n;
    for (int q = 0; q < 3; ++q) {
      S->box.L[q] = box[q];
      S->box.half[q] = 0.5 * box[q];
      S->box.inv[q] = 1.0 / box[q];
      S->ll[q] = edge_module(box[q]);
    }
    hipchk(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking), "hipStreamCreate");
    hipchk(hipHostMalloc(&S->hctrl, sizeof(Ctrl)), "hipHostMalloc");
    hipchk(hipHostMalloc(&S->hred, 64 * 12 * sizeof(double)), "hipHostMalloc");
    const size_t m = n > 0 ? n : 1;
    for (DBuf<double>* b : {&S->x, &S->xn, &S->v, &S->a, &S->asum, &S->xb, &S->x2, &S->v2, &S->a2})
      b->alloc(3 * m);
    for (DBuf<double>* b : {&S->ffm, &S->mass, &S->ffm2, &S->mass2})
      b->alloc(m);
    for (DBuf<int>* b : {&S->z, &S->perm, &S->z2, &S->perm2})
      b->alloc(m);
    for (DBuf<uint32_t>* b : {&S->keys, &S->vals, &S->keys2, &S->vals2})
      b->alloc(m);
    S->ctrl.alloc(1);
    S->red.alloc(12 * 64);
    hipchk(hipMemsetAsync(S->ctrl.p, 0, sizeof(Ctrl), S->st), "memset");
    std::vector<double> ms(m, 1.0), ff(m, 0.0);
    std::vector<int> zz(m, 0), pm(m);
    for (long i = 0; i < n; ++i) {
      zz[i] = z[i];
      double mi = mass ? mass[i] : ElemMass(z[i]);
      if (!(mi > 0.0) || !std::isfinite(mi))
        fail(LPMDX_EINVAL, "no mass for atomic number " + std::to_string(z[i]));
      ms[i] = mi;
      ff[i] = kForceFactor / mi; // forcefactor / Mass(), pairpotential.cc:62
      pm[i] = (int)i;
    }
    S->hmass.assign(ms.begin(), ms.begin() + n);
    S->hz.assign(zz.begin(), zz.begin() + n);
    hipchk(hipMemcpyAsync(S->mass.p, ms.data(), m * sizeof(double), hipMemcpyHostToDevice, S->st), "memcpy");
    hipchk(hipMemcpyAsync(S->ffm.p, ff.data(), m * sizeof(double), hipMemcpyHostToDevice, S->st), "memcpy");
    hipchk(hipMemcpyAsync(S->z.p, zz.data(), m * sizeof(int), hipMemcpyHostToDevice, S->st), "memcpy");
    hipchk(hipMemcpyAsync(S->perm.p, pm.data(), m * sizeof(int), hipMemcpyHostToDevice, S->st), "memcpy");
    for (DBuf<double>* b : {&S->x, &S->v, &S->a})
      hipchk(hipMemsetAsync(b->p, 0, 3 * m * sizeof(double), S->st), "memset");
    S->sync();
    *out = S.release();
  });
}

int lpmdx_system_destroy(lpmdx_system* sys) {
  return guard([&] { delete sys; });
}

long lpmdx_system_size(const lpmdx_system* sys) { return sys ? sys->n : -1; }

}

extern "C" {

int lpmdx_system_set_state(lpmdx_system* S, const double* x, const double* v, const double* a) {
  return guard([&] {
    need(S, "system");
    if (S->n == 0)
      return;
    if (x)
      finite(x, 3 * S->n, "positions");
    if (v)
      finite(v, 3 * S->n, "velocities");
    if (a)
      finite(a, 3 * S->n, "accelerations");
    S->undrift();
    const double* const src[3] = {x, v, a};
    double* const dst[3] = {S->x.p, S->v.p, S->a.p};
    S->upload(src, dst);
    if (x) {
      S->forces_ok = false;
      S->pos_version++;
    }
    if (a)
      S->forces_ok = false;
  });
}

int lpmdx_system_get_state(lpmdx_system* S, double* x, double* v, double* a) {
  return guard([&] {
    need(S, "system");
    if (S->n == 0)
      return;
    const double* xu = S->drifted ? S->xn.p : S->x.p;
    const double* src[3] = {xu, S->v.p, S->a.p};
    double* const dst[3] = {x, v, a};
    S->download(src, dst);
  });
}

int lpmdx_set_cellmanager(lpmdx_system* S, const char* name, const char* args) {
  return guard([&] {
    need(S, "system");
    need(name, "name");
    std::string nm(name);
    CellMgr c;
    if (nm == "linkedcell") {
      // plugins/linkedcell.cc:17-41 (the mode keyword alone decides auto)
      Keywords k(args, {{"cutoff", "7.0"}, {"nx", "0"}, {"ny", "0"}, {"nz", "0"}, {"mode", "noauto"},
                        {"warn-outside", "true"}});
      c.kind = kCmLinkedCell;
      c.cutoff = k.d("cutoff");
      c.nx = k.i("nx");
      c.ny = k.i("ny");
      c.nz = k.i("nz");
      std::string mode = k.s("mode");
      c.mode_auto = (mode == "auto" || mode == "AUTO");
      c.warn_outside = k.b("warn-outside");
    } else if (nm == "verletlist") {
      // plugins/verletlist.cc:12-29; extcut = 1.08*cutoff (the unread extcutoff
      // keyword stays unread), `skin` switches to a skin-triggered list
      Keywords k(args, {{"cutoff", "0.0"}, {"extcutoff", "0.0"}, {"maxneighbors", "80"}, {"each", "5"}});
      c.kind = kCmVerlet;
      c.cutoff = k.d("cutoff");
      c.maxneighbors = k.i("maxneighbors");
      c.each = std::max(1, k.i("each"));
      c.extcut = 1.08 * c.cutoff;
      if (k.has("skin")) {
        c.skin = k.d("skin");
        if (!(c.skin >= 0.0))
          fail(LPMDX_EINVAL, "verletlist: skin must be non-negative");
      }
      if (fabs(c.cutoff) < 1E-10)
        fail(LPMDX_EINVAL, "verletlist: No cutoff was defined");
    } else if (nm == "minimumimage") {
      // plugins/minimumimage.cc:12-19; `skin` is this engine's extension, as
      // for verletlist (without it the lists are rebuilt on every UpdateCell)
      Keywords k(args, {{"cutoff", "0.0"}});
      c.kind = kCmMinImage;
      c.cutoff = k.d("cutoff");
      if (k.has("skin")) {
        c.skin = k.d("skin");
        if (!(c.skin >= 0.0))
          fail(LPMDX_EINVAL, "minimumimage: skin must be non-negative");
      }
    } else if (nm == "lcbinary") {
      // plugins/lcbinary.cc:9-33: automatic subdivisions when one is 0 or with mode auto
      Keywords k(args, {{"cutoff", "7.0"}, {"nx", "0"}, {"ny", "0"}, {"nz", "0"}, {"mode", "noauto"}});
      c.kind = kCmLcBinary;
      c.cutoff = k.d("cutoff");
      c.nx = k.i("nx");
      c.ny = k.i("ny");
      c.nz = k.i("nz");
      const std::string mode = k.s("mode");
      c.mode_auto = c.nx == 0 || c.ny == 0 || c.nz == 0 || mode == "auto" || mode == "AUTO";
    } else {
      fail(LPMDX_EINVAL, "unknown cell manager '" + nm + "'");
    }
    S->undrift();
    S->cm = c;
    S->list_built = false;
    S->forces_ok = false;
    S->update_cell(); // Configuration::SetCellManager calls UpdateCell
  });
}

int lpmdx_update_cell(lpmdx_system* S) {
  return guard([&] {
    need(S, "system");
    S->update_cell();
  });
}

int lpmdx_potential_append(lpmdx_system* S, const char* name, const char* args, int s1, int s2) {
  return guard([&] {
    need(S, "system");
    need(name, "name");
    S->pots.push_back(make_potential(name, args, s1, s2));
    S->forces_ok = false;
    S->lists_follow_potentials();
    S->undrift();
  });
}

int lpmdx_potential_clear(lpmdx_system* S) {
  return guard([&] {
    need(S, "system");
    S->pots.clear();
    S->forces_ok = false;
    S->lists_follow_potentials();
    S->undrift();
  });
}

int lpmdx_update_forces(lpmdx_system* S, double* epot, double* virial, double stress[9]) {
  return guard([&] {
    need(S, "system");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (S->pots.empty())
      fail(LPMDX_ESTATE, "no potential set");
    if (S->n == 0) {
      if (epot)
        *epot = 0.0;
      if (virial)
        *virial = 0.0;
      if (stress)
        for (int k = 0; k < 9; ++k)
          stress[k] = 0.0;
      return;
    }
    S->update_cell();
    S->forces(kForcesOnly, true, 0.0, S->ctrl.p->maxd2);
    S->read_energies(epot, virial, stress);
    S->forces_ok = true;
  });
}

int lpmdx_potential_forces(lpmdx_system* S, int k, double* acc, double* epot, double* virial, double stress[9]) {
  return guard([&] {
    need(S, "system");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (k < 0 || k >= (int)S->pots.size())
      fail(LPMDX_EINVAL, "no potential with index " + std::to_string(k));
    double e = 0.0, w = 0.0, t[9] = {0};
    if (S->n > 0) {
      S->undrift();
      if (!S->list_built || S->cell_version != S->pos_version)
        S->update_cell();
      S->forces_term(k);
      const double* pr = S->hred;
      e = pr[0] + pr[12 + 1];
      w = pr[2];
      for (int q = 0; q < 9; ++q)
        t[q] = pr[3 + q];
      if (acc) {
        const double* src[3] = {S->asum.p, nullptr, nullptr};
        double* const dst[3] = {acc, nullptr, nullptr};
        S->download(src, dst);
      }
    }
    if (epot)
      *epot = e;
    if (virial)
      *virial = w;
    if (stress)
      for (int q = 0; q < 9; ++q)
        stress[q] = t[q];
  });
}

int lpmdx_velocityverlet(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t,
                         double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->vv(dt, steps, obs_every, epot_t, ekin_t);
  });
}

int lpmdx_leapfrog_init(lpmdx_system* S, double dt) {
  return guard([&] {
    need(S, "system");
    S->lf_init(dt);
  });
}

int lpmdx_leapfrog(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t, double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->lf(dt, steps, obs_every, epot_t, ekin_t);
  });
}

int lpmdx_beeman_init(lpmdx_system* S, double dt) {
  return guard([&] {
    need(S, "system");
    S->bm_init(dt);
  });
}

int lpmdx_beeman(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t, double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->vv(dt, steps, obs_every, epot_t, ekin_t, true);
  });
}

int lpmdx_nosehoover_init(lpmdx_system* S, double dt, double fmass, double t, double kboltzmann) {
  return guard([&] {
    need(S, "system");
    S->nh_init(dt, fmass, t, kboltzmann);
  });
}

int lpmdx_nosehoover(lpmdx_system* S, double dt, long steps, long obs_every, double* epot_t, double* ekin_t) {
  return guard([&] {
    need(S, "system");
    if (steps < 0 || obs_every < 0)
      fail(LPMDX_EINVAL, "negative step count");
    S->vv(dt, steps, obs_every, epot_t, ekin_t, false, true);
  });
}

int lpmdx_set_temperature(lpmdx_system* S, double temp) {
  return guard([&] {
    need(S, "system");
    if (!std::isfinite(temp) || temp < 0.0)
      fail(LPMDX_EINVAL, "temperature must be finite and non-negative");
    if (S->n == 0)
      return;
    S->undrift(); // the look-ahead positions were made from the old velocities
    S->partials.alloc(std::max<size_t>((size_t)blocks_for(S->n) * 4, S->partials.n));
    if (!S->velocities_set)
      S->draw_velocities();
    S->scale_to_temperature(temp);
    S->sync();
  });
}

int lpmdx_atom_energies(lpmdx_system* S, int k, double* e) {
  return guard([&] {
    need(S, "system");
    need(e, "e");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (k < 0 || k >= (int)S->pots.size())
      fail(LPMDX_EINVAL, "no potential with index " + std::to_string(k));
    if (S->n == 0)
      return;
    S->undrift();
    if (S->cell_version != S->pos_version || !S->list_built)
      S->update_cell();
    S->eatom.alloc(S->n);
    S->forces_term(k, S->eatom.p);
    const long n = S->n;
    std::vector<int> perm(n);
    std::vector<double> tmp(n);
    hipchk(hipMemcpyAsync(perm.data(), S->perm.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(tmp.data(), S->eatom.p, n * sizeof(double), hipMemcpyDeviceToHost, S->st), "memcpy");
    S->sync();
    for (long i = 0; i < n; ++i)
      e[perm[i]] = tmp[i];
  });
}

int lpmdx_seed48(lpmdx_system* S, long seed) {
  return guard([&] {
    need(S, "system");
    S->rng48.seed(seed);
  });
}

int lpmdx_init_velocities(lpmdx_system* S) {
  return guard([&] {
    need(S, "system");
    if (S->n == 0) {
      S->velocities_set = true;
      return;
    }
    S->undrift(); // the look-ahead positions were made from the old velocities
    S->draw_velocities();
    S->sync();
  });
}

int lpmdx_velocities_set(lpmdx_system* S, int set) {
  return guard([&] {
    need(S, "system");
    S->velocities_set = set != 0;
  });
}

int lpmdx_system_set_fixed(lpmdx_system* S, const int* flags) {
  return guard([&] {
    need(S, "system");
    const long n = S->n;
    long tagged = 0;
    bool any = false;
    if (flags)
      for (long i = 0; i < n; ++i) {
        if (flags[i] & ~(LPMDX_FIX_VEL | LPMDX_FIX_POS | LPMDX_FIX_TAGGED))
          fail(LPMDX_EINVAL, "fixed-atom flags: unknown bit");
        if ((flags[i] & (LPMDX_FIX_VEL | LPMDX_FIX_POS)) && !(flags[i] & LPMDX_FIX_TAGGED))
          fail(LPMDX_EINVAL, "fixed-atom flags: a fixedvel or fixedpos atom carries the tag (LPMDX_FIX_TAGGED)");
        any = any || (flags[i] & (LPMDX_FIX_VEL | LPMDX_FIX_POS)) != 0;
        tagged += (flags[i] & LPMDX_FIX_TAGGED) ? 1 : 0;
      }
    S->undrift(); // a look-ahead was made without these flags
    if (flags)
      S->hfix.assign(flags, flags + n);
    else
      S->hfix.clear();
    bool anypos = false;
    if (flags)
      for (long i = 0; i < n; ++i)
        anypos = anypos || (flags[i] & LPMDX_FIX_POS) != 0;
    S->fixpos_any = anypos;
    if (anypos) {
      // the Nose-Hoover friction grows once per moving atom, in caller order
      std::vector<int> rank(n);
      long moving = 0;
      for (long i = 0; i < n; ++i) {
        rank[i] = (int)moving;
        moving += (flags[i] & LPMDX_FIX_POS) ? 0 : 1;
      }
      S->mv_rank.alloc(n);
      hipchk(hipMemcpyAsync(S->mv_rank.p, rank.data(), n * sizeof(int), hipMemcpyHostToDevice, S->st), "memcpy");
      S->sync();
      S->n_moving = moving;
    }
    if (tagged == 0) {
      S->any_fix = false;
      S->n_tagged = 0;
      return;
    }
    std::vector<int> perm(n), f(n);
    hipchk(hipMemcpyAsync(perm.data(), S->perm.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    S->sync();
    for (long i = 0; i < n; ++i)
      f[i] = flags[perm[i]];
    S->fix.alloc(n);
    hipchk(hipMemcpyAsync(S->fix.p, f.data(), n * sizeof(int), hipMemcpyHostToDevice, S->st), "memcpy");
    S->sync();
    S->any_fix = any;
    S->n_tagged = tagged;
  });
}

int lpmdx_tempscaling(lpmdx_system* S, double from, double to, long start, long end, long each) {
  return guard([&] {
    need(S, "system");
    if (each == 0) {
      S->tsched.on = false;
      return;
    }
    if (!std::isfinite(from) || !std::isfinite(to) || from < 0.0 || to < 0.0)
      fail(LPMDX_EINVAL, "tempscaling: temperatures must be finite and non-negative");
    if (each < 0 || start < 0 || end < -1 || (end != -1 && end < start))
      fail(LPMDX_EINVAL, "tempscaling: need each >= 1, start >= 0 and end == -1 or end >= start");
    if (end == -1 && from != to)
      fail(LPMDX_EINVAL, "tempscaling: a ramp (from != to) needs its last step as end");
    S->tsched.on = true;
    S->tsched.from = from;
    S->tsched.to = to;
    S->tsched.start = start;
    S->tsched.end = end;
    S->tsched.each = each;
  });
}

int lpmdx_berendsen(lpmdx_system* S, double from, double to, double tau, long start, long end, long each) {
  return guard([&] {
    need(S, "system");
    if (each == 0) {
      S->bsched = {};
      return;
    }
    if (!std::isfinite(from) || !std::isfinite(to) || from < 0.0 || to < 0.0 || !std::isfinite(tau) || !(tau > 0.0))
      fail(LPMDX_EINVAL, "berendsen: temperatures must be finite and non-negative, tau positive");
    if (each < 0 || start < 0 || end < -1 || (end != -1 && end < start))
      fail(LPMDX_EINVAL, "berendsen: need each >= 1, start >= 0 and end == -1 or end >= start");
    if (end == -1 && from != to)
      fail(LPMDX_EINVAL, "berendsen: a ramp (from != to) needs its last step as end");
    S->bsched = {};
    S->bsched.on = true;
    S->bsched.from = from;
    S->bsched.to = to;
    S->bsched.tau = tau;
    S->bsched.start = start;
    S->bsched.end = end;
    S->bsched.each = each;
  });
}

namespace {

// SetSimulationTags (simulationbuilder.cc:84-124) from the tag-aware kinetic
// energy ke and momentum p of the system's current velocities and the
// energies of its last force evaluation.
void fill_observables(const lpmdx_system* S, double ke, const double p[3], double out[LPMDX_NOBS]) {
  {
    const double vol = S->ll[0] * S->ll[1] * S->ll[2]; // OrthogonalCell::Volume (orthogonalcell.h:168)
    out[LPMDX_OBS_KINETIC] = ke;
    out[LPMDX_OBS_TEMP] = (2.0 / 3.0) * ke / (kBoltzmann * double(S->n - S->n_tagged));
    for (int q = 0; q < 3; ++q)
      out[LPMDX_OBS_PX + q] = p[q];
    out[LPMDX_OBS_MOMENTUM] = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    out[LPMDX_OBS_VOLUME] = vol;
    out[LPMDX_OBS_VOLUME_PER_ATOM] = vol / double(S->n);
    for (int q = 0; q < 3; ++q)
      out[LPMDX_OBS_CELL_A + q] = S->ll[q];
    out[LPMDX_OBS_PARTICLE_DENSITY] = double(S->n) / vol;
    double mtot = 0.0;
    for (long i = 0; i < S->n; ++i)
      mtot += S->hmass[i];
    out[LPMDX_OBS_DENSITY] = kUa3ToGrCm3 * (mtot / vol);
    out[LPMDX_OBS_KINETIC_PRESSURE] = (kPressFactor / vol) * (2.0 / 3.0) * ke;
    if (S->last_e_valid) {
      out[LPMDX_OBS_POTENTIAL] = S->last_e[0];
      out[LPMDX_OBS_TOTAL] = S->last_e[0] + ke;
      out[LPMDX_OBS_VIRIAL] = S->last_e[1];
      out[LPMDX_OBS_VIRIAL_PRESSURE] = (kPressFactor / vol) * (1.0 / 3.0) * S->last_e[1];
      out[LPMDX_OBS_PRESSURE] = (kPressFactor / vol) * ((2.0 / 3.0) * ke + (1.0 / 3.0) * S->last_e[1]);
      for (int k = 0; k < 9; ++k)
        out[LPMDX_OBS_STRESS + k] = (kPressFactor / vol) * S->last_e[2 + k];
    }
  }
}

// Scratch of lpmdx_observables_noalloc: the per-block kinetic partials (four
// per block of kBlock atoms) and their reduced sums, for up to
// kNoAllocAtoms atoms.  Each GPU holds its own copy; host code reaches it
// only through the address hipGetSymbolAddress returns for that device.
constexpr int kMaxDevices = 64;
constexpr long kNoAllocAtoms = 1L << 22;
constexpr long kNoAllocRows = kNoAllocAtoms / kBlock;
static __device__ __align__(256) double scratch[4 * kNoAllocRows + 8];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

// The device address of `scratch` on `device` (looked up once per device,
// with that device current).
double* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices)
    fail(LPMDX_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    hipchk(hipGetDevice(&cur), "hipGetDevice");
    hipchk(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    hipchk(hipSetDevice(cur), "hipSetDevice");
    hipchk(rc, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}

} // namespace

int lpmdx_observables(lpmdx_system* S, double out[LPMDX_NOBS]) {
  return guard([&] {
    need(S, "system");
    need(out, "out");
    for (int k = 0; k < LPMDX_NOBS; ++k)
      out[k] = 0.0;
    if (S->n == 0)
      return;
    // tag-aware sums when some atom carries fixedvel or fixedpos
    double p[3];
    const double ke = S->kinetic(p, true);
    fill_observables(S, ke, p, out);
  });
}

int lpmdx_observables_noalloc(lpmdx_system* S, double out[LPMDX_NOBS]) {
  return guard([&] {
    need(S, "system");
    need(out, "out");
    for (int k = 0; k < LPMDX_NOBS; ++k)
      out[k] = 0.0;
    if (S->n == 0)
      return;
    if (S->n > kNoAllocAtoms)
      fail(LPMDX_EINVAL, "lpmdx_observables_noalloc: " + std::to_string(S->n) + " atoms, at most " +
                             std::to_string(kNoAllocAtoms));
    double* sp = scratch_on(S->device);
    // the kernels and the reduction of lpmdx_observables (the same sums in the
    // same order), with the partials in the scratch array; every partial row
    // and the four sums are written before they are read, so nothing needs
    // clearing
    const long nb = blocks_for(S->n);
    const int* fx = S->n_tagged > 0 ? S->fix.p : nullptr;
    double* sums = sp + 4 * kNoAllocRows;
    k_kinetic<<<nb, kBlock, 0, S->st>>>(S->v.p, S->mass.p, S->n, sp, fx, kFixTagged);
    k_reduce_partials<<<1, kBlock, 0, S->st>>>(sp, nb, 4, sums);
    S->launch_check();
    hipchk(hipMemcpyAsync(S->hred, sums, 4 * sizeof(double), hipMemcpyDeviceToHost, S->st), "memcpy");
    S->sync();
    const double p[3] = {S->hred[1], S->hred[2], S->hred[3]};
    fill_observables(S, S->hred[0] * kKin2eV, p, out);
  });
}

int lpmdx_system_stats(const lpmdx_system* S, long out[9]) {
  return guard([&] {
    need(S, "system");
    need(out, "out");
    out[0] = S->n_builds;
    out[1] = S->n_forces;
    out[2] = S->n_steps;
    out[3] = S->cap;
    out[4] = S->max_list;
    out[5] = S->lpa;
    out[6] = S->staged ? S->grid.brick : 0;
    out[7] = S->sum_list;
    out[8] = S->n_cut_chunks;
  });
}

int lpmdx_system_profile(lpmdx_system* S, int enable) {
  return guard([&] {
    need(S, "system");
    if (enable < 0)
      fail(LPMDX_EINVAL, "profile: enable must be 0 (off) or the step sampling period (>= 1)");
    S->sync();
    S->profiling = enable != 0;
    S->prof_every = enable > 0 ? enable : 1;
    for (int k = 0; k < kProfSlots; ++k) {
      S->prof_ms[k] = 0.0;
      S->prof_n[k] = 0;
    }
  });
}

int lpmdx_system_profile_read(lpmdx_system* S, double ms[4], long launches[4]) {
  return guard([&] {
    need(S, "system");
    need(ms, "ms");
    need(launches, "launches");
    S->sync();
    for (int k = 0; k < kProfSlots; ++k) {
      ms[k] = S->prof_ms[k];
      launches[k] = S->prof_n[k];
    }
  });
}

int lpmdx_system_set_lanes(lpmdx_system* S, int lanes) {
  return guard([&] {
    need(S, "system");
    if (!(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16))
      fail(LPMDX_EINVAL, "lanes per atom must be 0, 1, 2, 4, 8 or 16");
    S->lpa_req = lanes;
    S->list_built = false;
    S->cap = 0;
  });
}

int lpmdx_neighbor_lists(lpmdx_system* S, int full, double rcut, long* counts, long cap, long* js,
                         double* r2, long* total) {
  return guard([&] {
    need(S, "system");
    need(counts, "counts");
    need(total, "total");
    if (S->cm.kind == kNone)
      fail(LPMDX_ESTATE, "cellmanager not set");
    if (!S->list_built)
      S->update_cell();
    S->undrift();
    const long n = S->n;
    if (n == 0) {
      *total = 0;
      return;
    }
    if (S->cm.kind == kCmMinImage && rcut < 1.0E-10)
      fail(LPMDX_EINVAL, "MinimumImage cutoff equals zero..."); // minimumimage.cc:76-77
    if (S->cm.kind == kCmLinkedCell && rcut > S->rl) {
      // the cached lists are built on the first request's rcut (linkedcell.cc:315-354)
      S->rl = rcut;
      S->build_list();
    }
    if ((S->cm.kind == kCmMinImage || S->cm.kind == kCmLcBinary) && rcut > S->rl) {
      // these list afresh with the caller's radius: lists of that radius (the
      // next UpdateCell returns to the potentials' radius)
      S->rl_min = rcut;
      S->rebuild();
      S->rl_min = 0.0;
    }
    std::vector<int> perm(n), cntv(n), nb((size_t)S->cap * S->stride);
    std::vector<double> x(3 * n);
    hipchk(hipMemcpyAsync(perm.data(), S->perm.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(cntv.data(), S->cnt.p, n * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(nb.data(), S->nbr.p, nb.size() * sizeof(int), hipMemcpyDeviceToHost, S->st), "memcpy");
    hipchk(hipMemcpyAsync(x.data(), S->x.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, S->st), "memcpy");
    const int H = S->staged ? halo_cells(S->grid.brick) : 0;
    std::vector<int2> bh(S->staged ? (size_t)S->nbrick * (H + 1) : 0);
    std::vector<uint32_t> keys(S->staged ? n : 0);
    if (S->staged) {
      hipchk(hipMemcpyAsync(bh.data(), S->bhalo.p, bh.size() * sizeof(int2), hipMemcpyDeviceToHost, S->st),
             "memcpy");
      hipchk(hipMemcpyAsync(keys.data(), S->keys.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, S->st), "memcpy");
    }
    S->sync();
    // staged lists hold slots of the atom's brick halo: back to sorted indices
    const int B3 = S->grid.brick * S->grid.brick * S->grid.brick;
    auto atom_of = [&](long i, long e) -> long {
      if (!S->staged)
        return e;
      const int2* row = bh.data() + (size_t)(keys[i] / B3) * (H + 1);
      int h = 0; // the last halo cell whose first slot is <= e
      while (h + 1 < H && row[h + 1].x <= e)
        ++h;
      return (long)row[h].y + (e - row[h].x);
    };
    std::vector<long> iperm(n);
    for (long i = 0; i < n; ++i)
      iperm[perm[i]] = i;
    const double rc2 = rcut * rcut;
    long k = 0;
    std::vector<std::pair<long, double>> row;
    for (long o = 0; o < n; ++o) {
      const long i = iperm[o];
      row.clear();
      for (int c = 0; c < cntv[i]; ++c) {
        const long jj = atom_of(i, nb[list_index(i, c, S->lpa, S->cap)]);
        const long oj = perm[jj];
        if (!full && S->stencil_grid() && oj > o)
          continue; // linkedcell / lcbinary half lists keep j < i (linkedcell.cc:323, lcbinary.cc:291)
        if (!full && (S->cm.kind == kCmVerlet || S->cm.kind == kCmMinImage) && oj < o)
          continue; // verletlist / minimumimage half lists keep j > i (verletlist.cc:157, minimumimage.cc:90)
        double d[3];
        for (int q = 0; q < 3; ++q) {
          d[q] = x[3 * jj + q] - x[3 * i + q];
          if (d[q] >= S->box.half[q])
            d[q] -= S->box.L[q];
          else if (d[q] < -S->box.half[q])
            d[q] += S->box.L[q];
        }
        const double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (rr < rc2)
          row.emplace_back(oj, rr);
      }
      if (S->cm.kind == kCmVerlet || S->cm.kind == kCmMinImage)
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<long, double>& p, const std::pair<long, double>& q) {
                           return p.first < q.first;
                         });
      counts[o] = (long)row.size();
      for (auto& e : row) {
        if (k < cap) {
          if (js)
            js[k] = e.first;
          if (r2)
            r2[k] = e.second;
        }
        ++k;
      }
    }
    *total = k;
  });
}

} // extern "C"

// ---------------------------------------------------------------- snapshots
// An asynchronous copy of a system's state for checkpoints and trajectory
// files (Dump, lpmd / zlp, xyz; lpmd2_amd/csrc/formats.cpp): x, v and a are
// gathered into caller order on the system's stream, into the snapshot's own
// device buffer, and copied to pinned host memory on the snapshot's stream,
// so the system's next steps run while the copy is in flight.
struct lpmdx_snapshot {
  int device = -1;
  long n = 0, step = 0;
  double cell[9] = {0};
  std::vector<int> z, fix;
  DBuf<double> d;
  double* h = nullptr;
  size_t hn = 0;
  hipStream_t cs = nullptr;
  hipEvent_t ready = nullptr, done = nullptr;
  bool pending = false;

  ~lpmdx_snapshot() {
    if (pending && done)
      (void)hipEventSynchronize(done);
    if (cs)
      (void)hipStreamSynchronize(cs);
    if (h)
      (void)hipHostFree(h);
    if (ready)
      (void)hipEventDestroy(ready);
    if (done)
      (void)hipEventDestroy(done);
    if (cs)
      (void)hipStreamDestroy(cs);
  }

  void wait() {
    if (pending) {
      hipchk(hipEventSynchronize(done), "hipEventSynchronize");
      pending = false;
    }
  }
};

extern "C" {

int lpmdx_snapshot_create(lpmdx_snapshot** out) {
  return guard([&] {
    need(out, "out");
    *out = new lpmdx_snapshot();
  });
}

int lpmdx_snapshot_destroy(lpmdx_snapshot* P) {
  return guard([&] { delete P; });
}

int lpmdx_snapshot_take(lpmdx_system* S, lpmdx_snapshot* P, long step) {
  return guard([&] {
    need(S, "system");
    need(P, "snapshot");
    // the previous copy lands before its buffers are reused
    P->wait();
    if (P->device >= 0 && P->device != S->device)
      fail(LPMDX_EINVAL, "snapshot: taken on another device before");
    P->device = S->device;
    if (!P->cs) {
      hipchk(hipStreamCreateWithFlags(&P->cs, hipStreamNonBlocking), "hipStreamCreate");
      hipchk(hipEventCreateWithFlags(&P->ready, hipEventDisableTiming), "hipEventCreate");
      hipchk(hipEventCreateWithFlags(&P->done, hipEventDisableTiming), "hipEventCreate");
    }
    const long n = S->n;
    P->n = n;
    P->step = step;
    for (int k = 0; k < 9; ++k)
      P->cell[k] = (k % 4 == 0) ? S->box.L[k / 4] : 0.0;
    P->z = S->hz;
    P->fix = S->hfix;
    if (n == 0)
      return;
    const size_t m = (size_t)9 * n;
    P->d.alloc(m);
    if (P->hn < m) {
      if (P->h)
        (void)hipHostFree(P->h);
      P->h = nullptr;
      P->hn = 0;
      hipchk(hipHostMalloc(&P->h, m * sizeof(double)), "hipHostMalloc");
      P->hn = m;
    }
    const double* src[3] = {S->drifted ? S->xn.p : S->x.p, S->v.p, S->a.p};
    for (int j = 0; j < 3; ++j)
      k_to_caller<<<blocks_for(n), kBlock, 0, S->st>>>(src[j], S->perm.p, n, P->d.p + (size_t)3 * n * j);
    S->launch_check();
    hipchk(hipEventRecord(P->ready, S->st), "hipEventRecord");
    hipchk(hipStreamWaitEvent(P->cs, P->ready, 0), "hipStreamWaitEvent");
    hipchk(hipMemcpyAsync(P->h, P->d.p, m * sizeof(double), hipMemcpyDeviceToHost, P->cs), "memcpy");
    hipchk(hipEventRecord(P->done, P->cs), "hipEventRecord");
    P->pending = true;
  });
}

int lpmdx_snapshot_wait(lpmdx_snapshot* P) {
  return guard([&] {
    need(P, "snapshot");
    P->wait();
  });
}

int lpmdx_snapshot_info(lpmdx_snapshot* P, long* n, long* step, double cell[9]) {
  return guard([&] {
    need(P, "snapshot");
    P->wait();
    if (n)
      *n = P->n;
    if (step)
      *step = P->step;
    if (cell)
      memcpy(cell, P->cell, sizeof P->cell);
  });
}

int lpmdx_snapshot_state(lpmdx_snapshot* P, int* z, double* x, double* v, double* a, int* flags) {
  return guard([&] {
    need(P, "snapshot");
    P->wait();
    const size_t n = (size_t)P->n;
    if (!n)
      return;
    if (z)
      memcpy(z, P->z.data(), n * sizeof(int));
    double* dst[3] = {x, v, a};
    for (int j = 0; j < 3; ++j)
      if (dst[j])
        memcpy(dst[j], P->h + 3 * n * j, 3 * n * sizeof(double));
    if (flags) {
      if (P->fix.size() == n)
        memcpy(flags, P->fix.data(), n * sizeof(int));
      else
        memset(flags, 0, n * sizeof(int));
    }
  });
}

} // extern "C"

"""ctypes front end of oracle/_ref/libref_harness.so: the reference's own
linkedcell, verletlist, potential, crystal3d and velocityverlet plugin code,
compiled from its sources by oracle/ref.mk.  Test infrastructure only; absent
unless that recipe has been run (it needs the reference sources)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_lib = None


def available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libref_harness.so"))


def so(name: str) -> bytes:
    return os.path.join(REF_DIR, name + ".so").encode()


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(os.path.join(REF_DIR, "libref_harness.so"))
        vp, d, l, i = C.c_void_p, C.c_double, C.c_long, C.c_int
        L.ref_last_error.restype = C.c_char_p
        L.ref_conf_new.restype = vp
        L.ref_conf_new.argtypes = [l, vp, vp, vp]
        L.ref_conf_set_positions.argtypes = [vp, vp]
        L.ref_conf_set_velocities.argtypes = [vp, vp]
        L.ref_conf_get.argtypes = [vp, vp, vp, vp]
        L.ref_conf_free.argtypes = [vp]
        L.ref_cm_open.restype = vp
        L.ref_cm_open.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_verlet_open.restype = vp
        L.ref_verlet_open.argtypes = [C.c_char_p]
        L.ref_cm_update.argtypes = [vp, vp, i]
        L.ref_cm_lists.argtypes = [vp, vp, i, d, l, vp, vp, vp, vp, vp]
        L.ref_cm_close.argtypes = [vp]
        L.ref_pot_open.restype = vp
        L.ref_pot_open.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_pot_cutoff.restype = d
        L.ref_pot_cutoff.argtypes = [vp]
        L.ref_pair_eval.argtypes = [vp, l, vp, vp, vp]
        L.ref_metal_eval.argtypes = [vp, l, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ref_pot_close.argtypes = [vp]
        L.ref_crystal.argtypes = [C.c_char_p, C.c_char_p, l, vp, vp]
        L.ref_vv_run.argtypes = [C.c_char_p, C.c_char_p, vp, vp, l, l, vp]
        L.ref_plugin_open.restype = vp
        L.ref_plugin_open.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_plugin_close.argtypes = [vp]
        L.ref_combined_forces.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.ref_integrate.argtypes = [vp, vp, i, vp, vp, vp, l, vp]
        L.ref_atom_energies.argtypes = [vp, vp, i, i, l, vp]
        L.ref_set_atom_tag.argtypes = [vp, l, C.c_char_p, C.c_char_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().ref_last_error().decode())


class RefConf:
    def __init__(self, x, L, z=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self.n = self.x.shape[0]
        self.L = np.ascontiguousarray(L, dtype=np.float64)
        zz = None if z is None else np.ascontiguousarray(np.broadcast_to(z, (self.n,)), dtype=np.int32)
        self.h = lib().ref_conf_new(self.n, _p(self.L), _p(self.x), None if zz is None else _p(zz))
        if not self.h:
            raise RuntimeError(lib().ref_last_error().decode())

    def set_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().ref_conf_set_positions(self.h, _p(x)))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(lib().ref_conf_set_velocities(self.h, _p(v)))

    def get(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        a = np.zeros((self.n, 3))
        _check(lib().ref_conf_get(self.h, _p(x), _p(v), _p(a)))
        return x, v, a

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_conf_free(self.h)
            self.h = None


class RefCellManager:
    def __init__(self, name: str, args: str):
        if name == "verletlist":
            self.h = lib().ref_verlet_open((args + " debug none").encode())
        else:
            self.h = lib().ref_cm_open(so(name), (args + " debug none").encode())
        if not self.h:
            raise RuntimeError(lib().ref_last_error().decode())

    def update(self, conf: RefConf, attach=False):
        _check(lib().ref_cm_update(self.h, conf.h, int(attach)))

    def lists(self, conf: RefConf, full: bool, rcut: float, cap=None):
        n = conf.n
        cap = cap or n * 400
        counts = np.zeros(n, dtype=np.int64)
        js = np.zeros(cap, dtype=np.int64)
        r2 = np.zeros(cap)
        rij = np.zeros((cap, 3))
        tot = np.zeros(1, dtype=np.int64)
        _check(lib().ref_cm_lists(self.h, conf.h, int(full), rcut, cap, _p(counts), _p(js), _p(r2), _p(rij), _p(tot)))
        t = int(tot[0])
        if t > cap:
            return self.lists(conf, full, rcut, cap=t)
        return counts, js[:t], r2[:t], rij[:t]

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_cm_close(self.h)
            self.h = None


class RefPot:
    """A potential plugin loaded through its create(args): the reference's of
    that name, or the .so at `path` (a drop-in)."""

    def __init__(self, name: str, args: str, path: str | None = None):
        self.name = name
        self.h = lib().ref_pot_open(path.encode() if path else so(name), (args + " debug none").encode())
        if not self.h:
            raise RuntimeError(lib().ref_last_error().decode())

    @property
    def cutoff(self):
        return lib().ref_pot_cutoff(self.h)

    def pair(self, rij):
        rij = np.ascontiguousarray(np.atleast_2d(rij), dtype=np.float64)
        m = rij.shape[0]
        e = np.zeros(m)
        f = np.zeros((m, 3))
        _check(lib().ref_pair_eval(self.h, m, _p(rij), _p(e), _p(f)))
        return e, f

    def metal(self, rij, rhoi, rhoj):
        rij = np.ascontiguousarray(np.atleast_2d(rij), dtype=np.float64)
        m = rij.shape[0]
        rhoi = np.ascontiguousarray(np.broadcast_to(rhoi, (m,)), dtype=np.float64)
        rhoj = np.ascontiguousarray(np.broadcast_to(rhoj, (m,)), dtype=np.float64)
        ep, rho, fr = np.zeros(m), np.zeros(m), np.zeros(m)
        pf, mb = np.zeros((m, 3)), np.zeros((m, 3))
        _check(lib().ref_metal_eval(self.h, m, _p(rij), _p(rhoi), _p(rhoj), _p(ep), _p(rho), _p(fr), _p(pf), _p(mb)))
        return ep, rho, fr, pf, mb

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_pot_close(self.h)
            self.h = None


def crystal(args: str, n: int, L):
    L = np.ascontiguousarray(L, dtype=np.float64)
    pos = np.zeros((n, 3))
    _check(lib().ref_crystal(so("crystal3d"), (args + " debug none").encode(), n, _p(L), _p(pos)))
    return pos


def vv_run(args: str, conf: RefConf, cm: RefCellManager, steps: int, acc_tables, integrator="velocityverlet",
           path=None):
    """Initialize + `steps` Advance calls of an integrator plugin (the
    reference's of that name, or the .so at `path`), its force calls
    replaying acc_tables in order."""
    acc = np.ascontiguousarray(acc_tables, dtype=np.float64)
    ncalls = acc.shape[0]
    lp = path.encode() if path else so(integrator)
    _check(lib().ref_vv_run(lp, (args + " debug none").encode(), conf.h, cm.h, steps, ncalls, _p(acc)))


# ---- plugins loaded through the reference's create(args), and the reference's
# ---- composition driving them (the MI355X drop-in plugins among them)
DROPIN_DIR = os.path.join(os.path.dirname(_HERE), "lpmd2_amd", "plugins", "_build")


def dropin_available() -> bool:
    return available() and os.path.exists(os.path.join(DROPIN_DIR, "liblpmdx_lpmd.so"))


class RefPlugin:
    """A plugin object made by the create() of a .so file (the reference's
    own from oracle/_ref, or a drop-in from lpmd2_amd/plugins/_build)."""

    def __init__(self, path: str, args: str):
        self.path = path
        self.h = lib().ref_plugin_open(path.encode(), (args + " debug none").encode())
        if not self.h:
            raise RuntimeError(lib().ref_last_error().decode())

    @classmethod
    def reference(cls, name: str, args: str):
        return cls(so(name).decode(), args)

    @classmethod
    def dropin(cls, name: str, args: str):
        return cls(os.path.join(DROPIN_DIR, name + ".so"), args)

    def attach(self, conf: RefConf):
        """Configuration::SetCellManager(this) (runs its UpdateCell)."""
        _check(lib().ref_cm_update(self.h, conf.h, 1))

    def update(self, conf: RefConf):
        _check(lib().ref_cm_update(self.h, conf.h, 0))

    def lists(self, conf: RefConf, full: bool, rcut: float, cap=None):
        return RefCellManager.lists(self, conf, full, rcut, cap)

    def close(self):
        if getattr(self, "h", None):
            lib().ref_plugin_close(self.h)
            self.h = None


def _handles(pots):
    arr = (C.c_void_p * len(pots))(*[p.h for p in pots])
    return C.cast(arr, C.c_void_p), arr


def _species(pots, species):
    s1 = np.array([s[0] for s in species] if species else [-1] * len(pots), dtype=np.int32)
    s2 = np.array([s[1] for s in species] if species else [-1] * len(pots), dtype=np.int32)
    return s1, s2


def combined_forces(conf: RefConf, pots, species=None):
    """CombinedPotential::UpdateForces over loaded potential plugins on conf
    (whose cell manager is attached): (acc, energy, virial, stress[3x3])."""
    hp, keep = _handles(pots)
    s1, s2 = _species(pots, species)
    acc = np.zeros((conf.n, 3))
    out = np.zeros(11)
    _check(lib().ref_combined_forces(conf.h, len(pots), hp, _p(s1), _p(s2), _p(acc), _p(out)))
    return acc, float(out[0]), float(out[1]), out[2:].reshape(3, 3)


def atom_energies(pot: RefPlugin, conf: RefConf, s1: int = -1, s2: int = -1):
    """Potential::AtomEnergy(conf, i) of every atom through the plugin, its
    species rule (s1, s2) (-1, -1: every pair)."""
    e = np.zeros(conf.n)
    _check(lib().ref_atom_energies(pot.h, conf.h, s1, s2, conf.n, _p(e)))
    return e


def set_atom_tag(conf: RefConf, i: int, key: str, value: str = "true"):
    """BasicParticleSet::SetTag(atoms[i], key, value)."""
    _check(lib().ref_set_atom_tag(conf.h, int(i), key.encode(), value.encode()))


def integrate(integrator: RefPlugin, conf: RefConf, pots, steps: int, species=None):
    """Integrator::Initialize + `steps` Advance calls; the potential-energy tag per step."""
    hp, keep = _handles(pots)
    s1, s2 = _species(pots, species)
    ep = np.zeros(max(steps, 1))
    _check(lib().ref_integrate(integrator.h, conf.h, len(pots), hp, _p(s1), _p(s2), steps, _p(ep)))
    return ep[:steps]

"""MI355X-native force-and-integration engine for LPMD.

Python front end over the C ABI declared in include/lpmdx.h (liblpmdx.so,
hand-written HIP for gfx950).  The objects mirror the reference's plugin
interface: a System is a Configuration, `set_cellmanager(name, args)` is
Configuration::SetCellManager(LoadPluginAs<CellManager>(name, args)),
`append_potential(name, args, s1, s2)` is CombinedPotential::Append with
SetValidSpecies, `update_forces()` is CombinedPotential::UpdateForces and
`velocityverlet(dt, steps)` runs the lpmd loop with the velocityverlet
integrator.  There is no CPU fallback: without the built library or without a
GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build
from .lattice import fcc, maxwell_boltzmann, ELEMENT_MASS  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblpmdx.so")

OK, EINVAL, ENONFINITE, EDEVICE, ECAPACITY, ESTATE, ENOMEM = 0, -1, -2, -3, -4, -5, -6
NOBS = 28
# lpmdx_observables slots (include/lpmdx.h LPMDX_OBS_*), the SetSimulationTags tag names
OBS = {"kinetic-energy": 0, "temperature": 1, "px": 2, "py": 3, "pz": 4, "potential-energy": 5, "virial": 6,
       "pressure": 7, "total-energy": 17, "momentum": 18, "volume": 19, "volume-per-atom": 20,
       "particle-density": 21, "density": 22, "virial-pressure": 23, "kinetic-pressure": 24, "cell-a": 25,
       "cell-b": 26, "cell-c": 27}
for _q, _a in enumerate("xyz"):
    for _p_, _b in enumerate("xyz"):
        # the tag s<q><p> is pressfactor / V * stress[q][p]
        OBS[f"s{_a}{_b}"] = 8 + 3 * _q + _p_
FIX_VEL, FIX_POS, FIX_TAGGED = 1, 2, 4


class LpmdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load liblpmdx.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp, d, l, i, s = C.c_void_p, C.c_double, C.c_long, C.c_int, C.c_char_p
        L.lpmdx_last_error.restype = s
        L.lpmdx_abi_version.restype = i
        L.lpmdx_device_count.argtypes = [vp]
        L.lpmdx_system_create.argtypes = [vp, l, vp, vp, vp]
        L.lpmdx_system_destroy.argtypes = [vp]
        L.lpmdx_system_size.argtypes = [vp]
        L.lpmdx_system_size.restype = l
        L.lpmdx_system_set_state.argtypes = [vp, vp, vp, vp]
        L.lpmdx_system_get_state.argtypes = [vp, vp, vp, vp]
        L.lpmdx_set_cellmanager.argtypes = [vp, s, s]
        L.lpmdx_update_cell.argtypes = [vp]
        L.lpmdx_neighbor_lists.argtypes = [vp, i, d, vp, l, vp, vp, vp]
        L.lpmdx_potential_append.argtypes = [vp, s, s, i, i]
        L.lpmdx_potential_clear.argtypes = [vp]
        L.lpmdx_update_forces.argtypes = [vp, vp, vp, vp]
        L.lpmdx_velocityverlet.argtypes = [vp, d, l, l, vp, vp]
        L.lpmdx_leapfrog_init.argtypes = [vp, d]
        L.lpmdx_leapfrog.argtypes = [vp, d, l, l, vp, vp]
        L.lpmdx_beeman_init.argtypes = [vp, d]
        L.lpmdx_beeman.argtypes = [vp, d, l, l, vp, vp]
        L.lpmdx_set_temperature.argtypes = [vp, d]
        L.lpmdx_tempscaling.argtypes = [vp, d, d, l, l, l]
        L.lpmdx_berendsen.argtypes = [vp, d, d, d, l, l, l]
        L.lpmdx_observables.argtypes = [vp, vp]
        L.lpmdx_system_stats.argtypes = [vp, vp]
        L.lpmdx_system_set_lanes.argtypes = [vp, i]
        L.lpmdx_system_profile.argtypes = [vp, i]
        L.lpmdx_system_profile_read.argtypes = [vp, vp, vp]
        L.lpmdx_atom_energies.argtypes = [vp, i, vp]
        L.lpmdx_seed48.argtypes = [vp, l]
        L.lpmdx_init_velocities.argtypes = [vp]
        L.lpmdx_velocities_set.argtypes = [vp, i]
        L.lpmdx_system_set_fixed.argtypes = [vp, vp]
        L.lpmdx_nosehoover_init.argtypes = [vp, d, d, d, d]
        L.lpmdx_nosehoover.argtypes = [vp, d, l, l, vp, vp]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != OK:
        raise LpmdxError(rc, lib().lpmdx_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    c = C.c_int(0)
    _check(lib().lpmdx_device_count(C.byref(c)))
    return c.value


def _aos(a, n, what):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != 3 * n:
        raise ValueError(f"{what}: expected {n} x 3 values")
    return a.reshape(n, 3)


class System:
    """One Configuration resident in HBM (SimulationBuilder::CreateFixedOrthogonal)."""

    def __init__(self, x, box, z, mass=None, v=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        self.n = n
        self.box = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
        zz = np.ascontiguousarray(np.broadcast_to(np.asarray(z, dtype=np.int32), (n,)))
        mm = None if mass is None else np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        h = C.c_void_p()
        _check(lib().lpmdx_system_create(C.byref(h), n, _p(self.box), _p(zz), _p(mm)))
        self._h = h
        self.set_state(x=x, v=v)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().lpmdx_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x=None, v=None, a=None):
        xs = None if x is None else _aos(x, self.n, "x")
        vs = None if v is None else _aos(v, self.n, "v")
        as_ = None if a is None else _aos(a, self.n, "a")
        _check(lib().lpmdx_system_set_state(self._h, _p(xs), _p(vs), _p(as_)))

    def get_state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        a = np.zeros((self.n, 3))
        _check(lib().lpmdx_system_get_state(self._h, _p(x), _p(v), _p(a)))
        return x, v, a

    def set_cellmanager(self, name: str, args: str):
        _check(lib().lpmdx_set_cellmanager(self._h, name.encode(), args.encode()))

    def update_cell(self):
        _check(lib().lpmdx_update_cell(self._h))

    def append_potential(self, name: str, args: str, s1: int = -1, s2: int = -1):
        _check(lib().lpmdx_potential_append(self._h, name.encode(), args.encode(), s1, s2))

    def clear_potentials(self):
        _check(lib().lpmdx_potential_clear(self._h))

    def update_forces(self):
        e = C.c_double()
        w = C.c_double()
        S = np.zeros(9)
        _check(lib().lpmdx_update_forces(self._h, C.byref(e), C.byref(w), _p(S)))
        return e.value, w.value, S.reshape(3, 3)

    def velocityverlet(self, dt: float, steps: int, obs_every: int = 0):
        k = steps // obs_every if obs_every else 0
        ep = np.zeros(max(k, 1))
        ek = np.zeros(max(k, 1))
        _check(lib().lpmdx_velocityverlet(self._h, dt, steps, obs_every, _p(ep), _p(ek)))
        return ep[:k], ek[:k]

    def leapfrog_init(self, dt: float):
        """Leapfrog::Initialize: half-step velocities from the old configuration."""
        _check(lib().lpmdx_leapfrog_init(self._h, dt))

    def leapfrog(self, dt: float, steps: int, obs_every: int = 0):
        k = steps // obs_every if obs_every else 0
        ep = np.zeros(max(k, 1))
        ek = np.zeros(max(k, 1))
        _check(lib().lpmdx_leapfrog(self._h, dt, steps, obs_every, _p(ep), _p(ek)))
        return ep[:k], ek[:k]

    def beeman_init(self, dt: float):
        """Beeman::Initialize: the previous acceleration from the old configuration."""
        _check(lib().lpmdx_beeman_init(self._h, dt))

    def beeman(self, dt: float, steps: int, obs_every: int = 0):
        k = steps // obs_every if obs_every else 0
        ep = np.zeros(max(k, 1))
        ek = np.zeros(max(k, 1))
        _check(lib().lpmdx_beeman(self._h, dt, steps, obs_every, _p(ep), _p(ek)))
        return ep[:k], ek[:k]

    def nosehoover_init(self, dt: float, fmass: float, t: float, kboltzmann: float = 0.00008617):
        """NoseHoover::Initialize (keywords dt, fmass, t of plugins/nosehoover.cc;
        kboltzmann the session's value, "0.00008617" unless changed)."""
        _check(lib().lpmdx_nosehoover_init(self._h, dt, fmass, t, kboltzmann))

    def nosehoover(self, dt: float, steps: int, obs_every: int = 0):
        k = steps // obs_every if obs_every else 0
        ep = np.zeros(max(k, 1))
        ek = np.zeros(max(k, 1))
        _check(lib().lpmdx_nosehoover(self._h, dt, steps, obs_every, _p(ep), _p(ek)))
        return ep[:k], ek[:k]

    def set_temperature(self, temp: float):
        """SetTemperature: scale every velocity to the temperature `temp` (K),
        after drawing them with InitVelocities while none were set."""
        _check(lib().lpmdx_set_temperature(self._h, temp))

    def tempscaling(self, from_: float, to: float, start: int = 0, end: int = -1, each: int = 1):
        """The tempscaling modifier after every step active for (start, end, each); each=0 switches it off."""
        _check(lib().lpmdx_tempscaling(self._h, from_, to, start, end, each))

    def berendsen(self, from_: float, to: float, tau: float = 400.0, start: int = 0, end: int = -1, each: int = 1):
        """The berendsen modifier on the step schedule (start, end, each); each=0 switches it off."""
        _check(lib().lpmdx_berendsen(self._h, from_, to, tau, start, end, each))

    def neighbor_lists(self, full: bool, rcut: float):
        counts = np.zeros(self.n, dtype=np.int64)
        total = C.c_long(0)
        _check(lib().lpmdx_neighbor_lists(self._h, int(full), rcut, _p(counts), 0, None, None, C.byref(total)))
        cap = max(1, total.value)
        js = np.zeros(cap, dtype=np.int64)
        r2 = np.zeros(cap)
        _check(lib().lpmdx_neighbor_lists(self._h, int(full), rcut, _p(counts), cap, _p(js), _p(r2), C.byref(total)))
        return counts, js[: total.value], r2[: total.value]

    def observables(self):
        out = np.zeros(NOBS)
        _check(lib().lpmdx_observables(self._h, _p(out)))
        return out

    def tags(self):
        """The SetSimulationTags tags by name (include/lpmdx.h LPMDX_OBS_*)."""
        out = self.observables()
        return {k: float(out[q]) for k, q in OBS.items()}

    def atom_energies(self, k: int = 0):
        """Potential::AtomEnergy of potential k for every atom."""
        e = np.zeros(self.n)
        _check(lib().lpmdx_atom_energies(self._h, k, _p(e)))
        return e

    def seed48(self, seed: int):
        """Seed the system's drand48 stream as srand48(seed) does."""
        _check(lib().lpmdx_seed48(self._h, seed))

    def init_velocities(self):
        """SimulationEngine::InitVelocities from the system's drand48 stream."""
        _check(lib().lpmdx_init_velocities(self._h))

    def velocities_set(self, flag: bool = True):
        """Mark the velocities as set (SetTemperature then draws none), as Restore does."""
        _check(lib().lpmdx_velocities_set(self._h, int(flag)))

    def set_fixed(self, fixedvel=None, fixedpos=None, fixedpos_true=None):
        """The fixedvel / fixedpos atom tags: boolean masks of the atoms carrying
        each tag; fixedpos_true marks which fixedpos tags read "true" (default:
        all of them).  No argument clears the tags."""
        if fixedvel is None and fixedpos is None:
            _check(lib().lpmdx_system_set_fixed(self._h, None))
            return
        fv = np.zeros(self.n, bool) if fixedvel is None else np.asarray(fixedvel, bool)
        fp = np.zeros(self.n, bool) if fixedpos is None else np.asarray(fixedpos, bool)
        ft = fp if fixedpos_true is None else (fp & np.asarray(fixedpos_true, bool))
        flags = (np.where(fv, FIX_VEL, 0) | np.where(ft, FIX_POS, 0) | np.where(fv | fp, FIX_TAGGED, 0))
        flags = np.ascontiguousarray(flags, dtype=np.int32)
        _check(lib().lpmdx_system_set_fixed(self._h, _p(flags)))

    def snapshot(self, step: int = 0, into=None):
        """Copy x, v, a to the host without waiting for it (lpmdx_snapshot_take):
        a snapshots.Snapshot whose state() waits for the copy; `into` reuses one."""
        from .snapshots import Snapshot

        snap = into if into is not None else Snapshot()
        return snap.take(self, step)

    def stats(self):
        out = np.zeros(9, dtype=np.int64)
        _check(lib().lpmdx_system_stats(self._h, _p(out)))
        keys = ("builds", "force_evals", "steps", "list_cap", "max_list", "lanes", "brick", "list_entries",
                "cut_chunks")
        return dict(zip(keys, (int(v) for v in out)))

    def profile(self, enable: bool = True, every: int = 1):
        """Per-kernel HIP-event timing on (the launches of every `every`-th
        integrator step, and every list build) or off; clears the totals."""
        _check(lib().lpmdx_system_profile(self._h, int(every) if enable else 0))

    def profile_read(self):
        """{slot: (ms, launches)} for force, rho, build, other."""
        ms = np.zeros(4)
        n = np.zeros(4, dtype=np.int64)
        _check(lib().lpmdx_system_profile_read(self._h, _p(ms), _p(n)))
        return {k: (float(ms[q]), int(n[q])) for q, k in enumerate(("force", "rho", "build", "other"))}

    def set_lanes(self, lanes: int):
        _check(lib().lpmdx_system_set_lanes(self._h, lanes))


def build(force: bool = True) -> str:
    """Compile liblpmdx.so for gfx950 in-tree."""
    return _build.build_library(force=force)

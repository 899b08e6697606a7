"""ctypes wrapper of tests/dropin_driver/sim_driver.cc (built into
lpmd2_amd/plugins/_build/libsim_driver.so): a DeviceSimulation made by
DeviceSimulationBuilder, with plugin objects loaded through their create(args).
Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from oracle import refcheck as R

PATH = os.path.join(R.DROPIN_DIR, "libsim_driver.so")
_lib = None


def available() -> bool:
    return R.dropin_available() and os.path.exists(PATH)


def lib():
    global _lib
    if _lib is None:
        R.lib()  # the reference core the plugins resolve against
        L = C.CDLL(PATH)
        vp, d, l, i, s = C.c_void_p, C.c_double, C.c_long, C.c_int, C.c_char_p
        L.sim_last_error.restype = s
        L.sim_new.restype = vp
        L.sim_new.argtypes = [l, vp, vp, vp, vp]
        L.sim_free.argtypes = [vp]
        L.sim_set_cellmanager.argtypes = [vp, s, s]
        L.sim_add_potential.argtypes = [vp, s, s, i, i]
        L.sim_set_integrator.argtypes = [vp, s, s]
        L.sim_do_steps.argtypes = [vp, l]
        L.sim_do_step.argtypes = [vp]
        L.sim_get.argtypes = [vp, vp, vp, vp]
        L.sim_touch.argtypes = [vp]
        L.sim_set_velocity.argtypes = [vp, l, vp]
        L.sim_set_atom_tag.argtypes = [vp, l, s, s]
        L.sim_set_temperature.argtypes = [vp, d, i]
        L.sim_tag.argtypes = [vp, s, vp]
        L.sim_counters.argtypes = [vp, vp]
        L.sim_dump.argtypes = [vp, s]
        L.sim_restore.argtypes = [vp, s]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().sim_last_error().decode())


# SetSimulationTags' tag names in the lpmdx_observables layout (the virial, [6], is no tag)
TAGS = {"kinetic-energy": 0, "temperature": 1, "px": 2, "py": 3, "pz": 4, "potential-energy": 5, "pressure": 7,
        "total-energy": 17, "momentum": 18, "volume": 19, "volume-per-atom": 20, "particle-density": 21,
        "density": 22, "virial-pressure": 23, "kinetic-pressure": 24, "cell-a": 25, "cell-b": 26, "cell-c": 27}
for _a, _ca in enumerate("xyz"):
    for _b, _cb in enumerate("xyz"):
        TAGS["s" + _ca + _cb] = 8 + 3 * _a + _b


class Sim:
    """DeviceSimulationBuilder::CreateFixedOrthogonal with the given atoms."""

    def __init__(self, x, L, z, v=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self.n = x.shape[0]
        Lb = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.float64), (3,)))
        zz = np.ascontiguousarray(np.broadcast_to(np.asarray(z, dtype=np.int32), (self.n,)))
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(self.n, 3)
        self.h = lib().sim_new(self.n, _p(Lb), _p(zz), _p(x), None if vv is None else _p(vv))
        if not self.h:
            raise RuntimeError(lib().sim_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().sim_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_cellmanager(self, path: str, args: str):
        _check(lib().sim_set_cellmanager(self.h, path.encode(), (args + " debug none").encode()))

    def add_potential(self, path: str, args: str, s1: int = -1, s2: int = -1):
        _check(lib().sim_add_potential(self.h, path.encode(), (args + " debug none").encode(), s1, s2))

    def set_integrator(self, path: str, args: str):
        _check(lib().sim_set_integrator(self.h, path.encode(), (args + " debug none").encode()))

    def do_steps(self, n: int):
        _check(lib().sim_do_steps(self.h, int(n)))

    def do_step(self):
        _check(lib().sim_do_step(self.h))

    def get(self):
        x, v, a = (np.zeros((self.n, 3)) for _ in range(3))
        _check(lib().sim_get(self.h, _p(x), _p(v), _p(a)))
        return x, v, a

    def touch(self):
        """Atoms() once (a loop that reads or tags the atoms each step)."""
        _check(lib().sim_touch(self.h))

    def set_velocity(self, i: int, v):
        vv = np.ascontiguousarray(v, dtype=np.float64).reshape(3)
        _check(lib().sim_set_velocity(self.h, int(i), _p(vv)))

    def set_atom_tag(self, i: int, key: str, value: str = "true"):
        _check(lib().sim_set_atom_tag(self.h, int(i), key.encode(), value.encode()))

    def set_temperature(self, t: float, tag: bool = False):
        _check(lib().sim_set_temperature(self.h, float(t), int(tag)))

    def tag(self, name: str) -> float:
        out = C.c_double()
        _check(lib().sim_tag(self.h, name.encode(), C.byref(out)))
        return out.value

    def counters(self):
        out = np.zeros(5, dtype=np.int64)
        _check(lib().sim_counters(self.h, _p(out)))
        return dict(zip(("refreshes", "device_runs", "uploads", "host_steps", "step"), (int(q) for q in out)))

    def dump(self, path: str):
        _check(lib().sim_dump(self.h, path.encode()))

    def restore(self, path: str):
        _check(lib().sim_restore(self.h, path.encode()))

"""Snapshots and configuration files over the C ABI (include/lpmdx.h).

`System.snapshot(step)` copies a system's state to the host without stopping
its step loop; the writers and readers below are the reference's
configuration formats: `SimulationEngine::Dump` / `Restore`
(liblpmd/simulation/simulationbuilder.cc:195-293), the lpmd / zlp format of
plugins/lpmd.cc and the xyz writer of plugins/xyz.cc.  Cells are (3, 3)
arrays of cell vectors (rows); positions are Cartesian, in the caller's atom
order.  The file routines run on the host; only `take` needs the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import LpmdxError, _check, _p, lib as _engine_lib

_ready = False


def lib():
    global _ready
    L = _engine_lib()
    if not _ready:
        vp, d, l, i, s = C.c_void_p, C.c_double, C.c_long, C.c_int, C.c_char_p
        L.lpmdx_snapshot_create.argtypes = [vp]
        L.lpmdx_snapshot_destroy.argtypes = [vp]
        L.lpmdx_snapshot_take.argtypes = [vp, vp, l]
        L.lpmdx_snapshot_wait.argtypes = [vp]
        L.lpmdx_snapshot_info.argtypes = [vp, vp, vp, vp]
        L.lpmdx_snapshot_state.argtypes = [vp, vp, vp, vp, vp, vp]
        L.lpmdx_dump_write.argtypes = [s, l, vp, l, vp, vp, vp, vp, vp]
        L.lpmdx_dump_read.argtypes = [s, vp, vp, vp, l, vp, vp, vp, vp, vp]
        L.lpmdx_lpmd_open_write.argtypes = [vp, s, s]
        L.lpmdx_lpmd_write.argtypes = [vp, l, vp, vp, vp, vp, vp]
        L.lpmdx_lpmd_open_read.argtypes = [vp, s, vp]
        L.lpmdx_lpmd_next.argtypes = [vp, vp]
        L.lpmdx_lpmd_frame.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.lpmdx_lpmd_close.argtypes = [vp]
        L.lpmdx_xyz_write.argtypes = [s, i, l, vp, vp, vp, vp, vp, s]
        _ready = True
    return L


def _cell(cell):
    c = np.asarray(cell, dtype=np.float64)
    if c.shape == (3,):
        c = np.diag(c)
    return np.ascontiguousarray(c.reshape(3, 3))


def _vec(a, n):
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n, 3))


def _pp(a):
    return None if a is None else _p(a)


class Snapshot:
    """A state copy of a System, taken asynchronously (lpmdx_snapshot_*)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().lpmdx_snapshot_create(C.byref(h)))
        self._h = h

    def take(self, system, step: int = 0):
        """Enqueue the copy of the system's current x, v, a; returns at once."""
        _check(lib().lpmdx_snapshot_take(system._h, self._h, int(step)))
        return self

    def wait(self):
        _check(lib().lpmdx_snapshot_wait(self._h))
        return self

    def state(self):
        """dict(step, cell, z, x, v, a, flags), waiting for the copy first."""
        n, step = C.c_long(), C.c_long()
        cell = np.zeros((3, 3))
        _check(lib().lpmdx_snapshot_info(self._h, C.byref(n), C.byref(step), _p(cell)))
        m = n.value
        z = np.zeros(m, dtype=np.int32)
        x, v, a = (np.zeros((m, 3)) for _ in range(3))
        flags = np.zeros(m, dtype=np.int32)
        _check(lib().lpmdx_snapshot_state(self._h, _p(z), _p(x), _p(v), _p(a), _p(flags)))
        return dict(step=step.value, cell=cell, z=z, x=x, v=v, a=a, flags=flags)

    def close(self):
        if getattr(self, "_h", None):
            lib().lpmdx_snapshot_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- Dump / Restore
def write_dump(path, step, cell, z, x, v, a, flags=None):
    """SimulationEngine::Dump of a state (numbers at six significant digits,
    as the reference writes them)."""
    z = np.ascontiguousarray(z, dtype=np.int32)
    n = z.shape[0]
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.int32)
    _check(lib().lpmdx_dump_write(str(path).encode(), int(step), _p(_cell(cell)), n, _p(z), _p(_vec(x, n)),
                                  _p(_vec(v, n)), _p(_vec(a, n)), _pp(fl)))


def read_dump(path):
    """SimulationEngine::Restore: dict(step, cell, z, x, v, a, flags)."""
    n, step = C.c_long(), C.c_long()
    cell = np.zeros((3, 3))
    _check(lib().lpmdx_dump_read(str(path).encode(), C.byref(step), _p(cell), C.byref(n), 0, None, None, None,
                                 None, None))
    m = n.value
    z = np.zeros(m, dtype=np.int32)
    x, v, a = (np.zeros((m, 3)) for _ in range(3))
    flags = np.zeros(m, dtype=np.int32)
    _check(lib().lpmdx_dump_read(str(path).encode(), C.byref(step), _p(cell), C.byref(n), m, _p(z), _p(x), _p(v),
                                 _p(a), _p(flags)))
    return dict(step=step.value, cell=cell, z=z, x=x, v=v, a=a, flags=flags)


# ---------------------------------------------------------------- lpmd / zlp
class LpmdWriter:
    """The lpmd plugin as an output module: header at open, one frame per write."""

    def __init__(self, path, args: str = ""):
        h = C.c_void_p()
        _check(lib().lpmdx_lpmd_open_write(C.byref(h), str(path).encode(), args.encode()))
        self._h = h

    def write(self, cell, z, x, v=None, a=None):
        z = np.ascontiguousarray(z, dtype=np.int32)
        n = z.shape[0]
        _check(lib().lpmdx_lpmd_write(self._h, n, _p(_cell(cell)), _p(z), _p(_vec(x, n)), _pp(_vec(v, n)),
                                      _pp(_vec(a, n))))

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().lpmdx_lpmd_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_lpmd(path):
    """Every frame of an lpmd / zlp file: a list of dict(cell, z, s, v, a,
    level), s the scaled positions as written (cartesian(s, cell) gives the
    positions in an orthogonal cell)."""
    h = C.c_void_p()
    level = C.c_int()
    _check(lib().lpmdx_lpmd_open_read(C.byref(h), str(path).encode(), C.byref(level)))
    frames = []
    try:
        while True:
            n = C.c_long()
            _check(lib().lpmdx_lpmd_next(h, C.byref(n)))
            if n.value < 0:
                break
            m = n.value
            cell = np.zeros((3, 3))
            z = np.zeros(m, dtype=np.int32)
            s, v, a = (np.zeros((m, 3)) for _ in range(3))
            lv = C.c_int()
            _check(lib().lpmdx_lpmd_frame(h, _p(cell), _p(z), _p(s), _p(v), _p(a), C.byref(lv)))
            frames.append(dict(cell=cell, z=z, s=s, v=v, a=a, level=lv.value))
    finally:
        lib().lpmdx_lpmd_close(h)
    return frames


def cartesian(s, cell):
    """OrthogonalCell::Cartesian (orthogonalcell.h:96-100): s_q |cell row q|."""
    c = _cell(cell)
    return np.asarray(s) * np.sqrt((c * c).sum(axis=1))


# ---------------------------------------------------------------- xyz
def write_xyz(path, cell, z, x, v=None, a=None, args: str = "", append: bool = False):
    """One xyz frame (XYZFormat::WriteCell); args: level, precout, inside."""
    z = np.ascontiguousarray(z, dtype=np.int32)
    n = z.shape[0]
    _check(lib().lpmdx_xyz_write(str(path).encode(), int(bool(append)), n, _p(_cell(cell)), _p(z),
                                 _p(_vec(x, n)), _pp(_vec(v, n)), _pp(_vec(a, n)), args.encode()))


__all__ = ["Snapshot", "write_dump", "read_dump", "LpmdWriter", "read_lpmd", "cartesian", "write_xyz",
           "LpmdxError"]

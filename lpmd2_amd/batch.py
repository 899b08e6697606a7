"""Batched engine: many independent periodic systems advanced together
(lpmdx_batch_* in include/lpmdx.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _p, lib

KINDS = {"lennardjones": 1, "fastlj": 2, "morse": 3, "buckingham": 4, "suttonchen": 5, "finnissinclair": 6,
         "gupta": 7, "harmonic": 8, "lennardjonesMod": 9, "finnissinclair-ext": 10,
         "tabulatedpair": 2}  # tabulatedpair runs the fastlj table kernels
NKINDS = 10  # LPMDX_NKINDS

_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, d, l, i, s = C.c_void_p, C.c_double, C.c_long, C.c_int, C.c_char_p
    L.lpmdx_batch_create.argtypes = [vp, l, l, l]
    L.lpmdx_batch_destroy.argtypes = [vp]
    L.lpmdx_batch_add.argtypes = [vp, l, vp, vp, vp, vp, vp, s, i, vp, vp, vp, vp, d, l, vp]
    L.lpmdx_batch_remove.argtypes = [vp, l]
    L.lpmdx_batch_run.argtypes = [vp, l]
    L.lpmdx_batch_get.argtypes = [vp, l, vp, vp, vp, vp]
    L.lpmdx_batch_energies.argtypes = [vp, l, vp]
    L.lpmdx_batch_autotune.argtypes = [vp, l, vp, vp, vp]
    L.lpmdx_batch_set_tiles.argtypes = [vp, i, i]
    L.lpmdx_batch_set_lanes.argtypes = [vp, i, i]
    L.lpmdx_batch_set_cells.argtypes = [vp, i, i]
    L.lpmdx_batch_info.argtypes = [vp, vp]
    L.lpmdx_batch_bytes_per_atom.restype = l
    L.lpmdx_batch_progress.argtypes = [vp, l, vp, vp]
    _bound = True
    return L


class Batch:
    """Shared HBM arenas for atoms, neighbour lists and cells; systems are added
    with their own box, species, potentials, dt and step count."""

    def __init__(self, atom_capacity: int, list_words: int | None = None, cell_slots: int | None = None):
        L = _bind()
        list_words = list_words or atom_capacity * 256
        cell_slots = cell_slots or max(4096, atom_capacity)
        h = C.c_void_p()
        _check(L.lpmdx_batch_create(C.byref(h), atom_capacity, list_words, cell_slots))
        self._h = h
        self.sizes = {}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().lpmdx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, x, box, z, potentials, cm="", v=None, mass=None, dt=1.0, steps=0) -> int:
        """potentials: list of (name, args) or (name, args, s1, s2)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        box = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
        zz = np.ascontiguousarray(np.broadcast_to(np.asarray(z, dtype=np.int32), (n,)))
        mm = None if mass is None else np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        k = len(potentials)
        names = (C.c_char_p * k)(*[p[0].encode() for p in potentials])
        args = (C.c_char_p * k)(*[p[1].encode() for p in potentials])
        s1 = np.array([p[2] if len(p) > 2 else -1 for p in potentials], dtype=np.int32)
        s2 = np.array([p[3] if len(p) > 3 else -1 for p in potentials], dtype=np.int32)
        sid = C.c_long()
        _check(lib().lpmdx_batch_add(self._h, n, _p(box), _p(zz), _p(mm), _p(x), _p(vv), cm.encode(), k,
                                     C.cast(names, C.c_void_p), C.cast(args, C.c_void_p), _p(s1), _p(s2),
                                     float(dt), int(steps), C.byref(sid)))
        self.sizes[sid.value] = n
        return sid.value

    def remove(self, sid: int):
        _check(lib().lpmdx_batch_remove(self._h, sid))
        self.sizes.pop(sid, None)

    def run(self, steps: int):
        _check(lib().lpmdx_batch_run(self._h, steps))

    def get(self, sid: int):
        n = self.sizes[sid]
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        a = np.zeros((n, 3))
        info = np.zeros(8, dtype=np.int64)
        _check(lib().lpmdx_batch_get(self._h, sid, _p(x), _p(v), _p(a), _p(info)))
        keys = ("done", "left", "builds", "started", "max_list", "cap", "lanes", "overflow")
        return x, v, a, dict(zip(keys, (int(q) for q in info)))

    def energies(self, sid: int):
        out = np.zeros(3)
        _check(lib().lpmdx_batch_energies(self._h, sid, _p(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def autotune(self, trial_steps: int = 10, cells: bool = False):
        """Lanes per atom, cells per list radius and tiles per workgroup of each
        potential kind, timed on a snapshot of the batch: {kind: lanes}, or
        with cells=True {kind: (lanes, cells per list radius, tiles per workgroup)}."""
        lanes = np.zeros(NKINDS + 1, dtype=np.int32)
        div = np.zeros(NKINDS + 1, dtype=np.int32)
        tpw = np.zeros(NKINDS + 1, dtype=np.int32)
        _check(lib().lpmdx_batch_autotune(self._h, trial_steps, _p(lanes), _p(div), _p(tpw)))
        if cells:
            return {name: (int(lanes[k]), int(div[k]), int(tpw[k])) for name, k in KINDS.items()}
        return {name: int(lanes[k]) for name, k in KINDS.items()}

    def set_tiles(self, kind: str, tiles_per_group: int):
        """Tiles of 64 atoms per workgroup (1, 2 or 4) of one kind's force and rho launches."""
        _check(lib().lpmdx_batch_set_tiles(self._h, KINDS[kind], tiles_per_group))

    def set_lanes(self, kind: str, lanes: int):
        _check(lib().lpmdx_batch_set_lanes(self._h, KINDS[kind], lanes))

    def set_cells(self, kind: str, per_radius: int):
        """Cells per list radius (1 or 2) of the grids of one kind's systems."""
        _check(lib().lpmdx_batch_set_cells(self._h, KINDS[kind], per_radius))

    def progress(self, nslots: int):
        """(left, flags) per slot: graph steps still to run; bit 0 live, bit 1
        overflow, bit 2 initial force evaluation pending."""
        left = np.zeros(nslots, dtype=np.int64)
        flags = np.zeros(nslots, dtype=np.int32)
        _check(lib().lpmdx_batch_progress(self._h, nslots, _p(left), _p(flags)))
        return left, flags

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        _check(lib().lpmdx_batch_info(self._h, _p(out)))
        keys = ("systems", "atoms", "atom_capacity", "free_atoms", "free_list_words", "free_cells", "steps_run",
                "slots")
        return dict(zip(keys, (int(q) for q in out)))


def bytes_per_atom() -> int:
    return int(_bind().lpmdx_batch_bytes_per_atom())

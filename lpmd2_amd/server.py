"""Serving loop: MD jobs submitted at any time run together on the device with
continuous batching (lpmdx_server_* in include/lpmdx.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _p, lib
from .batch import KINDS

_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    vp, d, l, i, s = C.c_void_p, C.c_double, C.c_long, C.c_int, C.c_char_p
    L.lpmdx_server_create.argtypes = [vp, l, l]
    L.lpmdx_server_destroy.argtypes = [vp]
    L.lpmdx_server_submit.argtypes = [vp, l, vp, vp, vp, vp, vp, s, i, vp, vp, vp, vp, d, l, vp]
    L.lpmdx_server_wait.argtypes = [vp, l, d, vp]
    L.lpmdx_server_result.argtypes = [vp, l, vp, vp, vp, vp]
    L.lpmdx_server_release.argtypes = [vp, l]
    L.lpmdx_server_stats.argtypes = [vp, vp]
    L.lpmdx_server_set_lanes.argtypes = [vp, i, i]
    _bound = True
    return L


STAT_KEYS = ("done", "failed", "p50_s", "p99_s", "mean_s", "atom_steps", "span_s", "rounds", "max_resident",
             "queued", "running", "graph_steps", "admit_s", "run_s", "retire_s", "reserved")


class Server:
    """A worker thread that admits jobs into HBM arenas sized from `hbm_budget`
    bytes and advances all admitted jobs together."""

    def __init__(self, hbm_budget: int, max_round_steps: int = 256):
        L = _bind()
        h = C.c_void_p()
        _check(L.lpmdx_server_create(C.byref(h), int(hbm_budget), int(max_round_steps)))
        self._h = h
        self.sizes = {}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().lpmdx_server_destroy(self._h)
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

    def submit(self, x, box, z, potentials, cm="", v=None, mass=None, dt=1.0, steps=0) -> int:
        """potentials: list of (name, args) or (name, args, s1, s2); returns the job id."""
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
        job = C.c_long()
        _check(lib().lpmdx_server_submit(self._h, n, _p(box), _p(zz), _p(mm), _p(x), _p(vv), cm.encode(), k,
                                         C.cast(names, C.c_void_p), C.cast(args, C.c_void_p), _p(s1), _p(s2),
                                         float(dt), int(steps), C.byref(job)))
        self.sizes[job.value] = n
        return job.value

    def submit_job(self, job: dict) -> int:
        """Submits a job dict as made by lpmd2_amd.workloads."""
        return self.submit(job["x"], job["box"], job["z"], job["potentials"], cm=job["cm"], v=job.get("v"),
                           dt=job["dt"], steps=job["steps"])

    def wait(self, job: int = -1, timeout: float = -1.0) -> bool:
        done = C.c_int(0)
        _check(lib().lpmdx_server_wait(self._h, job, float(timeout), C.byref(done)))
        return bool(done.value)

    def result(self, job: int):
        """(x, v, a, info) of a finished job; raises if the job failed."""
        n = self.sizes[job]
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        a = np.zeros((n, 3))
        out = np.zeros(8)
        _check(lib().lpmdx_server_result(self._h, job, _p(x), _p(v), _p(a), _p(out)))
        info = dict(zip(("epot", "virial", "ekin", "queued_s", "latency_s", "builds", "steps", "status"),
                        (float(q) for q in out)))
        if int(info["status"]) != 0:
            from . import LpmdxError
            raise LpmdxError(int(info["status"]), lib().lpmdx_last_error().decode())
        return x, v, a, info

    def release(self, job: int):
        _check(lib().lpmdx_server_release(self._h, job))
        self.sizes.pop(job, None)

    def stats(self):
        out = np.zeros(len(STAT_KEYS))
        _check(lib().lpmdx_server_stats(self._h, _p(out)))
        return {k: float(q) for k, q in zip(STAT_KEYS, out)}

    def set_lanes(self, kind: str, lanes: int):
        _check(lib().lpmdx_server_set_lanes(self._h, KINDS[kind], lanes))

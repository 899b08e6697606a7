"""Benchmark of the MI355X LPMD engine.

Default workload (one GPU): Sutton-Chen Cu bulk, fcc 32x32x32 cells = 131072
atoms, a = 3.61 A, n 9 m 6 e 1.2382e-2 eV c 39.432, rc = 2a = 7.22 A, seeded
300 K velocities plus 0.01 A Gaussian position noise, Verlet list with a
0.3 A skin, velocity Verlet dt = 2 fs.  A step is one velocity-Verlet step of
the whole system (both EAM passes, list rebuilds when the skin is exceeded),
with the state resident in HBM.

Prints one JSON line (value = atom-steps/s) with the dominant kernel's
roofline (HIP events on the engine's stream) and a CPU baseline (the FP64
oracle, a restatement of the reference path, on one host core, timed on a
bounded sample and extrapolated to one reference step).

Every plain-path workload reports, after its timed steps, its relative force
RMS against the FP64 oracle (all atoms of a small box, a seeded sample of a
large one) and its total-energy drift over the timed steps ("parity"); "legs"
holds the same record for the 4,000,000-atom LJ liquid (the HBM-roofline
case) and the reference's 2048-atom argon case.

The same line carries the two other legs of the metric:
  "batched": 512 periodic systems, sizes log-uniform in [256, 16384] atoms,
      half LJ argon, half Sutton-Chen Cu/Ni/Ag, 1000 steps each in one batch
      (graph-replayed steps, autotuned lanes), beside the plain path running
      the same systems one at a time (timed over --plain-steps steps each);
  "served": Poisson arrivals of 500-65536-atom Gupta / Finnis-Sinclair /
      Buckingham rocksalt jobs of 100-5000 steps at a few rates, p50/p99
      completion latency and atom-steps/s, beside the same jobs run one at a
      time on the plain path (service times measured, FIFO queue replayed on
      the same arrival times).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32
FP64_PEAK_TFLOPS = 78.6   # MI355X vector FP64

# Algorithmic flops of the force pass: per listed pair the FP64 displacement and
# r^2 (3 sub, 3 mul, 2 add); per pair inside the cutoff the FP64 force sum
# (3 FMA = 6 flops) and the FP32 pair function (LJ: rcp, 3 mul for s6, 1 for
# s12, 4 for the force scalar = 9; Sutton-Chen force pass: rsq, 1 mul for t,
# 5 for t^9 and t^6 at the common exponents, 6 for the force scalar = 13).
FP64_PER_LISTED_PAIR = 8
FP64_PER_CUT_PAIR = 6
FP32_PER_CUT_PAIR = {"lennardjones": 9, "suttonchen": 13,
                     # harmonic: r from rsq, (r - a), k (r - a) / r; lennardjonesMod: t = sigma^2 / r^2,
                     # t^5 by squaring, t^2.5 through powf (counted as 4), the force scalar;
                     # finnissinclair-ext: r, the rho-derivative term (the pair polynomial is FP64)
                     "harmonic": 4, "lennardjonesMod": 14, "finnissinclair-ext": 9}

PROF_EVERY = 4  # step sampling period of the per-kernel HIP-event timing

WORKLOADS = {
    # the reference's own CPU-runnable case: 2048 argon atoms on the linkedcell
    # stencil of `mode auto` (12^3 cells, two a side), 1000 steps of 1 fs
    "lj_ar_2048": dict(
        desc="LJ argon fcc 8^3 = 2048 atoms, a 5.26, rc 8.5, linkedcell mode auto, dt 1 fs, 100 K",
        nx=8, a=5.26, z=18, T=100.0, noise=0.0, dt=1.0, seed=2048,
        pot=("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5"),
        cm=("linkedcell", "cutoff 8.5 mode auto"),
        rc=8.5, skin=0.0,
        metal=False, steps=1000,
    ),
    "sc_cu_131k": dict(
        desc="Sutton-Chen Cu fcc 32^3 = 131072 atoms, rc 7.22, skin 0.3, dt 2 fs, 300 K",
        nx=32, a=3.61, z=29, T=300.0, noise=0.01, dt=2.0, seed=2024,
        pot=("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22"),
        cm=("verletlist", "cutoff 7.22 maxneighbors 400 each 5 skin 0.3"),
        rc=7.22, skin=0.3,
        metal=True,
    ),
    "lj_4m": dict(
        desc="LJ liquid fcc 100^3 = 4,000,000 atoms, rho* 0.8442, T* 0.72, rc 2.5 sigma, skin 0.3 sigma, dt* 0.005",
        nx=100, a=1.679596 * 3.41, z=18, T=0.72 * 0.0103 / 8.6173422e-05, noise=0.0, dt=0.005 * 2161.98, seed=7,
        pot=("lennardjones", f"sigma 3.41 epsilon 0.0103 cutoff {2.5 * 3.41}"),
        cm=("verletlist", f"cutoff {2.5 * 3.41} maxneighbors 400 each 5 skin {0.3 * 3.41}"),
        rc=2.5 * 3.41, skin=0.3 * 3.41,
        metal=False,
    ),
    # the same LJ liquid under the leapfrog integrator (one force call per step)
    "lj_4m_leapfrog": dict(
        desc="LJ liquid fcc 100^3 = 4,000,000 atoms, rho* 0.8442, T* 0.72, rc 2.5 sigma, skin 0.3 sigma, dt* 0.005, "
             "leapfrog",
        nx=100, a=1.679596 * 3.41, z=18, T=0.72 * 0.0103 / 8.6173422e-05, noise=0.0, dt=0.005 * 2161.98, seed=7,
        pot=("lennardjones", f"sigma 3.41 epsilon 0.0103 cutoff {2.5 * 3.41}"),
        cm=("verletlist", f"cutoff {2.5 * 3.41} maxneighbors 400 each 5 skin {0.3 * 3.41}"),
        rc=2.5 * 3.41, skin=0.3 * 3.41,
        metal=False, integrator="leapfrog",
    ),
    "lj_4m_beeman": dict(
        desc="LJ liquid fcc 100^3 = 4,000,000 atoms, rho* 0.8442, T* 0.72, rc 2.5 sigma, skin 0.3 sigma, dt* 0.005, "
             "beeman",
        nx=100, a=1.679596 * 3.41, z=18, T=0.72 * 0.0103 / 8.6173422e-05, noise=0.0, dt=0.005 * 2161.98, seed=7,
        pot=("lennardjones", f"sigma 3.41 epsilon 0.0103 cutoff {2.5 * 3.41}"),
        cm=("verletlist", f"cutoff {2.5 * 3.41} maxneighbors 400 each 5 skin {0.3 * 3.41}"),
        rc=2.5 * 3.41, skin=0.3 * 3.41,
        metal=False, integrator="beeman",
    ),
    # the widened potentials, one 131072-atom box each
    "harmonic_131k": dict(
        desc="harmonic springs k 3.4 a 3.0 on fcc 32^3 = 131072 atoms, a 4.243 (nearest neighbours at the spring "
             "length), list cutoff 3.5, skin 0.3, dt 1 fs, 300 K",
        nx=32, a=4.243, z=18, T=300.0, noise=0.01, dt=1.0, seed=11,
        pot=("harmonic", "k 3.4 a 3.0"),
        cm=("verletlist", "cutoff 3.5 maxneighbors 100 each 5 skin 0.3"),
        rc=3.5, skin=0.3,
        metal=False,
    ),
    "ljmod_ar_131k": dict(
        desc="lennardjonesMod Ar m 10 n 5 cohesive 0.8 on fcc 32^3 = 131072 atoms, a 5.26, rc 8.5, skin 1.0, "
             "dt 2 fs, 100 K",
        nx=32, a=5.26, z=18, T=100.0, noise=0.01, dt=2.0, seed=12,
        pot=("lennardjonesMod", "sigma 3.41 epsilon 0.0103 m 10 n 5 cohesive 0.8 cutoff 8.5"),
        cm=("verletlist", "cutoff 8.5 maxneighbors 400 each 5 skin 1.0"),
        rc=8.5, skin=1.0,
        metal=False,
    ),
    "fsext_mo_131k": dict(
        desc="finnissinclair-ext Mo (the plugin's example: c 3.2572, d 4.1472, B 0) on fcc 32^3 = 131072 atoms, "
             "a 3.625 (zero virial of the static lattice), skin 0.3, dt 2 fs, 300 K",
        nx=32, a=3.625, z=42, T=300.0, noise=0.01, dt=2.0, seed=13,
        pot=("finnissinclair-ext", "c0 47.98066 c1 -34.09924 c2 5.832293 c3 0.017494 c4 0.020393 A 1.848648 "
                                   "B 0.0 c 3.2572 d 4.1472"),
        cm=("verletlist", "cutoff 4.1472 maxneighbors 200 each 5 skin 0.3"),
        rc=4.1472, skin=0.3,
        metal=True,
    ),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(w):
    import lpmd2_amd as E

    x, L = E.fcc(w["nx"], w["a"])
    n = x.shape[0]
    rng = np.random.default_rng(w["seed"])
    if w["noise"] > 0:
        x = np.mod(x + rng.normal(0.0, w["noise"], x.shape), L)
    m = np.full(n, E.ELEMENT_MASS[w["z"]])
    v = E.maxwell_boltzmann(m, w["T"], seed=w["seed"] + 1)
    s = E.System(x, L, w["z"], v=v)
    s.set_cellmanager(*w["cm"])
    s.append_potential(*w["pot"])
    return s, x, L, v


def cpu_baseline(w, budget_s: float):
    """The oracle (FP64 restatement of the reference path) on one core, on a
    bounded sample of the same workload: the fcc box with (nx/2)^3 cells (1/8 of
    the atoms, same density and parameters).  Timed there: one O(N^2)
    verletlist rebuild and force calls; extrapolated to the full system (rebuild
    x N^2, force call x N) and to the reference's step: 2 force calls and a
    rebuild every `each` = 5 UpdateCell calls, i.e. 0.4 rebuilds per step."""
    from oracle import oracle as O
    import lpmd2_amd as E

    nxs = max(2, w["nx"] // 2)
    x, L = E.fcc(nxs, w["a"])
    rng = np.random.default_rng(w["seed"])
    if w["noise"] > 0:
        x = np.mod(x + rng.normal(0.0, w["noise"], x.shape), L)
    ns = x.shape[0]
    n = 4 * w["nx"] ** 3
    conf = O.Conf(x, L, z=w["z"])
    kw = dict(zip(w["pot"][1].split()[0::2], (float(v) for v in w["pot"][1].split()[1::2])))
    pot = O.Pot(w["pot"][0], **kw)
    # a potential without a cutoff (harmonic) takes the workload's list cutoff
    cut = pot.cutoff if np.isfinite(pot.cutoff) else w["rc"]
    cm = O.CellManager("verletlist", cut, maxneighbors=400, each=5)
    t0 = time.perf_counter()
    cm.update(conf)  # the O(N^2) rebuild
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    ncalls = 0
    while True:
        O.update_forces(conf, cm, [pot])  # UpdateCell calls 2..: no rebuild under each 5
        ncalls += 1
        if time.perf_counter() - t0 > min(10.0, budget_s) or ncalls >= 3:
            break
    t_force = (time.perf_counter() - t0) / ncalls
    scale = n / ns
    # velocityverlet (TwoStepIntegrator) calls UpdateForces twice a step,
    # leapfrog (OneStepIntegrator) once; the list is rebuilt every 5 calls
    calls = 1 if w.get("integrator") == "leapfrog" else 2
    t_step = calls * t_force * scale + calls / 5 * t_build * scale * scale
    return {
        "value": n / t_step,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle, verletlist each 5, on the {ns}-atom box of the same workload: 1 O(N^2) rebuild "
                  f"({t_build:.2f} s, scaled x{scale * scale:.0f}) + {ncalls} force calls ({t_force:.3f} s each, "
                  f"scaled x{scale:.0f}); step = {calls} force calls + {calls / 5:.1f} rebuild",
        "cpu": cpu_model(),
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_probe(nbytes: int):
    """One device-to-device hipMemcpy of nbytes (reads and writes nbytes once)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy probe: {nbytes} bytes device to device")


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary profiles/<kernel>_<workload>_traffic.json (FETCH_SIZE +
    WRITE_SIZE, each scaled by the same counter on a copy of known size), or
    None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def plain_run(job, steps=None, warmup=0):
    """One job on the plain single-system path: (seconds to create and upload,
    seconds of `steps` velocity-Verlet steps after `warmup` steps)."""
    import lpmd2_amd as E

    t0 = time.perf_counter()
    s = E.System(job["x"], job["box"], job["z"], v=job["v"])
    s.set_cellmanager("verletlist", job["cm"] + " maxneighbors 4096")
    for p in job["potentials"]:
        s.append_potential(*p)
    t_create = time.perf_counter() - t0
    if warmup:
        s.velocityverlet(job["dt"], warmup)
    t0 = time.perf_counter()
    s.velocityverlet(job["dt"], job["steps"] if steps is None else steps)
    t_run = time.perf_counter() - t0
    return t_create, t_run


def batched_leg(args):
    """512 systems in one batch against the same systems one at a time."""
    from lpmd2_amd import workloads as W
    from lpmd2_amd.batch import Batch

    steps = args.batch_steps
    warm = 10
    ens = W.batched_ensemble(seed=3, n_systems=args.batch_systems, steps=steps + warm)
    atoms = sum(j["x"].shape[0] for j in ens)
    padded = sum(-(-j["x"].shape[0] // 64) * 64 for j in ens)
    log(f"batched: {len(ens)} systems, {atoms} atoms")
    b = Batch(atom_capacity=padded + 64, list_words=padded * 320, cell_slots=max(4096, padded))
    t0 = time.perf_counter()
    sids = [b.add(x=j["x"], box=j["box"], z=j["z"], potentials=j["potentials"], cm=j["cm"], v=j["v"], dt=j["dt"],
                  steps=j["steps"]) for j in ens]
    log(f"  added in {time.perf_counter() - t0:.2f} s")
    t0 = time.perf_counter()
    shape = b.autotune(trial_steps=5, cells=True)
    lanes = {k: v[0] for k, v in shape.items()}
    cells = {k: v[1] for k, v in shape.items()}
    tiles = {k: v[2] for k, v in shape.items()}
    log(f"  autotune {time.perf_counter() - t0:.2f} s: lanes {lanes}, cells per list radius {cells}, "
        f"tiles per workgroup {tiles}")
    b.run(warm + 1)  # initial force evaluation + warmup steps
    t0 = time.perf_counter()
    b.run(steps)
    t_b = time.perf_counter() - t0
    info = b.info()
    done = [b.get(sid)[3] for sid in sids[:4]]
    assert all(d["left"] == 0 and d["done"] == steps + warm for d in done), done
    batched = atoms * steps / t_b
    log(f"  batched {steps} steps: {t_b:.3f} s, {batched:.4e} atom-steps/s, {info}")
    b.close()
    out = {
        "value": batched,
        "unit": "atom-steps/s",
        "systems": len(ens),
        "atoms": atoms,
        "steps": steps,
        "ms_per_step": 1e3 * t_b / steps,
        "lanes": {k: v for k, v in lanes.items() if k in ("lennardjones", "suttonchen")},
        "cells_per_list_radius": {k: v for k, v in cells.items() if k in ("lennardjones", "suttonchen")},
        "tiles_per_workgroup": {k: v for k, v in tiles.items() if k in ("lennardjones", "suttonchen")},
        "workload": "512 systems, N log-uniform [256, 16384], half LJ Ar rho* 0.8-1.0 (skin 1.0), "
                    "half Sutton-Chen Cu/Ni/Ag rc 2a (skin 0.3), 100-900 K, seed 3",
    }
    if args.no_plain:
        return out
    # the plain path, one system at a time
    t_plain = 0.0
    t0 = time.perf_counter()
    for j in ens:
        _, t_run = plain_run(j, steps=args.plain_steps, warmup=2)
        t_plain += t_run
    plain = atoms * args.plain_steps / t_plain
    log(f"  plain one at a time ({args.plain_steps} steps each): {t_plain:.3f} s of steps "
        f"({time.perf_counter() - t0:.1f} s wall), {plain:.4e} atom-steps/s")
    out["plain_value"] = plain
    out["plain_steps_per_system"] = args.plain_steps
    out["speedup_vs_plain"] = batched / plain
    return out


def served_leg(args):
    """Poisson arrivals into the server at a few rates, and the same jobs run
    one at a time on the plain path."""
    from lpmd2_amd import workloads as W
    from lpmd2_amd.server import Server

    out = {"rates": [], "workload": "Poisson arrivals, N log-uniform [500, 65536], Gupta/Finnis-Sinclair Cu/Al/Ni "
                                    "fcc or Buckingham rocksalt NaCl, 300-1200 K, 100-5000 steps, seed 5"}
    plain_service = None
    for rate in args.serve_rates:
        jobs = W.served_jobs(seed=5, n_jobs=args.serve_jobs, rate_hz=rate)
        work = sum(j["x"].shape[0] * j["steps"] for _, j in jobs)
        with Server(args.serve_budget << 30, max_round_steps=256) as srv:
            t0 = time.perf_counter()
            ids = []
            for t_arr, j in jobs:
                d = t_arr - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                ids.append(srv.submit_job(j))
            assert srv.wait(timeout=600.0)
            st = srv.stats()
            lat = []
            failed = 0
            for i in ids:
                try:
                    lat.append(srv.result(i)[3]["latency_s"])
                except Exception as e:  # a failed job is reported, not hidden
                    failed += 1
                    log(f"  job {i} failed: {e}")
        rec = {
            "rate_hz": rate,
            "jobs": len(jobs),
            "failed": failed,
            "p50_s": st["p50_s"],
            "p99_s": st["p99_s"],
            "value": st["atom_steps"] / st["span_s"],
            "span_s": st["span_s"],
            "max_resident": int(st["max_resident"]),
            "rounds": int(st["rounds"]),
        }
        log(f"served at {rate} jobs/s: {rec}")
        out["rates"].append(rec)
        if args.no_plain:
            continue
        if plain_service is None:
            plain_service = []
            for _, j in jobs:
                tc, tr = plain_run(j)
                plain_service.append(tc + tr)
            plain_rate = work / sum(plain_service)
            log(f"  plain one at a time: {sum(plain_service):.3f} s, {plain_rate:.4e} atom-steps/s")
        # the same arrivals served first-come first-served by the plain path
        done = 0.0
        plat = []
        for (t_arr, _), ts in zip(jobs, plain_service):
            done = max(done, t_arr) + ts
            plat.append(done - t_arr)
        rec["plain_fifo_p50_s"] = float(np.percentile(plat, 50))
        rec["plain_fifo_p99_s"] = float(np.percentile(plat, 99))
        rec["speedup_vs_plain"] = rec["value"] / plain_rate
    if plain_service is not None:
        out["plain_value"] = plain_rate
    out["value"] = max(r["value"] for r in out["rates"])
    out["unit"] = "atom-steps/s"
    return out


def served_batch_leg(args):
    """The served jobs (all arrived at once) in one batch on the calling
    thread, run until every job is done: the served mix's batched rate
    without the serving layer (no admission rounds; finished jobs stay in the
    batch and their tiles return at once)."""
    from lpmd2_amd import workloads as W
    from lpmd2_amd.batch import Batch

    jobs = [j for _, j in W.served_jobs(seed=5, n_jobs=args.serve_jobs, rate_hz=1e6)]
    work = sum(j["x"].shape[0] * j["steps"] for j in jobs)
    padded = sum(-(-j["x"].shape[0] // 64) * 64 for j in jobs)
    b = Batch(atom_capacity=padded + 64, list_words=padded * 400, cell_slots=max(4096, padded))
    sids = [b.add(x=j["x"], box=j["box"], z=j["z"], potentials=j["potentials"], cm=j["cm"], v=j["v"], dt=j["dt"],
                  steps=j["steps"]) for j in jobs]
    t0 = time.perf_counter()
    graph_steps = 0
    while True:
        left, _ = b.progress(max(sids) + 1)
        if not np.any(left > 0):
            break
        b.run(64)
        graph_steps += 64
    t = time.perf_counter() - t0
    b.close()
    out = {"value": work / t, "unit": "atom-steps/s", "jobs": len(jobs), "atom_steps": work, "seconds": t,
           "graph_steps": graph_steps,
           "workload": "the served jobs (seed 5) all resident in one batch on the calling thread"}
    log(f"served jobs in one batch: {work:.3e} atom-steps in {t:.3f} s, {work / t:.4e} atom-steps/s")
    return out


def force_check(w, s, n_sample=20000, seed=99):
    """The relative force RMS ||F_gpu - F_64|| / ||F_64|| of the plain path at
    its current positions against the FP64 oracle (the checker): the
    accelerations of the last step's force evaluation against the oracle's on
    the same positions, on every atom of a small box (through the oracle's own
    cell manager of the workload) or on a seeded sample of atoms of a large
    one (the exact minimum-image set the skin list holds)."""
    from oracle import oracle as O

    x, _, a = s.get_state()
    n = x.shape[0]
    kw = dict(zip(w["pot"][1].split()[0::2], (float(q) for q in w["pot"][1].split()[1::2])))
    pot = O.Pot(w["pot"][0], **kw)
    L = np.full(3, w["nx"] * w["a"])
    conf = O.Conf(x, L, z=w["z"])
    if w["cm"][0] == "linkedcell":
        args = dict(zip(w["cm"][1].split()[0::2], w["cm"][1].split()[1::2]))
        O.update_forces(conf, O.CellManager("linkedcell", float(args["cutoff"]), mode=args.get("mode", "noauto")),
                        [pot])
        ref, got, sample = conf.a, a, f"all {n} atoms, oracle linkedcell {w['cm'][1]}"
    else:
        idx = np.sort(np.random.default_rng(seed).choice(n, size=min(n, n_sample), replace=False))
        ref = O.sample_forces(conf, [pot], w["rc"], idx)
        got, sample = a[idx], f"{len(idx)} seeded atoms, exact minimum-image set within rc"
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return {"rel_force_rms": err, "tolerance": 1e-5, "sample": sample}


def plain_leg(name: str, steps: int, warmup: int, lanes: int):
    """One workload on the plain single-system path: `warmup` untimed steps,
    then `steps` timed ones (one integrator call, the state resident in HBM),
    the force kernel's roofline from HIP events on the engine's stream, the
    VALU rates, and after the timed region the relative force RMS against
    the FP64 oracle and the total-energy drift over the timed steps."""
    w = WORKLOADS[name]
    log(f"workload {name}: {w['desc']}")
    s, x, L, v = make_system(w)
    if lanes:
        s.set_lanes(lanes)
    n = s.n
    if w.get("integrator") == "leapfrog":
        s.leapfrog_init(w["dt"])
        advance = s.leapfrog
    elif w.get("integrator") == "beeman":
        s.beeman_init(w["dt"])
        advance = s.beeman
    else:
        advance = s.velocityverlet
    t0 = time.perf_counter()
    ep0, ek0 = advance(w["dt"], max(warmup, 1), obs_every=max(warmup, 1))
    log(f"warmup {warmup} steps: {time.perf_counter() - t0:.3f} s; stats {s.stats()}")

    # HIP events bracket the launches of every PROF_EVERY-th step (and every
    # list build): each event costs the stream about 2.5 us, so bracketing
    # every launch would add about 10 us to each EAM step
    s.profile(True, every=PROF_EVERY)
    st0 = s.stats()
    t0 = time.perf_counter()
    advance(w["dt"], steps)
    dt_wall = time.perf_counter() - t0  # the integrator returns after the stream is synchronised
    st1 = s.stats()
    prof = s.profile_read()
    s.profile(False)
    value = n * steps / dt_wall
    log(f"timed {steps} steps: {dt_wall:.4f} s, {value:.4e} atom-steps/s, profile {prof}, stats {st1}")

    # dominant kernel: the force pass (EAM pass 2 or the pair kernel)
    nbar = st1["list_entries"] / n
    f_ms, f_n = prof["force"]
    r_ms, r_n = prof["rho"]
    if w["metal"]:
        bytes_force = n * (156 + 4 * nbar)
        kname = "k_eam_force"
    else:
        bytes_force = n * (148 + 4 * nbar)
        kname = "k_pair_force"
    if w.get("integrator") == "beeman":
        bytes_force += n * 48  # the previous step's accelerations, read and written
    if st1["brick"]:
        kname += "_bk"  # the LDS-staged variant, one workgroup per brick of cells
    avg_force_s = f_ms / max(f_n, 1) / 1e3
    achieved = bytes_force / avg_force_s / 1e9
    roofline = {
        "kernel": kname,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(name, kname),
        "avg_launch_us": round(avg_force_s * 1e6, 2),
        "launches": f_n,
        "events_every_steps": PROF_EVERY,
        "bytes_per_launch": int(bytes_force),
        "mean_list_len": round(nbar, 2),
    }
    if w["metal"] and r_n:
        roofline["rho_pass_avg_us"] = round(r_ms / r_n * 1e3, 2)
    # VALU rates of the same launch from the algorithmic flop counts above; the
    # pairs inside the cutoff are the listed ones scaled by (rc / r_list)^3
    rc, rlist = w["rc"], w["rc"] + w["skin"]
    ncut = nbar * (rc / rlist) ** 3
    f64 = n * (nbar * FP64_PER_LISTED_PAIR + ncut * FP64_PER_CUT_PAIR) / avg_force_s / 1e12
    f32 = n * ncut * FP32_PER_CUT_PAIR[w["pot"][0]] / avg_force_s / 1e12
    valu = {
        "fp32_tflops": round(f32, 3),
        "fp32_frac": round(f32 / FP32_PEAK_TFLOPS, 4),
        "fp64_tflops": round(f64, 3),
        "fp64_frac": round(f64 / FP64_PEAK_TFLOPS, 4),
        "pairs_in_cutoff_per_atom": round(ncut, 2),
    }
    step_bytes = (160 + 8 * nbar) if w["metal"] else (144 + 4 * nbar)
    # accuracy after the timed steps (outside the timed region): forces
    # against FP64, and the total energy one step on against the warmup's last
    parity = force_check(w, s)
    ep1, ek1 = advance(w["dt"], 1, obs_every=1)
    e0, e1 = float(ep0[-1] + ek0[-1]), float(ep1[-1] + ek1[-1])
    parity["energy_drift_rel"] = (e1 - e0) / abs(e0)
    parity["energy_drift_steps"] = steps + 1
    log(f"  parity {parity}")
    out = {
        "value": value,
        "ms_per_step": 1e3 * dt_wall / steps,
        "steps": steps,
        "config": {"workload": name, "atoms": n, "path": "plain single-system", "desc": w["desc"],
                   "integrator": w.get("integrator", "velocityverlet")},
        "roofline": roofline,
        "step_hbm_fraction": round(value * step_bytes / 1e9 / HBM_PEAK_GBS, 4),
        "valu": valu,
        "list_builds": st1["builds"] - st0["builds"],
        "parity": parity,
    }
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="sc_cu_131k", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=30.0, help="seconds of CPU baseline sampling")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per atom (0 = automatic)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--plain-steps", type=int, default=100, help="steps per system of the plain comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the served leg")
    ap.add_argument("--served-batch", action="store_true",
                    help="also run the served jobs all at once in one batch on this thread")
    ap.add_argument("--serve-jobs", type=int, default=64)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[25.0, 100.0, 1e6])
    ap.add_argument("--serve-budget", type=int, default=16, help="server HBM budget in GiB")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a copy of known size that calibrates "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-at-a-time plain comparisons of the batched and served legs")
    ap.add_argument("--legs", nargs="*", default=["lj_4m", "lj_ar_2048"], choices=sorted(WORKLOADS),
                    help="further plain-path workloads reported beside the main one")
    ap.add_argument("--no-legs", action="store_true", help="skip the further plain-path workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: each box has one MI355X and multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import lpmd2_amd as E

    if E.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
    out = plain_leg(args.workload, args.steps, args.warmup, args.lanes)
    out = {
        "metric": "atom-steps/s plain vs batched/served at <=1e-5 rel force err; p50/p99 latency",
        "value": out["value"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        **{k: v for k, v in out.items() if k not in ("value", "ms_per_step")},
    }
    w = WORKLOADS[args.workload]
    # the other plain legs the targets are stated on: the 4M-atom LJ step
    # (HBM roofline) and the reference's 2048-atom argon case
    if not args.no_legs:
        out["legs"] = {}
        for name in args.legs:
            if name == args.workload:
                continue
            lw = WORKLOADS[name]
            out["legs"][name] = plain_leg(name, lw.get("steps", 200), 20, 0)
    if not args.no_batch:
        out["batched"] = batched_leg(args)
    if not args.no_serve:
        out["served"] = served_leg(args)
    if args.served_batch:
        out["served_in_one_batch"] = served_batch_leg(args)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

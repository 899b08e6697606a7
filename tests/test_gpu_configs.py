"""GPU parity at the sizes the benchmark times, against the FP64 oracle.

- 2048-atom argon on linkedcell `mode auto` (the reference's CPU-runnable
  case): forces, energy, virial and stress at every atom, and a 1000-step
  velocity-Verlet run whose total energy follows the oracle's.
- 131072-atom Sutton-Chen Cu (LDS-staged EAM kernels) and the 4,000,000-atom
  LJ liquid (LDS-staged pair kernel): the step kernels' accelerations and the
  force call's energy, virial and stress against the oracle on an exact cell
  list (linkedcell with one cell a side and cells no smaller than the
  cutoff lists every pair within it once).
- A sample of the 512-system batched ensemble across its size range, and the
  Finnis-Sinclair / Gupta / rocksalt Buckingham jobs of the served workload
  through the batch engine and the server.

Bars: relative force RMS <= 1e-5 against FP64; energy 1e-6 relative; the
virial and stress 1e-5 of the energy scale (they sum terms that largely
cancel near the lattice constant).
"""
import os
import sys

import numpy as np
import pytest

import lpmd2_amd as E
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  the benchmark's workload definitions

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rms(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_pot(name, args, s1=-1, s2=-1):
    kw = dict(zip(args.split()[0::2], (float(q) for q in args.split()[1::2])))
    return O.Pot(name, s1, s2, **kw)


def check_energies(e, w, S, e0, w0, S0):
    esc = abs(e0)
    assert abs(e - e0) <= 1e-6 * esc, (e, e0)
    assert abs(w - w0) <= 1e-5 * max(abs(w0), esc), (w, w0)
    assert np.abs(S - S0).max() <= 1e-5 * max(np.abs(S0).max(), esc)


def test_c0_argon_2048_forces_and_trajectory():
    """Forces, energy, virial and stress of the 2048-atom argon box on the
    reference's truncated linkedcell stencil, then 1000 steps of 1 fs: the
    sampled total energy follows the oracle's to 1e-6 and fluctuates by the
    same amount."""
    w = bench.WORKLOADS["lj_ar_2048"]
    s, x, L, v = bench.make_system(w)
    say("C0", s.n)
    e, wv, S = s.update_forces()
    _, _, a = s.get_state()
    conf = O.Conf(x, L, z=18, v=v)
    cm = O.CellManager("linkedcell", 8.5, mode="auto")
    pots = [oracle_pot(*w["pot"])]
    e0, w0, S0 = O.update_forces(conf, cm, pots)
    err = rel_rms(a, conf.a)
    say(f"  rel force rms {err:.2e}, E {e:.9g} vs {e0:.9g}")
    assert err <= 1e-5
    check_energies(e, wv, S, e0, w0, S0)
    steps = 1000
    ep, ek = s.velocityverlet(w["dt"], steps, obs_every=50)
    et = ep + ek
    conf = O.Conf(x, L, z=18, v=v)
    e_o, k_o, _ = O.md_run(conf, O.CellManager("linkedcell", 8.5, mode="auto"), pots, w["dt"], steps)
    et0 = (e_o + k_o)[49::50]
    spread, spread0 = et.max() - et.min(), et0.max() - et0.min()
    say(f"  E_tot {et[0]:.6f}..{et[-1]:.6f} (spread {spread:.2e}) oracle {et0[0]:.6f}..{et0[-1]:.6f} "
        f"(spread {spread0:.2e})")
    np.testing.assert_allclose(et, et0, rtol=1e-6)
    assert abs(spread - spread0) <= 0.25 * spread0 + 1e-6 * abs(et0[0])


@pytest.mark.parametrize("name,nx", [("sc_cu_131k", 15), ("lj_4m", 67)])
def test_large_boxes_on_the_staged_kernels(name, nx):
    """The LDS-staged (bricked) kernels the benchmark times, at full size:
    after steps with list rebuilds, the accelerations the step kernels left
    and one force call's energy, virial and stress against the oracle's
    full evaluation on an exact cell list."""
    w = bench.WORKLOADS[name]
    s, x, L, v = bench.make_system(w)
    say(name, s.n)
    s.velocityverlet(w["dt"], 30)
    st = s.stats()
    say(f"  stats {st}")
    assert st["brick"] > 0 and st["builds"] >= 2
    xg, _, a_step = s.get_state()
    e, wv, S = s.update_forces()
    _, _, a_call = s.get_state()
    conf = O.Conf(xg, L, z=w["z"])
    pots = [oracle_pot(*w["pot"])]
    e0, w0, S0 = O.update_forces(conf, O.CellManager("linkedcell", w["rc"], nx, nx, nx, mode="noauto"), pots)
    err_step, err_call = rel_rms(a_step, conf.a), rel_rms(a_call, conf.a)
    say(f"  rel force rms step {err_step:.2e} call {err_call:.2e}; E {e:.10g} vs {e0:.10g}")
    assert err_step <= 1e-5 and err_call <= 1e-5
    check_energies(e, wv, S, e0, w0, S0)


def test_c3_batched_ensemble_sample():
    """24 systems spread over the size range of the 512-system ensemble (LJ
    and Sutton-Chen Cu/Ni/Ag) in one batch: accelerations after 20 steps
    against the oracle's exact-set forces at the same positions, energies
    against its full evaluation for the systems up to 4000 atoms."""
    from lpmd2_amd import workloads as W
    from lpmd2_amd.batch import Batch

    ens = W.batched_ensemble(seed=3, n_systems=512, steps=20)
    order = np.argsort([j["x"].shape[0] for j in ens])
    pick = [ens[k] for k in order[np.linspace(0, len(ens) - 1, 24).astype(int)]]
    padded = sum(-(-j["x"].shape[0] // 64) * 64 for j in pick)
    b = Batch(atom_capacity=padded + 64, list_words=padded * 320)
    sids = [b.add(x=j["x"], box=j["box"], z=j["z"], potentials=j["potentials"], cm=j["cm"], v=j["v"], dt=j["dt"],
                  steps=j["steps"]) for j in pick]
    b.run(25)
    worst = 0.0
    for sid, j in zip(sids, pick):
        x, _, a, info = b.get(sid)
        assert info["done"] == 20 and info["overflow"] == 0
        rc = float(dict(zip(j["cm"].split()[0::2], j["cm"].split()[1::2]))["cutoff"])
        pots = [oracle_pot(*p) for p in j["potentials"]]
        conf = O.Conf(x, j["box"], z=j["z"])
        n = x.shape[0]
        idx = np.arange(n) if n <= 3000 else np.sort(np.random.default_rng(n).choice(n, 3000, replace=False))
        ref = O.sample_forces(conf, pots, rc, idx)
        err = rel_rms(a[idx], ref)
        worst = max(worst, err)
        assert err <= 1e-5, (j["potentials"][0][0], n, err)
        if n <= 4000:
            ep, _, _ = b.energies(sid)
            e0, _, _ = O.update_forces(conf, O.CellManager("verletlist", rc, maxneighbors=600, each=5), pots)
            assert abs(ep - e0) <= 1e-6 * abs(e0), (n, ep, e0)
    say(f"  24 systems {pick[0]['x'].shape[0]}..{pick[-1]['x'].shape[0]} atoms, worst rel force rms {worst:.2e}")


def served_sample():
    """Jobs of the served workload (seed 5): the first of each kind
    (Gupta, Finnis-Sinclair, rocksalt Buckingham) and the largest, cut to 15
    steps."""
    from lpmd2_amd import workloads as W

    jobs = [j for _, j in W.served_jobs(seed=5, n_jobs=64, rate_hz=1e6)]
    kinds = {}
    for j in jobs:
        kinds.setdefault(j["potentials"][0][0], j)
    pick = list(kinds.values()) + [max(jobs, key=lambda j: j["x"].shape[0])]
    for j in pick:
        j["steps"] = 15
    return pick


def check_job(j, x, a):
    rc = float(dict(zip(j["cm"].split()[0::2], j["cm"].split()[1::2]))["cutoff"])
    pots = [oracle_pot(*p) for p in j["potentials"]]
    conf = O.Conf(x, j["box"], z=j["z"])
    n = x.shape[0]
    idx = np.arange(n) if n <= 4000 else np.sort(np.random.default_rng(n).choice(n, 4000, replace=False))
    err = rel_rms(a[idx], O.sample_forces(conf, pots, rc, idx))
    say(f"  {j['potentials'][0][0]} x{len(j['potentials'])}: {n} atoms, rel force rms {err:.2e}")
    assert err <= 1e-5


def test_c4_served_jobs_batch_and_server():
    """Finnis-Sinclair, Gupta and rocksalt Buckingham jobs of the served
    workload (the last on the batch's one-selected-term path) through the
    batch engine and through the server: accelerations after 15 steps against
    the oracle, and the same results from both."""
    from lpmd2_amd.batch import Batch
    from lpmd2_amd.server import Server

    pick = served_sample()
    padded = sum(-(-j["x"].shape[0] // 64) * 64 for j in pick)
    b = Batch(atom_capacity=padded + 64, list_words=padded * 400)
    sids = [b.add(x=j["x"], box=j["box"], z=j["z"], potentials=j["potentials"], cm=j["cm"], v=j["v"], dt=j["dt"],
                  steps=j["steps"]) for j in pick]
    b.run(20)
    res_b = []
    for sid, j in zip(sids, pick):
        x, v, a, info = b.get(sid)
        assert info["done"] == 15
        check_job(j, x, a)
        res_b.append((x, v, a))
    b.close()
    with Server(8 << 30, max_round_steps=64) as srv:
        ids = [srv.submit_job(j) for j in pick]
        assert srv.wait(timeout=120.0)
        for i, (xb, vb, ab) in zip(ids, res_b):
            x, v, a, info = srv.result(i)
            np.testing.assert_array_equal(x, xb)
            np.testing.assert_array_equal(v, vb)
            np.testing.assert_array_equal(a, ab)

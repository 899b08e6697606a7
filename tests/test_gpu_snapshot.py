"""Asynchronous state snapshots (lpmdx_snapshot_*) of a running system, and
the configuration files written from them.

A snapshot taken after 30 steps, with the system stepping on while the copy
is in flight, holds exactly (bit for bit) the state a twin system reaches
after 30 steps; its atomic numbers, cell and fixed-atom flags are the
system's; the lpmd / zlp and Dump files written from it read back to the
precision those formats keep."""
import sys

import numpy as np
import pytest

import lpmd2_amd as E
from lpmd2_amd import snapshots as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(fv=None, fp=None):
    x, L = E.fcc(4, 5.26)
    x = np.mod(x + np.random.default_rng(31).normal(0.0, 0.05, x.shape), L)
    v = E.maxwell_boltzmann(np.full(x.shape[0], 39.948), 120.0, seed=31)
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 0.4")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    if fv is not None:
        s.set_fixed(fixedvel=fv, fixedpos=fp)
    return s, L


def test_snapshot_mid_run_equals_the_state_at_its_step(tmp_path):
    say("snapshot mid-run")
    n = 256
    fv = (np.arange(n) % 17) == 3
    fp = (np.arange(n) % 23) == 5
    s1, L = make(fv, fp)
    s2, _ = make(fv, fp)
    s1.velocityverlet(2.0, 30)
    snap = s1.snapshot(step=30)
    # the system steps on while the copy is in flight
    s1.velocityverlet(2.0, 20)
    st = snap.state()
    s2.velocityverlet(2.0, 30)
    x2, v2, a2 = s2.get_state()
    assert st["step"] == 30
    assert np.array_equal(st["x"], x2) and np.array_equal(st["v"], v2) and np.array_equal(st["a"], a2)
    assert np.all(st["z"] == 18)
    np.testing.assert_array_equal(np.diag(st["cell"]), L)
    exp_flags = np.where(fv, E.FIX_VEL, 0) | np.where(fp, E.FIX_POS, 0) | np.where(fv | fp, E.FIX_TAGGED, 0)
    np.testing.assert_array_equal(st["flags"], exp_flags)
    # the system itself has moved on
    assert not np.array_equal(s1.get_state()[0], x2)
    # a reused snapshot object, taken again at step 50
    s1.snapshot(step=50, into=snap)
    st50 = snap.state()
    s2.velocityverlet(2.0, 20)
    assert st50["step"] == 50 and np.array_equal(st50["x"], s2.get_state()[0])

    # files written from the snapshots
    zlp = tmp_path / "traj.zlp"
    with S.LpmdWriter(zlp, "level 2") as w:
        for t in (st, st50):
            w.write(t["cell"], t["z"], t["x"], t["v"], t["a"])
    frames = S.read_lpmd(zlp)
    assert len(frames) == 2
    for fr, t in zip(frames, (st, st50)):
        assert np.abs(S.cartesian(fr["s"], fr["cell"]) - t["x"]).max() <= 5e-7 * L.max()
        assert np.abs(fr["v"] - t["v"]).max() <= 5e-7
    dump = tmp_path / "last.dump"
    S.write_dump(dump, st50["step"], st50["cell"], st50["z"], st50["x"], st50["v"], st50["a"], st50["flags"])
    r = S.read_dump(dump)
    assert r["step"] == 50
    np.testing.assert_array_equal(r["flags"], exp_flags)
    assert np.all(np.abs(r["x"] - st50["x"]) <= 5e-6 * np.abs(st50["x"]) + 1e-300)
    say("  ok")


def test_restored_dump_continues_the_run():
    """A system rebuilt from a Dump (six significant digits, as the reference
    writes it) continues close to the original run: after 20 more steps the
    positions differ by far less than an atom's step."""
    say("restore")
    s1, L = make()
    s1.velocityverlet(2.0, 10)
    st = s1.snapshot(step=10).state()
    import tempfile
    import os

    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "c.dump")
        S.write_dump(p, st["step"], st["cell"], st["z"], st["x"], st["v"], st["a"])
        r = S.read_dump(p)
    s2 = E.System(r["x"], np.diag(r["cell"]), r["z"], v=r["v"])
    s2.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 0.4")
    s2.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    s1.velocityverlet(2.0, 20)
    s2.velocityverlet(2.0, 20)
    dx = s1.get_state()[0] - s2.get_state()[0]
    dx -= L * np.round(dx / L)
    say(f"  max |dx| after 20 steps from the dump: {np.abs(dx).max():.2e}")
    assert np.abs(dx).max() < 1e-3

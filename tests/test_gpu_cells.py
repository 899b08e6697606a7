"""GPU checks of the batch's list build on half-radius cells (cells of at
least half the list radius; the rows of a 5 x 5 x 5 stencil trimmed per
atom), a shape the batch's autotuner may choose, on boxes where five such
cells fit along every axis: forces match the FP64 oracle, and the batch gives
the same trajectories with one or two cells per list radius, and lets the
autotuner choose between them without touching the state."""
import sys

import numpy as np
import pytest

import lpmd2_amd as E
from lpmd2_amd.batch import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # name, args, oracle keywords, Z, lattice constant, fcc cells, jitter, skin
    ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22",
     dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22), 29, 3.61, 8, 0.05, 0.3),
    ("gupta", "A 0.0855 r0 2.556 p 10.960 B 1.2240 q 2.278 cutoff 7.22",
     dict(A=0.0855, r0=2.556, p=10.960, B=1.2240, q=2.278, cutoff=7.22), 29, 3.61, 7, 0.05, 0.3),
    ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5", dict(sigma=3.41, epsilon=0.0103, cutoff=8.5),
     18, 5.26, 6, 0.1, 1.0),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def jittered(nx, a, jit, seed):
    x, L = E.fcc(nx, a)
    return np.mod(x + np.random.default_rng(seed).normal(0.0, jit, x.shape), L), L


def batch_system(name, args, z, a, nx, jit, skin, seed, T):
    x, L = jittered(nx, a, jit, seed)
    m = np.full(x.shape[0], E.ELEMENT_MASS[z])
    v = E.maxwell_boltzmann(m, T, seed=seed + 1)
    rc = float(args.split("cutoff")[1].split()[0])
    return dict(x=x, box=L, z=z, potentials=[(name, args)], cm=f"cutoff {rc} skin {skin}", v=v, dt=2.0)


def test_batch_half_radius_cells():
    """Forces of the initial evaluation against the oracle, then 60 steps
    (lists rebuilt on the way) with two and with one cell per list radius:
    the same neighbour sets, so the same trajectories up to the order of the
    force sums."""
    say("batch half-radius cells")
    specs = [batch_system(c[0], c[1], c[3], c[4], c[5], c[6], c[7], 40 + k, 900.0) for k, c in enumerate(CASES)]
    runs = {}
    for per_radius in (2, 1):
        b = Batch(atom_capacity=8192)
        for c in CASES:
            b.set_cells(c[0], per_radius)
        sids = [b.add(steps=60, **s) for s in specs]
        b.run(1)  # the initial force evaluation
        for sid, s, c in zip(sids, specs, CASES):
            _, _, a0, info = b.get(sid)
            conf = O.Conf(s["x"], s["box"], z=s["z"])
            O.update_forces(conf, O.CellManager("verletlist", c[2]["cutoff"], maxneighbors=400, each=5),
                            [O.Pot(c[0], **c[2])])
            err = np.linalg.norm(a0 - conf.a) / np.linalg.norm(conf.a)
            say(f"  {per_radius} cells per radius, {c[0]}: rel force rms {err:.2e}")
            assert err <= 1e-5 and info["overflow"] == 0
        b.run(60)
        runs[per_radius] = [b.get(sid) for sid in sids]
        b.close()
    for (x2, v2, _, i2), (x1, v1, _, i1), s in zip(runs[2], runs[1], specs):
        assert i2["done"] == i1["done"] == 60 and i2["builds"] > 1
        dx = x2 - x1
        dx -= s["box"] * np.round(dx / s["box"])
        say(f"  max |dx| {np.abs(dx).max():.2e}, builds {i2['builds']}")
        assert np.abs(dx).max() <= 1e-9 and np.abs(v2 - v1).max() <= 1e-11


def test_autotune_chooses_cells_and_keeps_state():
    say("autotune: lanes, cells per list radius, tiles per workgroup")
    specs = [batch_system(c[0], c[1], c[3], c[4], c[5], c[6], c[7], 60 + k, 600.0) for k, c in enumerate(CASES)]
    b = Batch(atom_capacity=8192)
    sids = [b.add(steps=30, **s) for s in specs]
    b.run(5)
    before = [b.get(s)[:2] for s in sids]
    shape = b.autotune(trial_steps=5, cells=True)
    after = [b.get(s)[:2] for s in sids]
    for (x0, v0), (x1, v1) in zip(before, after):
        np.testing.assert_array_equal(x0, x1)
        np.testing.assert_array_equal(v0, v1)
    say(f"  {shape}")
    for c in CASES:
        lanes, cells, tiles = shape[c[0]]
        assert lanes in (1, 2, 4) and cells in (1, 2) and tiles in (1, 2, 4)
    b.run(40)
    for s in sids:
        assert b.get(s)[3]["done"] == 30
    b.close()

"""GPU check of the brick-staged force kernels (bricked cell order, lists of
halo slots, each brick's halo copied into LDS by its workgroup).  Child
processes run the same systems with LPMDX_BRICK=0 (global-index lists, the
unstaged kernels), 2 and 3 (bricks of 2^3 and 3^3 cells):

- the exported neighbour lists must be identical (same pairs, same r^2);
- forces, energies and virials must agree to FP64 summation-order rounding
  (each pair term is computed from the same FP64 displacement in the same
  FP32 arithmetic; only the order of the FP64 per-atom sums changes with the
  atom order), and 40 velocity-Verlet steps with list rebuilds must stay
  within 1e-9 A and 1e-9 A/fs;
- a bricked run repeated in a fresh process must be bit-identical.

The cases cover every staged part: positions only (LJ), positions and
rho^-1/2 (Sutton-Chen, both EAM passes), species (rocksalt Buckingham, three
species-pair terms) and species with rho^-1/2 (Sutton-Chen on the Cu-Cu pairs
of a Cu/Ni mix).  The mix with 3^3-cell bricks needs more LDS than a
workgroup has (about 5100 halo atoms at 32 bytes): it keeps the bricked cell
order with global-index lists and the unstaged kernels, the fallback path."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu

SCRIPT = r'''
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import lpmd2_amd as E
from lpmd2_amd.workloads import rocksalt, BUCKINGHAM_NACL
SC = "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22"

def fcc_case(n, a, z, noise, seed):
    x, L = E.fcc(n, a)
    x = np.mod(x + np.random.default_rng(seed).normal(0.0, noise, x.shape), L)
    return x, L, np.full(x.shape[0], z, dtype=np.int32)

cases = []
x, L, z = fcc_case(13, 3.61, 29, 0.03, 4)
cases.append(("sc", x, L, z, [("suttonchen", SC)], "cutoff 7.22 maxneighbors 400 skin 0.3", 2))
x, L, z = fcc_case(10, 5.26, 18, 0.03, 5)
cases.append(("lj", x, L, z, [("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")],
              "cutoff 8.5 maxneighbors 400 skin 1.0", 1))
x, L, z = rocksalt(8, 8, 8, 5.64)
x = np.mod(x + np.random.default_rng(6).normal(0.0, 0.02, x.shape), L)
cases.append(("bk", x, L, z, [("buckingham", a, s1, s2) for a, s1, s2 in BUCKINGHAM_NACL],
              "cutoff 8.0 maxneighbors 400 skin 0.4", 1))
x, L, z = fcc_case(13, 3.61, 29, 0.03, 7)
z[np.random.default_rng(8).random(z.shape[0]) < 0.3] = 28
cases.append(("scmix", x, L, z, [("suttonchen", SC, 29, 29)], "cutoff 7.22 maxneighbors 400 skin 0.3", 2))

out = {}
for name, x, L, z, pots, cm, lanes in cases:
    mass = np.array([E.ELEMENT_MASS[int(q)] for q in z])
    v = E.maxwell_boltzmann(mass, 500.0, seed=2)
    s = E.System(x, L, z, v=v)
    s.set_lanes(lanes)
    s.set_cellmanager("verletlist", cm)
    for p in pots:
        s.append_potential(*p)
    e, w, S = s.update_forces()
    _, _, a0 = s.get_state()
    rc = float(cm.split()[1])
    counts, js, r2 = s.neighbor_lists(True, rc)
    s.velocityverlet(1.0 if name == "bk" else 2.0, 40)
    xs, vs, _ = s.get_state()
    out[name] = dict(e=e, w=w, S=np.asarray(S).tolist(), a=a0.tolist(), x=xs.tolist(), v=vs.tolist(),
                     counts=counts.tolist(), js=js.tolist(), r2=r2.tolist(), brick=s.stats()["brick"])
print("RESULT " + json.dumps(out))
'''

CASES = ("sc", "lj", "bk", "scmix")
# brick edge each case runs with under LPMDX_BRICK=2 / 3 (0: the fallback)
STAGED = {"2": dict(sc=2, lj=2, bk=2, scmix=2), "3": dict(sc=3, lj=3, bk=3, scmix=0)}


def run(brick: str):
    print(f"LPMDX_BRICK={brick}", flush=True)
    env = dict(os.environ, LPMDX_BRICK=brick)
    p = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    line = [l for l in p.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[7:])


@pytest.fixture(scope="module")
def runs():
    return {b: run(b) for b in ("0", "2", "3")}


@pytest.mark.parametrize("brick", ["2", "3"])
def test_bricked_lists_forces_and_trajectories(runs, brick):
    ref, got = runs["0"], runs[brick]
    for name in CASES:
        r, g = ref[name], got[name]
        assert r["brick"] == 0 and g["brick"] == STAGED[brick][name], (name, r["brick"], g["brick"])
        assert r["counts"] == g["counts"] and r["js"] == g["js"], name
        assert r["r2"] == g["r2"], name
        assert abs(g["e"] - r["e"]) <= 1e-12 * abs(r["e"]), name
        assert abs(g["w"] - r["w"]) <= 1e-12 * abs(r["w"]), name
        assert np.allclose(g["S"], r["S"], rtol=0, atol=1e-12 * np.abs(r["S"]).max()), name
        ar, ag = np.array(r["a"]), np.array(g["a"])
        assert np.abs(ag - ar).max() <= 1e-12 * np.abs(ar).max(), name
        assert np.abs(np.array(g["x"]) - np.array(r["x"])).max() <= 1e-9, name
        assert np.abs(np.array(g["v"]) - np.array(r["v"])).max() <= 1e-9, name


def test_bricked_runs_repeat_bit_for_bit(runs):
    again = run("3")
    for name in CASES:
        for key in ("e", "w", "S", "a", "x", "v"):
            assert again[name][key] == runs["3"][name][key], (name, key)

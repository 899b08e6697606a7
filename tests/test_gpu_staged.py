"""GPU check of the opt-in tile-staged kernels (LPMDX_STAGE=1: lists of halo
slots, neighbour positions copied into LDS per 64-atom tile): one child
process runs the same systems with staging on, and its forces, energies and
trajectories must be bit-identical to the default path's (the lists hold the
same neighbours in the same order, so every sum is taken in the same order)."""
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
from lpmd2_amd.batch import Batch
out = {}
for name, pot, a, z, cm in [
    ("sc", ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22"), 3.61, 29, "cutoff 7.22 maxneighbors 400 skin 0.3"),
    ("lj", ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5"), 5.26, 18, "cutoff 8.5 maxneighbors 400 skin 1.0"),
]:
    x, L = E.fcc(7, a)
    x = np.mod(x + np.random.default_rng(4).normal(0.0, 0.03, x.shape), L)
    v = E.maxwell_boltzmann(np.full(x.shape[0], E.ELEMENT_MASS[z]), 500.0, seed=2)
    s = E.System(x, L, z, v=v)
    s.set_lanes(4)
    s.set_cellmanager("verletlist", cm)
    s.append_potential(*pot)
    e, w, S = s.update_forces()
    _, _, a0 = s.get_state()
    s.velocityverlet(2.0, 40)
    xs, vs, _ = s.get_state()
    counts, js, r2 = s.neighbor_lists(True, float(cm.split()[1]))
    out[name] = dict(e=e, w=w, a=a0.tolist(), x=xs.tolist(), v=vs.tolist(), lists=[counts.tolist(), js.tolist()])
b = Batch(atom_capacity=8192)
x, L = E.fcc(6, 3.61)
x = np.mod(x + np.random.default_rng(9).normal(0.0, 0.02, x.shape), L)
sid = b.add(x=x, box=L, z=29, potentials=[("gupta", "A 0.0855 r0 2.556 p 10.960 B 1.2240 q 2.278 cutoff 7.22")],
            cm="cutoff 7.22 skin 0.3", v=E.maxwell_boltzmann(np.full(x.shape[0], 63.546), 800.0, seed=5),
            dt=2.0, steps=30)
b.set_lanes("gupta", 4)
b.run(31)
xb, vb, ab, info = b.get(sid)
out["batch"] = dict(x=xb.tolist(), v=vb.tolist(), e=list(b.energies(sid)))
print("RESULT " + json.dumps(out))
'''


def run(stage: str):
    env = dict(os.environ, LPMDX_STAGE=stage)
    p = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    line = [l for l in p.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[7:])


def test_staged_kernels_match_default_bit_for_bit():
    print("staged vs default", flush=True)
    ref = run("0")
    stg = run("1")
    for name in ("sc", "lj"):
        assert ref[name]["lists"] == stg[name]["lists"], name
        assert ref[name]["e"] == stg[name]["e"] and ref[name]["w"] == stg[name]["w"], name
        assert np.array_equal(np.array(ref[name]["a"]), np.array(stg[name]["a"])), name
        assert np.array_equal(np.array(ref[name]["x"]), np.array(stg[name]["x"])), name
        assert np.array_equal(np.array(ref[name]["v"]), np.array(stg[name]["v"])), name
    assert np.array_equal(np.array(ref["batch"]["x"]), np.array(stg["batch"]["x"]))
    assert np.array_equal(np.array(ref["batch"]["v"]), np.array(stg["batch"]["v"]))
    assert ref["batch"]["e"] == stg["batch"]["e"]

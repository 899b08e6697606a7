"""Times the partial reference build against the oracle on the same inputs
(this container only: it needs oracle/_ref, built from the reference sources
by oracle/ref.mk) and writes cpu_calibration.json.

The bench's CPU baseline times the oracle; the ratios here say how its speed
compares with the reference's own code on the parts both can run:
  - verletlist rebuild (O(N^2), plugins/verletlist.cc:67-104) on a 2048-atom
    Sutton-Chen Cu slice (8^3 fcc cells with 0.01 A noise), cutoff 7.22;
  - linkedcell UpdateCell + full and half BuildNeighborList of every atom
    (plugins/linkedcell.cc:112-360) on the 2048-atom LJ argon box (config 0).
(The reference's force composition needs its runtime session, which is not
built here, so force calls are not compared.)
ratio = oracle time / reference time (below 1: the oracle is faster, so a
baseline timed with it does not flatter the GPU)."""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lpmd2_amd.lattice import fcc  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import refcheck as R  # noqa: E402


def best(f, reps=5):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        f()
        t.append(time.perf_counter() - t0)
    return min(t)


def main():
    assert R.available(), "oracle/_ref is not built (make -f oracle/ref.mk)"
    out = {}
    # verletlist rebuild, SC Cu slice
    x, L = fcc(8, 3.61)
    x = np.mod(x + np.random.default_rng(1).normal(0.0, 0.01, x.shape), L)
    oc = O.Conf(x, L, z=29)
    rc = R.RefConf(x, L, z=np.full(x.shape[0], 29, dtype=np.int32))

    def o_vl():
        cm = O.CellManager("verletlist", 7.22, maxneighbors=400, each=5)
        cm.update(oc)

    def r_vl():
        cm = R.RefCellManager("verletlist", "cutoff 7.22 maxneighbors 400 each 5")
        cm.update(rc)

    to, tr = best(o_vl), best(r_vl)
    out["verletlist_rebuild_sc2048"] = dict(oracle_s=to, reference_s=tr, ratio=to / tr)
    # linkedcell update + lists, LJ argon config 0
    x, L = fcc(8, 5.26)
    oc = O.Conf(x, L, z=18)
    rc = R.RefConf(x, L, z=np.full(x.shape[0], 18, dtype=np.int32))

    def o_lc():
        cm = O.CellManager("linkedcell", 8.5, mode="auto")
        cm.update(oc)
        cm.lists(oc, True, 8.5)
        cm.lists(oc, False, 8.5)

    def r_lc():
        cm = R.RefCellManager("linkedcell", "cutoff 8.5 mode auto")
        cm.update(rc)
        cm.lists(rc, True, 8.5)
        cm.lists(rc, False, 8.5)

    to, tr = best(o_lc), best(r_lc)
    out["linkedcell_lists_ar2048"] = dict(oracle_s=to, reference_s=tr, ratio=to / tr)
    out["host"] = open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(" :\t")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_calibration.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

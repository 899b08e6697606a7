"""Oracle against the reference's own plugin code, live (oracle/_ref, built by
oracle/ref.mk).  Skipped where that partial build is absent (it needs the
reference sources); tests/test_oracle.py covers the same ground through the
committed golden fixtures."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import refcheck as R

pytestmark = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")


def fcc(nx, a):
    L = np.array([nx * a] * 3, dtype=np.float64)
    off = ((0.51, 0.5, 0.5), (0.51, 0.0, 1.0), (1.01, 0.0, 0.5), (1.01, 0.5, 1.0))
    pts = [[((i + o[0]) / nx) * L[0], ((j + o[1]) / nx) * L[1], ((k + o[2]) / nx) * L[2]]
           for k in range(nx) for j in range(nx) for i in range(nx) for o in off]
    return np.array(pts), L


CASES = [
    # (lattice cells, a, jitter, cell manager, args, rcut)
    (5, 5.26, 0.3, "linkedcell", "cutoff 8.5 mode auto", 8.5),
    (5, 5.26, 0.3, "linkedcell", "cutoff 8.5 nx 3 ny 3 nz 3", 8.5),
    (5, 5.26, 0.3, "linkedcell", "cutoff 8.5 nx 7 ny 5 nz 6", 6.0),
    (3, 3.61, 0.05, "linkedcell", "cutoff 7.22 mode auto", 7.22),
    (6, 3.61, 0.0, "verletlist", "cutoff 7.22 maxneighbors 200 each 5", 7.22),
    (4, 5.26, 0.4, "verletlist", "cutoff 8.5 maxneighbors 120 each 3", 8.5),
]


@pytest.mark.parametrize("nx,a,jit,cm,args,rcut", CASES)
@pytest.mark.parametrize("seed", [1, 2])
def test_lists_live(nx, a, jit, cm, args, rcut, seed):
    x, L = fcc(nx, a)
    x = np.mod(x + np.random.default_rng(seed).normal(0, jit, x.shape), L)
    rc = R.RefConf(x, L)
    rm = R.RefCellManager(cm, args)
    rm.update(rc, attach=True)
    kw = dict(zip(args.split()[0::2], args.split()[1::2]))
    if cm == "linkedcell":
        om = O.CellManager(cm, float(kw["cutoff"]), int(kw.get("nx", 0)), int(kw.get("ny", 0)), int(kw.get("nz", 0)),
                           mode=kw.get("mode", "noauto"))
    else:
        om = O.CellManager(cm, float(kw["cutoff"]), maxneighbors=int(kw["maxneighbors"]), each=int(kw["each"]))
    c = O.Conf(x, L)
    om.update(c)
    for full in (True, False):
        a_ = rm.lists(rc, full, rcut)
        b_ = om.lists(c, full, rcut)
        for p, q in zip(a_, b_):
            assert np.array_equal(p, q)


def test_verletlist_overflow_raises_like_reference():
    x, L = fcc(3, 3.61)
    rc = R.RefConf(x, L)
    rm = R.RefCellManager("verletlist", "cutoff 7.22 maxneighbors 100 each 5")
    with pytest.raises(RuntimeError, match="maxneighbors"):
        rm.update(rc, attach=True)
    om = O.CellManager("verletlist", 7.22, maxneighbors=100, each=5)
    with pytest.raises(O.OracleError, match="maxneighbors"):
        om.update(O.Conf(x, L))

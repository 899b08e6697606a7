"""lpmdx_observables_noalloc against lpmdx_observables through the C ABI: the
same SetSimulationTags values bit for bit on two calls in a row, the second on
a smaller system, and a system above its atom limit refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import lpmd2_amd as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(s):
    L = E.lib()
    L.lpmdx_observables_noalloc.argtypes = [C.c_void_p, C.c_void_p]
    out = np.zeros(E.NOBS)
    rc = L.lpmdx_observables_noalloc(s._h, out.ctypes.data_as(C.c_void_p))
    if rc != E.OK:
        raise E.LpmdxError(rc, L.lpmdx_last_error().decode())
    return out


def system(nx, seed, tagged):
    x, L = E.fcc(nx, 3.61)
    x = np.mod(x + np.random.default_rng(seed).normal(0.0, 0.02, x.shape), L)
    n = x.shape[0]
    v = E.maxwell_boltzmann(np.full(n, 63.546), 400.0, seed=seed) + 2e-4
    s = E.System(x, L, 29, v=v)
    s.set_cellmanager("verletlist", "cutoff 7.22 maxneighbors 300 each 5")
    s.append_potential("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22")
    if tagged:
        k = np.arange(n)
        s.set_fixed(fixedvel=(k % 5) == 1, fixedpos=(k % 9) == 2)
    s.update_forces()
    return s


def test_noalloc_observables_match_the_allocating_call():
    say("noalloc observables")
    big = system(12, 1, False)    # 6912 atoms: 27 partial rows
    small = system(5, 2, True)    # 500 atoms, tagged: 2 partial rows
    for s in (big, small):
        ref = s.observables()
        got = noalloc(s)
        say(f"  n {s.n}: T {got[1]:.6f} vs {ref[1]:.6f}")
        np.testing.assert_array_equal(got, ref)
    # the first system again after the smaller one (stale rows must not count)
    np.testing.assert_array_equal(noalloc(big), big.observables())


def test_noalloc_refuses_systems_above_its_limit():
    say("noalloc limit")
    n = (1 << 22) + 1
    x = np.random.default_rng(3).random((n, 3)) * 400.0
    s = E.System(x, [400.0] * 3, 18)
    with pytest.raises(E.LpmdxError) as e:
        noalloc(s)
    assert e.value.code == E.EINVAL and "at most 4194304" in str(e.value)

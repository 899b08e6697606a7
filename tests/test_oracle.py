"""Pins the FP64 oracle (oracle/lpmd_oracle.c) before anything is checked against it.

Three kinds of checks:
  * golden fixtures written from the reference's own plugin code
    (tests/golden/gen_golden.py): bit-exact neighbour lists, per-pair values,
    velocityverlet trajectories;
  * the reference's own known answers (liblpmd/demos/linkedcell.cc,
    liblpmd/tests/propertiestest.unit, liblpmd/tests/orthogonalcelltest.unit);
  * what the mathematics fixes: closed-form lattice sums, Newton's third law,
    finite-difference derivatives of the energy, the virial under dilation,
    velocity Verlet's exactness for constant acceleration and its O(dt^2)
    energy error.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "golden.json")))


def parse_cm(cm, args, ):
    kw = dict(zip(args.split()[0::2], args.split()[1::2]))
    cut = float(kw["cutoff"])
    if cm in ("linkedcell", "lcbinary"):
        return O.CellManager(cm, cut, int(kw.get("nx", 0)), int(kw.get("ny", 0)), int(kw.get("nz", 0)),
                             mode=kw.get("mode", "noauto"))
    if cm == "minimumimage":
        return O.CellManager("minimumimage", cut)
    return O.CellManager("verletlist", cut, maxneighbors=int(kw["maxneighbors"]), each=int(kw["each"]))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def fcc(nx, a):
    """crystal3d fcc lattice (plugins/crystal3d.cc:84-104) with nx^3 cells of edge a."""
    L = np.array([nx * a] * 3, dtype=np.float64)
    off = ((0.51, 0.5, 0.5), (0.51, 0.0, 1.0), (1.01, 0.0, 0.5), (1.01, 0.5, 1.0))
    ax = 1.0 / nx
    pts = [[((i + o[0]) * ax) * L[0], ((j + o[1]) * ax) * L[1], ((k + o[2]) * ax) * L[2]]
           for k in range(nx) for j in range(nx) for i in range(nx) for o in off]
    return np.array(pts), L


@pytest.mark.parametrize("case", sorted(META["crystal"]))
def test_fcc_helper_matches_crystal3d(case):
    d = np.load(os.path.join(GOLDEN, f"crystal_{case}.npz"))
    x, L = fcc(META["crystal"][case]["nx"], META["crystal"][case]["a"])
    assert np.array_equal(x, d["x"]) and np.array_equal(L, d["L"])


# ---------------------------------------------------------------- golden lists
@pytest.mark.parametrize("case", sorted(META["lists"]))
def test_lists_match_reference(case):
    m = META["lists"][case]
    d = np.load(os.path.join(GOLDEN, f"lists_{case}.npz"))
    conf = O.Conf(d["x"], d["L"])
    cm = parse_cm(m["cm"], m["cm_args"])
    cm.update(conf)
    for full, tag in ((True, "full"), (False, "half")):
        counts, js, r2, rij = cm.lists(conf, full, m["rcut"])
        assert np.array_equal(counts, d[f"{tag}_counts"])
        if f"{tag}_j" in d:
            assert np.array_equal(js, d[f"{tag}_j"])
            assert np.array_equal(r2, d[f"{tag}_r2"])
        assert digest(counts, js, r2, rij) == m[f"{tag}_digest"]


def test_duplicate_stencil_case_has_duplicates():
    d = np.load(os.path.join(GOLDEN, "lists_cu256j_lc_dup.npz"))
    counts = d["full_counts"]
    js = d["full_j"]
    start = np.concatenate([[0], np.cumsum(counts)])
    dup = sum(len(js[start[i]:start[i + 1]]) - len(np.unique(js[start[i]:start[i + 1]])) for i in range(len(counts)))
    assert dup > 0  # 2*side+1 = 5 > n = 4 visits cells twice


def test_lcbinary_cases_grow_and_repeat():
    # random atoms share subcells: the grid grows past the automatic choice
    d = np.load(os.path.join(GOLDEN, "lists_rand300_lcb.npz"))
    conf = O.Conf(d["x"], d["L"])
    cm = O.CellManager("lcbinary", 6.0)
    cm.update(conf)
    g = cm.grid()
    assert g[0] > O.lcb_auto_n(d["L"][0], 6.0)
    # the 256-atom Cu box: a stencil of 11 subcells on a grid of 10 repeats entries
    d = np.load(os.path.join(GOLDEN, "lists_cu256j_lcb.npz"))
    counts, js = d["full_counts"], d["full_j"]
    start = np.concatenate([[0], np.cumsum(counts)])
    dup = sum(len(js[start[i]:start[i + 1]]) - len(np.unique(js[start[i]:start[i + 1]])) for i in range(len(counts)))
    assert dup > 0


def test_minimumimage_matches_brute_force_set():
    # every pair within the radius, by its single minimum image, ascending j
    for case in ("rand300_mi", "rand300_mi_wide"):
        d = np.load(os.path.join(GOLDEN, f"lists_{case}.npz"))
        x, L = d["x"], d["L"]
        rc = META["lists"][case]["rcut"]
        dx = x[None, :, :] - x[:, None, :]
        dx -= L * np.where(dx >= L / 2, 1.0, np.where(dx < -L / 2, -1.0, 0.0))
        r2 = (dx * dx).sum(axis=2)
        np.fill_diagonal(r2, np.inf)
        expect = [np.nonzero(r2[i] < rc * rc)[0] for i in range(len(x))]
        start = np.concatenate([[0], np.cumsum(d["full_counts"])])
        for i in range(len(x)):
            assert np.array_equal(d["full_j"][start[i]:start[i + 1]], expect[i])


def test_shell_cutoff_counts():
    # rc = 2a lies on the 8th fcc shell; FP64 rounding of r2 decides membership
    # (134 within 7.22- in a perfect Cu lattice, 140 up to 7.52)
    d = np.load(os.path.join(GOLDEN, "lists_cu256_shell_vl.npz"))
    assert set(np.unique(d["full_counts"])) <= set(range(134, 141))


# ---------------------------------------------------------------- potentials
@pytest.mark.parametrize("name", sorted(META["pots"]))
def test_potentials_match_reference(name):
    d = np.load(os.path.join(GOLDEN, f"pot_{name}.npz"))
    params = dict(META["pots"][name]["params"])
    if name == "tabulatedpair":
        params["file"] = os.path.join(GOLDEN, params["file"])  # a fixture beside the vectors
    p = O.Pot(name, **params)
    assert p.cutoff == META["pots"][name]["cutoff"]
    rij = d["rij"]
    r = np.sqrt(rij[:, 0] * rij[:, 0] + rij[:, 1] * rij[:, 1] + rij[:, 2] * rij[:, 2])
    # same glibc libm: bit-exact; tolerance only covers a different libm build
    tol = dict(rtol=1e-13, atol=0)
    if p.metal:
        np.testing.assert_allclose(p.pair_energy(r), d["epair"], **tol)
        np.testing.assert_allclose(p.rhoij(r), d["rho"], **tol)
        np.testing.assert_allclose(p.embed(d["rhoi"]), d["frho"], **tol)
        pf, mb = p.metal_forces(rij, d["rhoi"], d["rhoj"])
        np.testing.assert_allclose(pf, d["pf"], **tol)
        np.testing.assert_allclose(mb, d["mb"], **tol)
    else:
        np.testing.assert_allclose(p.pair_energy(r), d["e"], **tol)
        np.testing.assert_allclose(p.pair_force(rij), d["f"], **tol)


# ---------------------------------------------------------------- integrator
def test_velocityverlet_matches_reference():
    d = np.load(os.path.join(GOLDEN, "vv_ar108.npz"))
    conf = O.Conf(d["x0"], d["L"], v=d["v0"])
    acc = d["acc"]
    steps = acc.shape[0] // 2
    for s in range(steps):
        conf.a[:] = acc[2 * s]
        aold = O.vv_position(conf, 1.0)
        conf.a[:] = acc[2 * s + 1]
        O.vv_velocity(conf, 1.0, aold)
    assert np.array_equal(conf.x, d["xf"])
    assert np.array_equal(conf.v, d["vf"])


def _replay(tables):
    calls = iter(tables)

    def forces(c):
        c.a[:] = next(calls)
        return 0.0

    return forces


def test_leapfrog_matches_reference():
    """Leapfrog::Initialize + 12 Advance calls of the reference plugin with
    replayed accelerations; one atom crosses the box several times per step."""
    d = np.load(os.path.join(GOLDEN, "lf_ar108.npz"))
    conf = O.Conf(d["x0"], d["L"], v=d["v0"])
    acc = d["acc"]
    O.leapfrog_run(conf, _replay(acc), 2.0, acc.shape[0] - 1)
    assert np.array_equal(conf.x, d["xf"])
    assert np.array_equal(conf.v, d["vf"])
    assert np.array_equal(conf.a, d["af"])


def test_leapfrog_constant_acceleration_closed_form():
    """With constant a the positions follow x0 + v0 t + a t^2 / 2 exactly and
    the reported velocity is the one at the last force evaluation, v0 + a (t - dt)."""
    L = [1e6] * 3
    x0 = np.array([[10.0, 20.0, 30.0]])
    v0 = np.array([[0.01, -0.02, 0.005]])
    a0 = np.array([[1e-4, 2e-4, -3e-4]])
    conf = O.Conf(x0, L, v=v0)
    conf.a[:] = a0
    dt, steps = 0.5, 40
    O.leapfrog_run(conf, _replay([a0] * (steps + 1)), dt, steps)
    t = dt * steps
    np.testing.assert_allclose(conf.x, x0 + v0 * t + 0.5 * a0 * t * t, rtol=1e-12)
    np.testing.assert_allclose(conf.v, v0 + a0 * (t - dt), rtol=1e-12)


def test_beeman_matches_reference():
    """Beeman::Initialize + 12 TwoStepIntegrator::Advance calls of the
    reference plugin with replayed accelerations (two per step)."""
    d = np.load(os.path.join(GOLDEN, "bm_ar108.npz"))
    conf = O.Conf(d["x0"], d["L"], v=d["v0"])
    acc = d["acc"]
    O.beeman_run(conf, _replay(acc), 2.0, (acc.shape[0] - 1) // 2)
    assert np.array_equal(conf.x, d["xf"])
    assert np.array_equal(conf.v, d["vf"])
    assert np.array_equal(conf.a, d["af"])


def test_beeman_constant_acceleration_closed_form():
    """Constant a: Beeman's weights sum to the exact kinematics, positions
    x0 + v0 t + a t^2 / 2 and velocities v0 + a t."""
    L = [1e6] * 3
    x0 = np.array([[10.0, 20.0, 30.0]])
    v0 = np.array([[0.01, -0.02, 0.005]])
    a0 = np.array([[1e-4, 2e-4, -3e-4]])
    conf = O.Conf(x0, L, v=v0)
    conf.a[:] = a0
    dt, steps = 0.5, 40
    O.beeman_run(conf, _replay([a0] * (2 * steps + 1)), dt, steps)
    t = dt * steps
    np.testing.assert_allclose(conf.x, x0 + v0 * t + 0.5 * a0 * t * t, rtol=1e-12)
    np.testing.assert_allclose(conf.v, v0 + a0 * t, rtol=1e-12)


def test_stepper_known_answers():
    """Stepper::IsActiveInStep / ValueAtStep (stepper.h:17-31) by hand."""
    st = O.Stepper(10, 50, 20)
    assert [k for k in range(0, 80) if st.active(k)] == [10, 30, 50]
    assert [k for k in range(0, 12) if O.Stepper(0, -1, 4).active(k)] == [0, 4, 8]
    ramp = O.Stepper(0, 100, 1)
    assert ramp.value(0, 300.0, 600.0) == 300.0
    assert ramp.value(50, 300.0, 600.0) == 450.0
    assert ramp.value(100, 300.0, 600.0) == 600.0
    assert O.Stepper(7, 7, 1).value(7, 300.0, 600.0) == 300.0


def test_set_temperature_closed_form():
    """Two argon atoms with |v|^2 = 1e-4 and 4e-4 A^2/fs^2: K = 0.5 m 5e-4 KIN2EV,
    T = (2/3) K / (2 kB); after SetTemperature(T') every velocity is scaled by
    sqrt(T' / T), so the directions stay and the temperature reads T'."""
    v = np.array([[0.01, 0.0, 0.0], [0.0, 0.0, -0.02]])
    conf = O.Conf(np.zeros((2, 3)) + [[1, 1, 1], [4, 4, 4]], [20.0] * 3, v=v)
    m = 39.948
    t0 = (2.0 / 3.0) * (0.5 * m * 5e-4 * O.KIN2EV) / (2 * O.KBOLTZMANN)
    assert abs(O.temperature(conf) - t0) <= 1e-12 * t0
    O.set_temperature(conf, 300.0)
    f = np.sqrt(300.0 / t0)
    np.testing.assert_allclose(conf.v, v * f, rtol=1e-14)
    assert abs(O.temperature(conf) - 300.0) <= 1e-12 * 300.0


def test_berendsen_schedule_and_closed_form():
    """berendsen.cc:45-80 by hand: start 0, each 10, tau 20 fs, dt 2 fs arms at
    step 0 (stop 10), scales at steps 1..10, disarms at 11, arms again at 20;
    one scaling moves T to T + (dt / tau) (T_set - T)."""
    v = np.array([[0.01, 0.0, 0.0], [0.0, 0.0, -0.02]])
    conf = O.Conf(np.array([[1.0, 1, 1], [4, 4, 4]]), [20.0] * 3, v=v)
    b = O.Berendsen(300.0, 300.0, tau=20.0, start=0, end=-1, each=10)
    scaled = [k for k in range(0, 35) if b.apply(k, conf, 2.0)]
    assert scaled == list(range(1, 11)) + list(range(21, 31))
    conf = O.Conf(np.array([[1.0, 1, 1], [4, 4, 4]]), [20.0] * 3, v=v)
    t0 = O.temperature(conf)
    b = O.Berendsen(300.0, 300.0, tau=20.0, each=1)
    assert not b.apply(0, conf, 2.0) and b.apply(1, conf, 2.0)
    assert abs(O.temperature(conf) - (t0 + 0.1 * (300.0 - t0))) <= 1e-12 * 300.0


def test_velocityverlet_exact_for_constant_acceleration():
    L = [1e6] * 3
    x0 = np.array([[10.0, 20.0, 30.0]])
    v0 = np.array([[0.01, -0.02, 0.005]])
    a0 = np.array([[1e-4, 2e-4, -3e-4]])
    conf = O.Conf(x0, L, v=v0)
    dt, steps = 0.5, 40
    for _ in range(steps):
        conf.a[:] = a0
        aold = O.vv_position(conf, dt)
        conf.a[:] = a0
        O.vv_velocity(conf, dt, aold)
    t = dt * steps
    np.testing.assert_allclose(conf.x, x0 + v0 * t + 0.5 * a0 * t * t, rtol=1e-12)
    np.testing.assert_allclose(conf.v, v0 + a0 * t, rtol=1e-12)


# ---------------------------------------------------------------- reference known answers
def test_demo_linkedcell_known_answer():
    # liblpmd/demos/linkedcell.cc:83-111: 108 Ar, box 17.1191, one cell, rcut 4.15:
    # 12 neighbours per atom at 4.03501, half list = full/2
    Lb = 17.1191
    pts = []
    for k in range(3):
        for j in range(3):
            for i in range(3):
                for o in ((0.5, 0.5, 0.5), (0.5, 0.0, 1.0), (1.0, 0.0, 0.5), (1.0, 0.5, 1.0)):
                    pts.append([(i + o[0]) / 3 * Lb, (j + o[1]) / 3 * Lb, (k + o[2]) / 3 * Lb])
    x = np.mod(np.array(pts), Lb)
    conf = O.Conf(x, [Lb] * 3)
    cm = O.CellManager("linkedcell", 8.5, 1, 1, 1, mode="noauto")
    cm.update(conf)
    counts, js, r2, _ = cm.lists(conf, True, 4.15)
    assert np.all(counts == 12)
    assert np.all(np.abs(np.sqrt(r2) - 4.03501) < 1e-4)
    hc, _, _, _ = cm.lists(conf, False, 4.15)
    assert counts.sum() == 2 * hc.sum()


def test_properties_known_answers():
    # liblpmd/tests/propertiestest.unit:19-21,61-63
    v = np.zeros((108, 3))
    v[:, 0] = 0.0017669141
    conf = O.Conf(np.zeros((108, 3)), [10.0] * 3, z=18, v=v)
    assert abs(O.kinetic_energy(conf) - 0.69800472) < 1e-5
    assert abs(O.temperature(conf) - 50.0) < 1e-5


def test_fitted_inside_known_answers():
    # liblpmd/tests/orthogonalcelltest.unit:97-116
    L = [20.0] * 3
    np.testing.assert_allclose(O.fitted_inside(L, [10, 10, 10]), [10, 10, 10], atol=1e-10)
    np.testing.assert_allclose(O.fitted_inside(L, [22, 23, 24]), [2, 3, 4], atol=1e-10)
    np.testing.assert_allclose(O.fitted_inside(L, [-3, -7, -4]), [17, 13, 16], atol=1e-10)
    np.testing.assert_allclose(O.fitted_inside(L, [63, 65, 61]), [3, 5, 1], atol=1e-10)
    np.testing.assert_allclose(O.fitted_inside(L, [-83, -87, -84]), [17, 13, 16], atol=1e-10)
    assert O.fitted_inside(L, [20.0, 0.0, 5.0])[0] == 20.0  # x == L is left as is


def test_displacement_long_distances():
    # orthogonalcelltest.unit:136-153: |d| <= L/2, d < 0, |d| + |b-a| = L
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = 1.5 + rng.uniform(-0.57, 0.57, 3)
        b = 8.5 + rng.uniform(-0.57, 0.57, 3)
        d = O.displacement([10.0] * 3, a, b)
        assert np.all(np.abs(d) <= 5.0) and np.all(d < 0)
        np.testing.assert_allclose(np.abs(d) + np.abs(b - a), 10.0, atol=1e-10)


def test_auto_subdivision():
    # C0: L = 42.08, rc 8.5 -> 12 cells per edge, side 2, 125 stencil cells
    assert O.lc_auto_n(42.08, 8.5) == 12
    x, L = fcc(8, 5.26)
    conf = O.Conf(x, L)
    cm = O.CellManager("linkedcell", 8.5, mode="auto")
    cm.update(conf)
    assert list(cm.grid()) == [12, 12, 12, 2, 125]


# ---------------------------------------------------------------- composition invariants
LJ = dict(sigma=3.41, epsilon=0.0103, cutoff=8.5)
FCC_SHELLS = [12, 6, 24, 12, 24, 8, 48, 6, 36, 24, 24, 24, 72, 0, 48, 12, 48, 30, 72, 24]


def lattice_sum(a, rc, fn):
    """Sum over fcc shells r_k = a*sqrt(k/2) below rc of n_k * fn(r_k)."""
    tot = 0.0
    for k, nk in enumerate(FCC_SHELLS, start=1):
        r = a * np.sqrt(k / 2.0)
        if r * r < rc * rc and nk:
            tot += nk * fn(r)
    return tot


def test_lj_lattice_energy_closed_form():
    x, L = fcc(8, 5.26)
    conf = O.Conf(x, L)
    cm = O.CellManager("verletlist", 8.5, maxneighbors=200, each=5)
    ep, vir, S = O.update_forces(conf, cm, [O.Pot("lennardjones", **LJ)])
    s, e = LJ["sigma"], LJ["epsilon"]
    per_atom = 0.5 * lattice_sum(5.26, 8.5, lambda r: 4 * e * ((s / r) ** 12 - (s / r) ** 6))
    assert abs(ep - 2048 * per_atom) < 1e-9 * abs(ep)
    assert abs(ep - (-170.724)) < 1e-3  # survey measurement of the exact set
    # perfect lattice: forces vanish by symmetry
    assert np.abs(conf.a).max() < 1e-14


def test_suttonchen_lattice_energy_closed_form():
    x, L = fcc(4, 3.61)
    conf = O.Conf(x, L, z=29)
    p = dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.0)
    cm = O.CellManager("minimumimage", 7.0)
    ep, vir, S = O.update_forces(conf, cm, [O.Pot("suttonchen", **p)])
    a = 3.61
    pair = 0.5 * lattice_sum(a, 7.0, lambda r: p["e"] * (p["a"] / r) ** p["n"])
    rho = lattice_sum(a, 7.0, lambda r: (p["a"] / r) ** p["m"])
    per_atom = pair - p["c"] * p["e"] * np.sqrt(rho)
    assert abs(ep - 256 * per_atom) < 1e-10 * abs(ep)


def jittered(nx, a, sigma, seed, z):
    x, L = fcc(nx, a)
    rng = np.random.default_rng(seed)
    return np.mod(x + rng.normal(0, sigma, x.shape), L), L


POT_CASES = [
    ("lennardjones", LJ, 18, 5.26, 8.5),
    ("fastlj", dict(LJ, bins=500), 18, 5.26, 8.5),
    ("morse", dict(depth=0.3429, a=1.3588, re=2.866), 29, 3.61, 7.0),
    ("buckingham", dict(B1=1254.5, B2=2.0, Ro=0.3249, cutoff=7.0), 11, 3.61, 7.0),
    ("suttonchen", dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.0), 29, 3.61, 7.0),
    ("finnissinclair", dict(c0=1.2371147, c1=-0.3592185, c2=-0.0385607, A=1.828905, B=1.8, c=3.4, d=3.569745), 26, 2.87 * 1.26, 4.0),
    ("gupta", dict(A=0.0855, r0=2.556, p=10.960, B=1.2240, q=2.278, cutoff=7.0), 29, 3.61, 7.0),
    ("harmonic", dict(k=3.4, a=3.0), 29, 3.61, 5.0),
    ("lennardjonesMod", dict(LJ, m=10, n=5, cohesive=0.8), 18, 5.26, 8.5),
    ("finnissinclair-ext", dict(c0=47.98066, c1=-34.09924, c2=5.832293, c3=0.017494, c4=0.020393, A=1.848648, B=1.2,
                                c=3.2572, d=4.1472), 42, 4.0, 4.2),
    ("tabulatedpair", dict(file=os.path.join(GOLDEN, "table_morse_cu.dat")), 29, 3.61, 7.0),
]


def energy_of(x, L, z, pot, rc):
    conf = O.Conf(x, L, z=z, mass=1.0)
    cm = O.CellManager("minimumimage", rc)
    ep, vir, S = O.update_forces(conf, cm, [pot])
    return ep, vir, S, conf


@pytest.mark.parametrize("name,params,z,a,rc", POT_CASES)
def test_forces_are_energy_gradients(name, params, z, a, rc):
    x, L = jittered(3, a, 0.05, 7, z)
    pot = O.Pot(name, **params)
    ep, vir, S, conf = energy_of(x, L, z, pot, rc)
    F = conf.a / O.FORCEFACTOR  # unit masses
    # Newton's third law
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * max(1.0, np.abs(F).max())
    if name in ("fastlj", "tabulatedpair"):
        return  # piecewise-constant tables: no derivative to compare
    h = 1e-6
    for i in (0, 17, 50):
        for q in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[i, q] += h
            xm[i, q] -= h
            dE = (energy_of(xp, L, z, pot, rc)[0] - energy_of(xm, L, z, pot, rc)[0]) / (2 * h)
            assert abs(-dE - F[i, q]) < 2e-5 * max(1e-3, np.abs(F).max()), (i, q, -dE, F[i, q])
    # virial: W = -dE/d(ln lambda) under a uniform dilation of positions and box
    lam = 1e-6
    Ep = energy_of(x * (1 + lam), L * (1 + lam), z, pot, rc)[0]
    Em = energy_of(x * (1 - lam), L * (1 - lam), z, pot, rc)[0]
    assert abs(vir + (Ep - Em) / (2 * lam)) < 1e-4 * max(1e-3, abs(vir))
    # stress diagonal: S[a][a] = -dE/d(eps_aa) for a dilation along one axis
    for q in range(3):
        sc = np.ones(3)
        sc[q] = 1 + lam
        Ep = energy_of(x * sc, L * sc, z, pot, rc)[0]
        sc[q] = 1 - lam
        Em = energy_of(x * sc, L * sc, z, pot, rc)[0]
        assert abs(S[q, q] + (Ep - Em) / (2 * lam)) < 1e-4 * max(1e-3, abs(S[q, q]))
    assert abs(np.trace(S) - vir) < 1e-10 * max(1.0, abs(vir))


def test_two_atom_stress_closed_form():
    # non-periodic pair along (1,1,0)/sqrt2 at r: S_ab = -f(r) rij_a rij_b
    r = 3.9
    d = np.array([1.0, 1.0, 0.0]) / np.sqrt(2) * r
    x = np.array([[50.0, 50.0, 50.0], 50.0 + d])
    pot = O.Pot("lennardjones", **LJ)
    conf = O.Conf(x, [100.0] * 3, per=(0, 0, 0), mass=1.0)
    ep, vir, S = O.update_forces(conf, O.CellManager("minimumimage", 8.5), [pot])
    s, e = LJ["sigma"], LJ["epsilon"]
    f = -48 * e / r**2 * ((s / r) ** 12 - 0.5 * (s / r) ** 6)  # force on atom 0 along rij
    assert abs(ep - 4 * e * ((s / r) ** 12 - (s / r) ** 6)) < 1e-15
    np.testing.assert_allclose(S, -f * np.outer(d, d), rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(conf.a[0] / O.FORCEFACTOR, f * d, rtol=1e-12)


def test_species_rule_sum_aliasing():
    # AppliesTo matches s1+s2 (potential.cc:23-27): (1,3) and (2,2) alias
    rng = np.random.default_rng(9)
    L = np.array([30.0] * 3)
    x = rng.uniform(0, 30, (40, 3))
    z = rng.choice([1, 2, 3], 40).astype(np.int32)
    pot = O.Pot("lennardjones", 2, 2, **dict(LJ, sigma=2.0))
    conf = O.Conf(x, L, z=z, mass=1.0)
    ep, _, _ = O.update_forces(conf, O.CellManager("minimumimage", 8.5), [pot])
    # brute force over all pairs with the minimum image
    ref = 0.0
    for i in range(40):
        for j in range(i + 1, 40):
            if z[i] + z[j] != 4:
                continue
            d = O.displacement(L, x[i], x[j])
            r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            if r2 < 8.5 * 8.5:
                ref += pot.pair_energy(np.sqrt(r2))[0]
    assert abs(ep - ref) < 1e-12 * max(1.0, abs(ref))


def test_md_energy_error_scales_with_dt_squared():
    # a 13-atom Ar cluster whose pairs all stay inside the cutoff, so that the
    # truncated potential has no jumps: velocity Verlet's energy error ~ dt^2
    s0 = 3.41 * 2 ** (1 / 6)
    pts = [[0, 0, 0]] + [[a, b, 0] for a in (-1, 1) for b in (-1, 1)] + \
          [[a, 0, b] for a in (-1, 1) for b in (-1, 1)] + [[0, a, b] for a in (-1, 1) for b in (-1, 1)]
    x = 200.0 + np.array(pts, dtype=np.float64) * s0 / np.sqrt(2)
    rng = np.random.default_rng(4)
    v = rng.normal(0, 2e-3, x.shape)
    v -= v.mean(axis=0)
    errs = []
    for dt, steps in ((8.0, 100), (4.0, 200)):
        conf = O.Conf(x, [400.0] * 3, v=v)
        cm = O.CellManager("minimumimage", 40.0)
        ep, ek, _ = O.md_run(conf, cm, [O.Pot("lennardjones", **dict(LJ, cutoff=40.0))], dt, steps)
        et = ep + ek
        errs.append(np.abs(et - et[0]).max())
    assert 3.0 < errs[0] / errs[1] < 5.0  # 4 for a second-order integrator

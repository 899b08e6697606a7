"""GPU parity of the HIP engine against the FP64 oracle and the golden fixtures,
through the C ABI.

Bars: neighbour lists bit-exact (indices, order, r2); forces within 1e-5
relative RMS of FP64 (‖F - F64‖2 / ‖F64‖2, the tolerance the workload
states), potential energy within 1e-6 relative; integrator bit-exact when the
forces are exactly zero.
"""
import json
import os
import sys

import numpy as np
import pytest

import lpmd2_amd as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "golden.json")))
FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def jittered(nx, a, sigma, seed):
    x, L = E.fcc(nx, a)
    return np.mod(x + np.random.default_rng(seed).normal(0.0, sigma, x.shape), L), L


def rel_rms(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_cm(cm, cm_args):
    """The oracle's cell manager for a plugin name and keyword string."""
    args = dict(zip(cm_args.split()[0::2], cm_args.split()[1::2]))
    cut = float(args.get("cutoff", 0.0))
    if cm in ("linkedcell", "lcbinary"):
        return O.CellManager(cm, cut, int(args.get("nx", 0)), int(args.get("ny", 0)), int(args.get("nz", 0)),
                             mode=args.get("mode", "noauto"))
    if cm == "minimumimage":
        return O.CellManager("minimumimage", cut)
    return O.CellManager("verletlist", cut, maxneighbors=int(args["maxneighbors"]), each=int(args["each"]))


# --------------------------------------------------------------- neighbour lists
@pytest.mark.parametrize("case", sorted(META["lists"]))
def test_lists_match_golden(case):
    m = META["lists"][case]
    say("lists", case)
    d = np.load(os.path.join(GOLDEN, f"lists_{case}.npz"))
    x, L = d["x"], d["L"]
    s = E.System(x, L, 18)
    s.set_cellmanager(m["cm"], m["cm_args"])
    oc = O.Conf(x, L)
    om = oracle_cm(m["cm"], m["cm_args"])
    om.update(oc)
    for full, tag in ((True, "full"), (False, "half")):
        counts, js, r2 = s.neighbor_lists(full, m["rcut"])
        assert np.array_equal(counts, d[f"{tag}_counts"])
        oc_counts, ojs, or2, _ = om.lists(oc, full, m["rcut"])
        assert np.array_equal(js, ojs)
        assert np.array_equal(r2, or2)
        if f"{tag}_j" in d:
            assert np.array_equal(js, d[f"{tag}_j"])


# --------------------------------------------------------------- forces
# Mo finnissinclair-ext (the plugin's help example) with B = 1.2
FSX = dict(c0=47.98066, c1=-34.09924, c2=5.832293, c3=0.017494, c4=0.020393, A=1.848648, B=1.2, c=3.2572,
           d=4.1472)
FSX_ARGS = " ".join(f"{k} {v}" for k, v in FSX.items())
# a Cu Morse pair table (r E ff rows) for tabulatedpair, a fixture in tests/golden/
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "table_morse_cu.dat")
POTS = [
    ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5", dict(sigma=3.41, epsilon=0.0103, cutoff=8.5), 18, 5.26, 0.1),
    ("fastlj", "sigma 3.41 epsilon 0.0103 cutoff 8.5", dict(sigma=3.41, epsilon=0.0103, cutoff=8.5, bins=500), 18, 5.26, 0.1),
    ("morse", "depth 0.3429 a 1.3588 re 2.866", dict(depth=0.3429, a=1.3588, re=2.866), 29, 3.61, 0.05),
    ("buckingham", "B1 1254.5 B2 2.0 Ro 0.3249 cutoff 7.0", dict(B1=1254.5, B2=2.0, Ro=0.3249, cutoff=7.0), 11, 3.61, 0.05),
    ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22",
     dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22), 29, 3.61, 0.05),
    ("finnissinclair", "c0 1.2371147 c1 -0.3592185 c2 -0.0385607 A 1.828905 B 1.8 c 3.4 d 3.569745",
     dict(c0=1.2371147, c1=-0.3592185, c2=-0.0385607, A=1.828905, B=1.8, c=3.4, d=3.569745), 26, 2.87 * 1.26, 0.05),
    ("gupta", "A 0.0855 r0 2.556 p 10.960 B 1.2240 q 2.278 cutoff 7.22",
     dict(A=0.0855, r0=2.556, p=10.960, B=1.2240, q=2.278, cutoff=7.22), 29, 3.61, 0.05),
    ("harmonic", "k 3.4 a 3.0", dict(k=3.4, a=3.0), 29, 3.61, 0.2),
    ("lennardjonesMod", "sigma 3.41 epsilon 0.0103 m 10 n 5 cohesive 0.8 cutoff 8.5",
     dict(sigma=3.41, epsilon=0.0103, m=10, n=5, cohesive=0.8, cutoff=8.5), 18, 5.26, 0.1),
    ("finnissinclair-ext", FSX_ARGS, FSX, 42, 4.0, 0.05),
    ("tabulatedpair", f"file {TABLE}", dict(file=TABLE), 29, 3.61, 0.05),
]


@pytest.mark.parametrize("cm", ["linkedcell", "verletlist", "minimumimage", "lcbinary"])
@pytest.mark.parametrize("name,args,kw,z,a,jit", POTS, ids=[p[0] for p in POTS])
def test_forces_match_oracle(name, args, kw, z, a, jit, cm):
    say("forces", name, cm)
    x, L = jittered(5, a, jit, 17)
    pot = O.Pot(name, **kw)
    rc = pot.cutoff if np.isfinite(pot.cutoff) else 7.0
    cm_args = {"linkedcell": f"cutoff {rc} mode auto", "verletlist": f"cutoff {rc} maxneighbors 400 each 5",
               "minimumimage": f"cutoff {rc}", "lcbinary": f"cutoff {rc} mode auto"}[cm]
    om = oracle_cm(cm, cm_args)
    s = E.System(x, L, z)
    s.set_cellmanager(cm, cm_args)
    s.append_potential(name, args)
    e, w, S = s.update_forces()
    _, _, acc = s.get_state()
    conf = O.Conf(x, L, z=z)
    e0, w0, S0 = O.update_forces(conf, om, [pot])
    err = rel_rms(acc, conf.a)
    say(f"  rel force rms {err:.2e}  E {e:.10g} vs {e0:.10g}  W {w:.6g} vs {w0:.6g}")
    assert err <= FORCE_TOL
    assert abs(e - e0) <= ENERGY_TOL * abs(e0)
    assert abs(w - w0) <= 1e-5 * max(abs(w0), 1e-3)
    assert np.abs(S - S0).max() <= 1e-5 * max(np.abs(S0).max(), 1e-3)


def test_species_pairs_buckingham_rocksalt():
    """Three species-pair potentials (Na-Na, Cl-Cl, Na-Cl) with SetValidSpecies."""
    say("species rocksalt")
    x, L = jittered(4, 5.64, 0.05, 3)
    z = np.where(np.arange(x.shape[0]) % 2 == 0, 11, 17).astype(np.int32)
    pots = [("buckingham", "B1 424.1 B2 1.05 Ro 0.3170 cutoff 7.0", dict(B1=424.1, B2=1.05, Ro=0.3170, cutoff=7.0), 11, 11),
            ("buckingham", "B1 1256.3 B2 2.0 Ro 0.3321 cutoff 7.0", dict(B1=1256.3, B2=2.0, Ro=0.3321, cutoff=7.0), 17, 17),
            ("buckingham", "B1 1254.5 B2 0.0 Ro 0.3249 cutoff 7.0", dict(B1=1254.5, B2=0.0, Ro=0.3249, cutoff=7.0), 11, 17)]
    s = E.System(x, L, z)
    s.set_cellmanager("verletlist", "cutoff 7.0 maxneighbors 300 each 5")
    op = []
    for name, args, kw, s1, s2 in pots:
        s.append_potential(name, args, s1, s2)
        op.append(O.Pot(name, s1, s2, **kw))
    e, w, S = s.update_forces()
    _, _, acc = s.get_state()
    conf = O.Conf(x, L, z=z)
    e0, w0, S0 = O.update_forces(conf, O.CellManager("verletlist", 7.0, maxneighbors=300, each=5), op)
    err = rel_rms(acc, conf.a)
    say(f"  rel force rms {err:.2e}  E {e} vs {e0}")
    assert err <= FORCE_TOL
    assert abs(e - e0) <= ENERGY_TOL * abs(e0)


# --------------------------------------------------------------- integrator
def test_free_flight_is_bit_exact():
    """epsilon = 0 makes every force exactly zero: the trajectory is the
    reference's AdvancePosition/FittedInside arithmetic alone."""
    say("free flight")
    x, L = jittered(3, 5.26, 0.3, 5)
    v = np.random.default_rng(2).normal(0, 0.05, x.shape)
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 1.0")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0 cutoff 8.5")
    s.velocityverlet(2.0, 37)
    xg, vg, ag = s.get_state()
    conf = O.Conf(x, L, v=v)
    for _ in range(37):
        conf.a[:] = 0.0
        aold = O.vv_position(conf, 2.0)
        conf.a[:] = 0.0
        O.vv_velocity(conf, 2.0, aold)
    assert np.array_equal(xg, conf.x)
    assert np.array_equal(vg, conf.v)


@pytest.mark.parametrize("cm", ["linkedcell", "verletlist"])
def test_short_trajectory_matches_oracle(cm):
    """20 steps of LJ argon at 100 K: per-atom positions stay within FP32
    force noise of the FP64 trajectory."""
    say("trajectory", cm)
    x, L = E.fcc(4, 5.26)
    m = np.full(x.shape[0], 39.948)
    v = E.maxwell_boltzmann(m, 100.0, seed=7)
    s = E.System(x, L, 18, v=v)
    if cm == "linkedcell":
        s.set_cellmanager("linkedcell", "cutoff 8.5 mode auto")
        om = O.CellManager("linkedcell", 8.5, mode="auto")
    else:
        s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 200 each 5")
        om = O.CellManager("verletlist", 8.5, maxneighbors=200, each=5)
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    ep, ek = s.velocityverlet(1.0, 20, obs_every=5)
    xg, vg, _ = s.get_state()
    conf = O.Conf(x, L, v=v)
    e0, k0, _ = O.md_run(conf, om, [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)], 1.0, 20)
    dx = xg - conf.x
    dx -= L * np.round(dx / L)
    say(f"  max |dx| {np.abs(dx).max():.2e}  E {ep} vs {e0[4::5]}")
    assert np.abs(dx).max() < 1e-8
    np.testing.assert_allclose(ep, e0[4::5], rtol=1e-6)
    np.testing.assert_allclose(ek, k0[4::5], rtol=1e-6)


@pytest.mark.parametrize("cm", ["linkedcell", "verletlist"])
def test_leapfrog_trajectory_matches_oracle(cm):
    """Leapfrog set-up (forces at the old configuration, one atom crossing
    the box edge there) and 20 + 10 steps of LJ argon at 100 K in two calls:
    positions within the FP32 force noise the trajectory accumulates
    (1e-7), sampled energies to 1e-6 relative."""
    say("leapfrog", cm)
    x, L = E.fcc(4, 5.26)
    m = np.full(x.shape[0], 39.948)
    v = E.maxwell_boltzmann(m, 100.0, seed=8)
    x[0] = [0.002, 0.003, 0.004]
    v[0] = [0.003, 0.004, 0.005]

    def cms():
        if cm == "linkedcell":
            return "cutoff 8.5 mode auto", O.CellManager("linkedcell", 8.5, mode="auto")
        return "cutoff 8.5 maxneighbors 200 each 5 skin 1.0", O.CellManager("verletlist", 8.5, maxneighbors=200, each=5)

    args, om = cms()
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager(cm, args)
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    s.leapfrog_init(1.0)
    ep, ek = s.leapfrog(1.0, 20, obs_every=5)
    xg, vg, ag = s.get_state()
    pots = [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)]
    conf = O.Conf(x, L, v=v)
    e0, k0 = O.leapfrog_run(conf, lambda c: O.update_forces(c, om, pots)[0], 1.0, 20)
    dx = xg - conf.x
    dx -= L * np.round(dx / L)
    say(f"  max |dx| {np.abs(dx).max():.2e} |dv| {np.abs(vg - conf.v).max():.2e}  E {ep} vs {e0[4::5]}")
    # FP32 force noise (1e-7 relative) over 20 steps, grown by the fast atom's
    # collisions; a dropped half kick would move atoms by ~1e-3
    assert np.abs(dx).max() < 1e-7
    assert np.abs(vg - conf.v).max() < 2e-8
    assert rel_rms(ag, conf.a) <= FORCE_TOL
    np.testing.assert_allclose(ep, e0[4::5], rtol=1e-6)
    np.testing.assert_allclose(ek, k0[4::5], rtol=1e-6)
    # a second call continues from the half-step velocities the first left
    ep2, ek2 = s.leapfrog(1.0, 10, obs_every=10)
    _, om30 = cms()
    e30, k30 = O.leapfrog_run(O.Conf(x, L, v=v), lambda c: O.update_forces(c, om30, pots)[0], 1.0, 30)
    np.testing.assert_allclose(ep2, e30[29:], rtol=1e-6)
    np.testing.assert_allclose(ek2, k30[29:], rtol=1e-6)


def test_leapfrog_free_flight_is_bit_exact():
    """Zero forces: the set-up's GoBack/FittedInside and every step's
    position and velocity arithmetic, bit for bit."""
    say("leapfrog free flight")
    x, L = jittered(3, 5.26, 0.3, 6)
    v = np.random.default_rng(3).normal(0, 0.05, x.shape)
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 1.0")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0 cutoff 8.5")
    s.leapfrog_init(2.0)
    s.leapfrog(2.0, 37)
    xg, vg, _ = s.get_state()
    conf = O.Conf(x, L, v=v)
    zero = np.zeros_like(x)

    def forces(c):
        c.a[:] = zero
        return 0.0

    O.leapfrog_run(conf, forces, 2.0, 37)
    assert np.array_equal(xg, conf.x)
    assert np.array_equal(vg, conf.v)


def test_leapfrog_errors():
    say("leapfrog errors")
    x, L = jittered(3, 5.26, 0.05, 3)
    s = E.System(x, L, 18)
    s.set_cellmanager("linkedcell", "cutoff 8.5 mode auto")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    with pytest.raises(E.LpmdxError) as e:
        s.leapfrog(2.0, 3)
    assert e.value.code == E.ESTATE
    s.leapfrog_init(2.0)
    with pytest.raises(E.LpmdxError) as e:
        s.leapfrog(1.0, 3)
    assert e.value.code == E.EINVAL
    s.leapfrog(2.0, 3)


@pytest.mark.parametrize("cm", ["linkedcell", "verletlist"])
def test_beeman_trajectory_matches_oracle(cm):
    """Beeman set-up (forces at the old configuration, one atom crossing the
    box edge there) and 20 + 10 steps of LJ argon at 100 K in two calls:
    positions within the FP32 force noise the trajectory accumulates (1e-7),
    sampled energies to 1e-6 relative."""
    say("beeman", cm)
    x, L = E.fcc(4, 5.26)
    m = np.full(x.shape[0], 39.948)
    v = E.maxwell_boltzmann(m, 100.0, seed=10)
    x[0] = [0.002, 0.003, 0.004]
    v[0] = [0.003, 0.004, 0.005]

    def cms():
        if cm == "linkedcell":
            return "cutoff 8.5 mode auto", O.CellManager("linkedcell", 8.5, mode="auto")
        return "cutoff 8.5 maxneighbors 200 each 5 skin 1.0", O.CellManager("verletlist", 8.5, maxneighbors=200, each=5)

    args, om = cms()
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager(cm, args)
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    s.beeman_init(1.0)
    ep, ek = s.beeman(1.0, 20, obs_every=5)
    xg, vg, ag = s.get_state()
    pots = [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)]
    conf = O.Conf(x, L, v=v)
    e0, k0 = O.beeman_run(conf, lambda c: O.update_forces(c, om, pots)[0], 1.0, 20)
    dx = xg - conf.x
    dx -= L * np.round(dx / L)
    say(f"  max |dx| {np.abs(dx).max():.2e} |dv| {np.abs(vg - conf.v).max():.2e}  E {ep} vs {e0[4::5]}")
    assert np.abs(dx).max() < 1e-7
    assert np.abs(vg - conf.v).max() < 2e-8
    assert rel_rms(ag, conf.a) <= FORCE_TOL
    np.testing.assert_allclose(ep, e0[4::5], rtol=1e-6)
    np.testing.assert_allclose(ek, k0[4::5], rtol=1e-6)
    ep2, ek2 = s.beeman(1.0, 10, obs_every=10)
    _, om30 = cms()
    e30, k30 = O.beeman_run(O.Conf(x, L, v=v), lambda c: O.update_forces(c, om30, pots)[0], 1.0, 30)
    np.testing.assert_allclose(ep2, e30[29:], rtol=1e-6)
    np.testing.assert_allclose(ek2, k30[29:], rtol=1e-6)


@pytest.mark.parametrize("cm", ["linkedcell", "verletlist"])
def test_nosehoover_trajectory_matches_oracle(cm):
    """The nosehoover integrator (plugins/nosehoover.cc:62-96): set-up on the
    old configuration, then 20 + 10 steps of LJ argon started at 100 K and
    driven toward 300 K (Q 1e7), in two calls, one fixedvel atom and two
    fixedpos atoms (which stay put and add no friction increment) among them:
    positions within the FP32 force noise the trajectory accumulates,
    velocities and the sampled energies against the oracle's sequential
    restatement; the friction has visibly heated the box.
    (The reference's friction increment per atom is (3 N / Q) k_B (T - t) dt,
    so a gentle thermostat for 256 atoms needs Q near 1e7.)"""
    say("nosehoover", cm)
    x, L = E.fcc(4, 5.26)
    m = np.full(x.shape[0], 39.948)
    v = E.maxwell_boltzmann(m, 100.0, seed=14)
    x[0] = [0.002, 0.003, 0.004]
    v[0] = [0.003, 0.004, 0.005]
    fv = np.zeros(x.shape[0], bool)
    fv[7] = True
    fp = np.zeros(x.shape[0], bool)
    fp[[20, 131]] = True
    Q, T0, dt = 1e7, 300.0, 2.0

    def cms():
        if cm == "linkedcell":
            return "cutoff 8.5 mode auto", O.CellManager("linkedcell", 8.5, mode="auto")
        return "cutoff 8.5 maxneighbors 200 each 5 skin 0.5", O.CellManager("minimumimage", 8.5)

    args, om = cms()
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager(cm, args)
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    s.set_fixed(fixedvel=fv, fixedpos=fp)
    s.nosehoover_init(dt, Q, T0)
    ep, ek = s.nosehoover(dt, 20, obs_every=5)
    xg, vg, ag = s.get_state()
    pots = [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)]
    conf = O.Conf(x, L, v=v)
    conf.set_fixed(fixedvel=fv, fixedpos=fp)
    e0, k0 = O.nosehoover_run(conf, lambda c: O.update_forces(c, om, pots)[0], dt, 20, Q, T0)
    dx = xg - conf.x
    dx -= L * np.round(dx / L)
    say(f"  max |dx| {np.abs(dx).max():.2e} |dv| {np.abs(vg - conf.v).max():.2e}  Ek {ek} vs {k0[4::5]}")
    assert np.abs(dx).max() < 1e-7
    assert np.abs(vg - conf.v).max() < 2e-8
    assert rel_rms(ag, conf.a) <= FORCE_TOL
    np.testing.assert_allclose(ep, e0[4::5], rtol=1e-6)
    np.testing.assert_allclose(ek, k0[4::5], rtol=1e-6)
    assert ek[-1] > 1.5 * ek[0]  # the thermostat pulls toward 300 K
    assert np.array_equal(xg[fp], x[fp]) and np.array_equal(vg[fp], v[fp])
    ep2, ek2 = s.nosehoover(dt, 10, obs_every=10)
    _, om30 = cms()
    c30 = O.Conf(x, L, v=v)
    c30.set_fixed(fixedvel=fv, fixedpos=fp)
    e30, k30 = O.nosehoover_run(c30, lambda c: O.update_forces(c, om30, pots)[0], dt, 30, Q, T0)
    np.testing.assert_allclose(ep2, e30[29:], rtol=1e-6)
    np.testing.assert_allclose(ek2, k30[29:], rtol=1e-6)
    if cm == "verletlist":
        # the skin test saw the Nose-Hoover position updates: the list was rebuilt
        assert s.stats()["builds"] >= 2
    with pytest.raises(E.LpmdxError) as e:
        s.nosehoover(1.0, 1)
    assert e.value.code == E.EINVAL


def test_beeman_free_flight_is_bit_exact():
    """Zero forces: the set-up and every step's Beeman position and velocity
    arithmetic, bit for bit."""
    say("beeman free flight")
    x, L = jittered(3, 5.26, 0.3, 7)
    v = np.random.default_rng(4).normal(0, 0.05, x.shape)
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 1.0")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0 cutoff 8.5")
    s.beeman_init(2.0)
    s.beeman(2.0, 37)
    xg, vg, _ = s.get_state()
    conf = O.Conf(x, L, v=v)
    zero = np.zeros_like(x)

    def forces(c):
        c.a[:] = zero
        return 0.0

    O.beeman_run(conf, forces, 2.0, 37)
    assert np.array_equal(xg, conf.x)
    assert np.array_equal(vg, conf.v)


def test_beeman_errors_and_scheme_switch():
    """Beeman without its set-up or with another dt is an error; switching
    between velocityverlet and beeman on one system continues each scheme
    from the state the other left (the look-ahead positions are taken back)."""
    say("beeman errors / switch")
    x, L = jittered(3, 5.26, 0.05, 3)
    s = E.System(x, L, 18)
    s.set_cellmanager("linkedcell", "cutoff 8.5 mode auto")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    with pytest.raises(E.LpmdxError) as e:
        s.beeman(2.0, 3)
    assert e.value.code == E.ESTATE
    s.beeman_init(2.0)
    with pytest.raises(E.LpmdxError) as e:
        s.beeman(1.0, 3)
    assert e.value.code == E.EINVAL
    # 5 velocity-Verlet steps, then Beeman from there, against the oracle
    s.velocityverlet(2.0, 5)
    x5, v5, a5 = s.get_state()
    s.beeman(2.0, 5)
    xg, vg, _ = s.get_state()
    om = O.CellManager("linkedcell", 8.5, mode="auto")
    pots = [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)]
    conf = O.Conf(x, L)
    O.md_run(conf, om, pots, 2.0, 5)
    # the Beeman history set at beeman_init (forces at x - v dt with v = 0: x itself)
    hist = O.Conf(x, L)
    O.update_forces(hist, om, pots)
    calls = iter([hist.a] + [None] * 10)

    def forces(c):
        h = next(calls)
        if h is not None:
            c.a[:] = h
            return 0.0
        return O.update_forces(c, om, pots)[0]

    b = O.Conf(conf.x, L, v=conf.v)
    O.beeman_run(b, forces, 2.0, 5)
    dx = xg - b.x
    dx -= L * np.round(dx / L)
    say(f"  after the switch max |dx| {np.abs(dx).max():.2e}")
    # FP32 force noise over 10 steps; taking velocity Verlet's look-ahead
    # instead of Beeman's would be off by (a_n - a_n-1) dt^2 / 6 ~ 1e-5
    assert np.abs(dx).max() < 5e-8


def test_set_temperature_matches_oracle():
    """SetTemperature on the device against the oracle's sequential sums: the
    same factor to FP64 rounding, and the temperature read back is the target."""
    say("set temperature")
    x, L = jittered(5, 3.61, 0.02, 21)
    m = np.full(x.shape[0], 63.546)
    v = E.maxwell_boltzmann(m, 300.0, seed=21)
    s = E.System(x, L, 29, v=v)
    s.set_temperature(512.0)
    _, vg, _ = s.get_state()
    conf = O.Conf(x, L, z=29, v=v)
    O.set_temperature(conf, 512.0)
    np.testing.assert_allclose(vg, conf.v, rtol=1e-13, atol=0)
    assert abs(O.temperature(O.Conf(x, L, z=29, v=vg)) - 512.0) <= 1e-10 * 512.0
    # the first call drew the velocities (InitVelocities): the given ones are gone
    assert np.abs(vg - v).max() > 1e-4
    # with the velocities marked as set, SetTemperature only scales them
    s1 = E.System(x, L, 29, v=v)
    s1.velocities_set(True)
    s1.set_temperature(512.0)
    c1 = O.Conf(x, L, z=29, v=v)
    c1.velocities_set = True
    O.set_temperature(c1, 512.0)
    np.testing.assert_allclose(s1.get_state()[1], c1.v, rtol=1e-13, atol=0)
    # a system at rest whose velocities count as set stays at rest
    s0 = E.System(x, L, 29)
    s0.velocities_set(True)
    s0.set_temperature(300.0)
    assert np.all(s0.get_state()[1] == 0.0)


def test_init_velocities_match_oracle():
    """InitVelocities (simulationbuilder.cc:167-185) from a seeded drand48
    stream: drawn on the host in the reference's order, so bit-identical to
    the oracle's restatement; a second draw continues the stream."""
    say("init velocities")
    x, L = jittered(4, 3.61, 0.02, 23)
    z = np.where(np.arange(x.shape[0]) % 4 == 0, 28, 29)
    s = E.System(x, L, z)
    conf = O.Conf(x, L, z=z)
    for seed in (1, 987654321):
        s.seed48(seed)
        O.seed48(conf, seed)
        for _ in range(2):
            s.init_velocities()
            O.init_velocities(conf)
            np.testing.assert_array_equal(s.get_state()[1], conf.v)


@pytest.mark.parametrize("mod", ["tempscaling", "berendsen"])
@pytest.mark.parametrize("cm", ["linkedcell", "verletlist"])
@pytest.mark.parametrize("scheme", ["velocityverlet", "beeman", "leapfrog", "nosehoover"])
def test_tempscaling_schedule_matches_oracle(scheme, cm, mod):
    """A temperature ramp 60 K -> 140 K applied every 4 steps over 40 steps of
    LJ argon, inside the device step loops, against the oracle running the
    same modifier after each active step (lpmd/lpmd.cc:89-93).  With a
    0.1 A skin the list is rebuilt inside the run, so speculative steps past
    the skin limit (thermostat kernels included) are launched and skipped."""
    say("tempscaling", scheme, cm)
    x, L = jittered(4, 5.26, 0.05, 22)
    m = np.full(x.shape[0], 39.948)
    v = E.maxwell_boltzmann(m, 100.0, seed=22)
    dt, steps = 2.0, 40
    s = E.System(x, L, 18, v=v)
    if cm == "linkedcell":
        s.set_cellmanager("linkedcell", "cutoff 8.5 mode auto")
    else:
        s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 0.1")
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    if mod == "tempscaling":
        s.tempscaling(60.0, 140.0, start=0, end=steps, each=4)
        st = O.Stepper(0, steps, 4)

        def modifier(k, c):
            if st.active(k):
                O.set_temperature(c, st.value(k, 60.0, 140.0))
    else:
        # berendsen with tau 16 fs from step 3 every 12: armed at 3 (stop 11),
        # scaling at 4..11, disarmed at 12 (each back to 12), so armed again
        # at 15, 27 and 39: scaling at 4..11, 16..23, 28..35 and 40
        s.berendsen(60.0, 140.0, tau=16.0, start=3, end=steps, each=12)
        ob = O.Berendsen(60.0, 140.0, tau=16.0, start=3, end=steps, each=12)

        def modifier(k, c):
            ob.apply(k, c, dt)

    # linkedcell: the reference's stencil (two cells a side here, so cells are
    # visited twice) on both sides; the skin list holds the exact
    # minimum-image set, which the oracle's minimumimage rebuilds every call
    om = (O.CellManager("linkedcell", 8.5, mode="auto") if cm == "linkedcell"
          else O.CellManager("minimumimage", 8.5))
    pots = [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)]
    conf = O.Conf(x, L, v=v)
    forces = lambda c: O.update_forces(c, om, pots)[0]
    if scheme == "velocityverlet":
        ep, ek = s.velocityverlet(dt, steps, obs_every=10)
        e0 = np.zeros(steps)
        k0 = np.zeros(steps)
        for k in range(1, steps + 1):
            forces(conf)
            aold = O.vv_position(conf, dt)
            e0[k - 1] = forces(conf)
            O.vv_velocity(conf, dt, aold)
            modifier(k, conf)
            k0[k - 1] = O.kinetic_energy(conf)
    elif scheme == "beeman":
        s.beeman_init(dt)
        ep, ek = s.beeman(dt, steps, obs_every=10)
        e0, k0 = O.beeman_run(conf, forces, dt, steps, after_step=modifier)
    elif scheme == "leapfrog":
        s.leapfrog_init(dt)
        ep, ek = s.leapfrog(dt, steps, obs_every=10)
        e0, k0 = O.leapfrog_run(conf, forces, dt, steps, after_step=modifier)
    else:
        # a Nose-Hoover bath at 100 K under the ramp: the friction of each
        # step reads the temperature the modifier left
        s.nosehoover_init(dt, 1e7, 100.0)
        ep, ek = s.nosehoover(dt, steps, obs_every=10)
        e0, k0 = O.nosehoover_run(conf, forces, dt, steps, 1e7, 100.0, after_step=modifier)
    xg, vg, _ = s.get_state()
    dx = xg - conf.x
    dx -= L * np.round(dx / L)
    say(f"  max |dx| {np.abs(dx).max():.2e} |dv| {np.abs(vg - conf.v).max():.2e}  Ek {ek} vs {k0[9::10]}")
    assert np.abs(dx).max() < 1e-7
    assert np.abs(vg - conf.v).max() < 2e-8
    np.testing.assert_allclose(ep, e0[9::10], rtol=1e-6)
    np.testing.assert_allclose(ek, k0[9::10], rtol=1e-6)
    if cm == "verletlist":
        st = s.stats()
        say(f"  list builds {st['builds']}, chunks cut short {st['cut_chunks']}")
        assert st["builds"] > 2
        # chunks ran past the skin limit: the modifier state was replayed
        assert st["cut_chunks"] >= 1
    # at the last step tempscaling set the ramp's end temperature (index 1:
    # LPMDX_OBS_TEMP)
    if mod == "tempscaling":
        assert abs(s.observables()[1] - 140.0) <= 1e-9 * 140.0
    with pytest.raises(E.LpmdxError) as e:
        s.tempscaling(60.0, 140.0, start=0, end=-1, each=4)
    assert e.value.code == E.EINVAL


def test_nve_energy_matches_oracle_suttonchen():
    """400 steps of Sutton-Chen Cu at 300 K, dt 2 fs, skin-triggered list:
    the sampled total energy follows the FP64 oracle (which runs the
    reference's verletlist cadence) and fluctuates by the same amount."""
    say("nve sc")
    x, L = jittered(5, 3.61, 0.01, 9)
    m = np.full(x.shape[0], 63.546)
    v = E.maxwell_boltzmann(m, 300.0, seed=9)
    s = E.System(x, L, 29, v=v)
    s.set_cellmanager("verletlist", "cutoff 7.22 maxneighbors 400 each 5 skin 0.3")
    s.append_potential("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22")
    ep, ek = s.velocityverlet(2.0, 400, obs_every=20)
    et = ep + ek
    conf = O.Conf(x, L, z=29, v=v)
    e0, k0, _ = O.md_run(conf, O.CellManager("verletlist", 7.22, maxneighbors=400, each=5),
                         [O.Pot("suttonchen", e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22)], 2.0, 400)
    et0 = (e0 + k0)[19::20]
    say(f"  E_tot spread gpu {et.max() - et.min():.4e} oracle {et0.max() - et0.min():.4e}, builds {s.stats()['builds']}")
    np.testing.assert_allclose(et, et0, rtol=1e-5)
    assert abs((et.max() - et.min()) - (et0.max() - et0.min())) < 0.2 * (et0.max() - et0.min())


def test_event_profile_samples_steps_and_leaves_results_alone():
    """Per-kernel HIP-event timing: every launch of every k-th step is
    bracketed (the EAM rho and force passes share the event between them),
    the counts follow the sampling period, and the trajectory is the same bit
    for bit with the events on or off."""
    say("event profile")
    x, L = jittered(5, 3.61, 0.01, 4)
    m = np.full(x.shape[0], 63.546)
    v = E.maxwell_boltzmann(m, 300.0, seed=4)
    out = []
    for every in (0, 1, 4):
        s = E.System(x, L, 29, v=v)
        s.set_cellmanager("verletlist", "cutoff 7.22 maxneighbors 400 each 5 skin 0.3")
        s.append_potential("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22")
        s.velocityverlet(2.0, 3)
        s.profile(every > 0, every=max(every, 1))
        s.velocityverlet(2.0, 16)
        p = s.profile_read()
        out.append(s.get_state())
        if every:
            assert p["force"][1] == 16 // every and p["rho"][1] == 16 // every, p
            assert p["force"][0] > 0.0 and p["rho"][0] > 0.0
        else:
            assert p["force"][1] == 0 and p["rho"][1] == 0
    with pytest.raises(E.LpmdxError) as e:
        E._check(E.lib().lpmdx_system_profile(s._h, -1))
    assert e.value.code == E.EINVAL
    for st in out[1:]:
        for a, b in zip(out[0], st):
            np.testing.assert_array_equal(a, b)


# --------------------------------------------------------------- error behaviour
def test_error_paths():
    say("errors")
    x, L = jittered(3, 3.61, 0.01, 1)
    s = E.System(x, L, 29)
    with pytest.raises(E.LpmdxError) as e:
        s.update_forces()
    assert e.value.code == E.ESTATE
    with pytest.raises(E.LpmdxError) as e:
        s.set_cellmanager("verletlist", "cutoff 7.22 maxneighbors 100 each 5")
    assert e.value.code == E.ECAPACITY and "maxneighbors" in str(e.value)
    with pytest.raises(E.LpmdxError) as e:
        s.set_cellmanager("linkedcell", "cutoff 7.22 mode")
    assert e.value.code == E.EINVAL
    with pytest.raises(E.LpmdxError) as e:
        s.set_state(x=np.full_like(x, np.nan))
    assert e.value.code == E.ENONFINITE
    with pytest.raises(E.LpmdxError) as e:
        s.append_potential("nosuchpotential", "")
    assert e.value.code == E.EINVAL


# --------------------------------------------------------------- fixed atoms, tags, per-atom energies
def fixed_masks(n, seed):
    """fixedvel on every 7th atom, fixedpos "true" on every 11th, a fixedpos
    tag reading "false" on every 13th (tagged, still integrated)."""
    k = np.arange(n)
    fv = (k % 7) == 3
    fp = ((k % 11) == 5) | ((k % 13) == 6)
    ft = (k % 11) == 5
    return fv, fp, ft


@pytest.mark.parametrize("cm", ["linkedcell", "verletlist"])
@pytest.mark.parametrize("scheme", ["velocityverlet", "beeman", "leapfrog"])
def test_fixed_atoms_match_oracle(scheme, cm):
    """The fixedvel / fixedpos tags inside the device step loops (a = 0 for
    fixedvel atoms after each force evaluation of a step; fixedpos "true"
    atoms neither move nor change velocity) against the oracle's
    TwoStepIntegrator / OneStepIntegrator restatements; with velocity Verlet
    a tempscaling ramp runs too, which with tags draws no velocities for and
    scales none of the tagged atoms and counts only the untagged ones."""
    say("fixed atoms", scheme, cm)
    x, L = jittered(4, 5.26, 0.05, 31)
    n = x.shape[0]
    m = np.full(n, 39.948)
    v = E.maxwell_boltzmann(m, 100.0, seed=31)
    dt, steps = 2.0, 30
    fv, fp, ft = fixed_masks(n, 31)
    s = E.System(x, L, 18, v=v)
    if cm == "linkedcell":
        s.set_cellmanager("linkedcell", "cutoff 8.5 mode auto")
        om = O.CellManager("linkedcell", 8.5, mode="auto")
    else:
        s.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 300 each 5 skin 0.1")
        om = O.CellManager("minimumimage", 8.5)
    s.append_potential("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    s.set_fixed(fixedvel=fv, fixedpos=fp, fixedpos_true=ft)
    s.velocities_set(True)
    pots = [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103, cutoff=8.5)]
    conf = O.Conf(x, L, v=v)
    conf.set_fixed(fixedvel=fv, fixedpos=fp, fixedpos_true=ft)
    conf.velocities_set = True
    forces = lambda c: O.update_forces(c, om, pots)[0]
    if scheme == "velocityverlet":
        s.tempscaling(80.0, 120.0, start=0, end=steps, each=5)
        st = O.Stepper(0, steps, 5)
        ep, ek = s.velocityverlet(dt, steps, obs_every=10)
        e0 = np.zeros(steps)
        k0 = np.zeros(steps)
        hold = (conf.fix & O.FIX_POS) != 0
        for k in range(1, steps + 1):
            forces(conf)
            conf.a[fv] = 0.0
            x_keep, v_keep = conf.x[hold].copy(), conf.v[hold].copy()
            aold = O.vv_position(conf, dt)
            conf.x[hold] = x_keep
            e0[k - 1] = forces(conf)
            conf.a[fv] = 0.0
            O.vv_velocity(conf, dt, aold)
            conf.v[hold] = v_keep
            if st.active(k):
                O.set_temperature(conf, st.value(k, 80.0, 120.0))
            k0[k - 1] = O.kinetic_energy_tag(conf, True)
    elif scheme == "beeman":
        s.beeman_init(dt)
        ep, ek = s.beeman(dt, steps, obs_every=10)
        e0, k0 = O.beeman_run(conf, forces, dt, steps)
    else:
        s.leapfrog_init(dt)
        ep, ek = s.leapfrog(dt, steps, obs_every=10)
        e0, k0 = O.leapfrog_run(conf, forces, dt, steps)
    xg, vg, ag = s.get_state()
    dx = xg - conf.x
    dx -= L * np.round(dx / L)
    say(f"  max |dx| {np.abs(dx).max():.2e} |dv| {np.abs(vg - conf.v).max():.2e}")
    hold = ft
    np.testing.assert_array_equal(xg[hold], x[hold])
    np.testing.assert_array_equal(vg[hold], v[hold])
    assert np.all(ag[fv] == 0.0)
    assert np.abs(dx).max() < 1e-7
    assert np.abs(vg - conf.v).max() < 2e-8
    np.testing.assert_allclose(ep, e0[9::10], rtol=1e-6)
    np.testing.assert_allclose(ek, k0[9::10], rtol=1e-6)
    if scheme == "velocityverlet":
        # the tag-aware temperature after the last step is the ramp's end
        assert abs(s.tags()["temperature"] - 120.0) <= 1e-9 * 120.0


@pytest.mark.parametrize("tagged", [False, True])
def test_observables_match_oracle(tagged):
    """SetSimulationTags on the device (lpmdx_observables) against the
    oracle's restatement after one force evaluation: kinetic energy,
    temperature and momentum (tag-aware when atoms carry fixedvel or
    fixedpos), potential energy, virial, pressure and its parts, the stress
    tags, total energy, |p|, volume, densities and cell edges."""
    say("observables", tagged)
    x, L = jittered(5, 3.61, 0.02, 41)
    L = L * np.array([1.0, 1.0, 1.0])
    n = x.shape[0]
    m = np.full(n, 63.546)
    v = E.maxwell_boltzmann(m, 450.0, seed=41) + 1e-3  # a net momentum
    sc = "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22"
    s = E.System(x, L, 29, v=v)
    s.set_cellmanager("verletlist", "cutoff 7.22 maxneighbors 300 each 5")
    s.append_potential("suttonchen", sc)
    conf = O.Conf(x, L, z=29, v=v)
    if tagged:
        fv, fp, ft = fixed_masks(n, 41)
        s.set_fixed(fixedvel=fv, fixedpos=fp, fixedpos_true=ft)
        conf.set_fixed(fixedvel=fv, fixedpos=fp, fixedpos_true=ft)
    eg, wg, Sg = s.update_forces()
    out = s.observables()
    e0, w0, S0 = O.update_forces(conf, O.CellManager("verletlist", 7.22, maxneighbors=300, each=5),
                                 [O.Pot("suttonchen", e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22)])
    # the energies, virial and stress of the FP32 pair arithmetic against FP64:
    # the virial and stress are sums of pair terms that largely cancel near
    # the lattice constant, so their bar is set by the energy scale
    esc = abs(e0)
    assert abs(eg - e0) <= 1e-6 * esc
    assert abs(wg - w0) <= 1e-5 * max(abs(w0), esc)
    assert np.abs(Sg - S0).max() <= 1e-5 * max(np.abs(S0).max(), esc)
    # every tag is the SetSimulationTags formula applied to this state and the
    # engine's own energies (sums in another order: 1e-12)
    ref = O.tags(conf, eg, wg, Sg)
    exact = [19, 20, 21, 22, 25, 26, 27]  # from the box and the masses
    np.testing.assert_array_equal(out[exact], ref[exact])
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-18)
    if tagged:
        untagged = conf.fix & O.FIX_TAGGED == 0
        assert abs(out[1] - O.temperature_tag(conf, True)) <= 1e-12 * out[1]
        assert untagged.sum() < n


@pytest.mark.parametrize("case", ["lennardjones", "suttonchen", "rocksalt"])
def test_atom_energies_match_oracle(case):
    """Potential::AtomEnergy on the device (lpmdx_atom_energies) against the
    oracle's restatement: pair energies over each atom's full list, plus
    F(rho_i) for EAM; with species rules for the rocksalt Buckingham terms."""
    from lpmd2_amd import workloads as W

    say("atom energies", case)
    if case == "lennardjones":
        x, L = jittered(4, 5.26, 0.05, 51)
        z = np.full(x.shape[0], 18)
        pots = [("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5", -1, -1)]
        cm = ("linkedcell", "cutoff 8.5 mode auto")
    elif case == "suttonchen":
        x, L = jittered(5, 3.61, 0.02, 52)
        z = np.full(x.shape[0], 29)
        pots = [("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22", -1, -1)]
        cm = ("verletlist", "cutoff 7.22 maxneighbors 300 each 5")
    else:
        x, L, z = W.rocksalt(3, 3, 3, 5.64)
        x = np.mod(x + np.random.default_rng(53).normal(0, 0.03, x.shape), L)
        pots = [("buckingham", a, s1, s2) for a, s1, s2 in W.BUCKINGHAM_NACL]
        cm = ("verletlist", "cutoff 8.0 maxneighbors 300 each 5")
    s = E.System(x, L, z)
    s.set_cellmanager(*cm)
    for p in pots:
        s.append_potential(*p)
    conf = O.Conf(x, L, z=z)
    ocm = oracle_cm(*cm)
    for k, (name, args, s1, s2) in enumerate(pots):
        kw = dict(zip(args.split()[0::2], (float(q) for q in args.split()[1::2])))
        pot = O.Pot(name, s1, s2, **kw)
        eg = s.atom_energies(k)
        e0 = O.atom_energies(conf, ocm, pot)
        err = rel_rms(eg, e0)
        say(f"  term {k}: rel rms {err:.2e}")
        assert err <= 1e-6

"""GPU checks of the drop-in plugins driven by the reference's own
composition: Configuration::SetCellManager, CombinedPotential::UpdateForces
(combinedpotential.h:39-52), CellManager::BuildNeighborList and
Integrator::Advance, with the MI355X plugins loaded through the reference's
create(args).  Bars: neighbour lists bit-exact against the golden fixtures;
forces within 1e-5 relative RMS of the FP64 oracle; the reference's own
velocityverlet (TwoStepIntegrator on the host) over the drop-in potentials,
and the drop-in velocityverlet, follow the engine's native trajectory."""
import json
import os
import sys

import numpy as np
import pytest

import lpmd2_amd as E
from oracle import oracle as O
from oracle import refcheck as R

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not R.dropin_available(), reason="drop-in plugins not built")]

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "golden.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rms(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", ["ar2048j_lc_auto", "cu256j_lc_dup", "rand300_vl", "rand300_mi_wide", "rand300_lcb",
                                  "cu256j_lcb"])
def test_dropin_lists_match_golden(case):
    m = META["lists"][case]
    say("drop-in lists", case)
    d = np.load(os.path.join(GOLDEN, f"lists_{case}.npz"))
    conf = R.RefConf(d["x"], d["L"], z=18)
    cm = R.RefPlugin.dropin(m["cm"], m["cm_args"])
    cm.attach(conf)
    for full, tag in ((True, "full"), (False, "half")):
        counts, js, r2, rij = cm.lists(conf, full, m["rcut"])
        assert np.array_equal(counts, d[f"{tag}_counts"])
        if f"{tag}_j" in d:
            assert np.array_equal(js, d[f"{tag}_j"])
        if f"{tag}_r2" in d:
            assert np.array_equal(r2, d[f"{tag}_r2"])


CASES = [
    ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5", dict(sigma=3.41, epsilon=0.0103, cutoff=8.5), 18, 5.26),
    ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22",
     dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22), 29, 3.61),
    ("gupta", "A 0.0855 r0 2.556 p 10.960 B 1.2240 q 2.278 cutoff 7.22",
     dict(A=0.0855, r0=2.556, p=10.960, B=1.2240, q=2.278, cutoff=7.22), 29, 3.61),
]


CM_ARGS = {"linkedcell": "cutoff {rc} mode auto", "verletlist": "cutoff {rc} maxneighbors 400 each 5",
           "minimumimage": "cutoff {rc}", "lcbinary": "cutoff {rc} mode auto"}


def oracle_cm(cmname, rc):
    if cmname in ("linkedcell", "lcbinary"):
        return O.CellManager(cmname, rc, mode="auto")
    if cmname == "minimumimage":
        return O.CellManager("minimumimage", rc)
    return O.CellManager("verletlist", rc, maxneighbors=400, each=5)


@pytest.mark.parametrize("cmname", ["linkedcell", "verletlist", "minimumimage", "lcbinary"])
@pytest.mark.parametrize("name,args,kw,z,a", CASES, ids=[c[0] for c in CASES])
def test_dropin_combined_forces_match_oracle(name, args, kw, z, a, cmname):
    say("drop-in forces", name, cmname)
    x, L = E.fcc(5, a)
    x = np.mod(x + np.random.default_rng(5).normal(0.0, 0.05, x.shape), L)
    rc = kw["cutoff"]
    cm_args = CM_ARGS[cmname].format(rc=rc)
    conf = R.RefConf(x, L, z=z)
    cm = R.RefPlugin.dropin(cmname, cm_args)
    cm.attach(conf)
    pot = R.RefPlugin.dropin(name, args)
    acc, e, w, S = R.combined_forces(conf, [pot])
    oc = O.Conf(x, L, z=z)
    om = oracle_cm(cmname, rc)
    e0, w0, S0 = O.update_forces(oc, om, [O.Pot(name, **kw)])
    err = rel_rms(acc, oc.a)
    say(f"  rel force rms {err:.2e}  E {e:.10g} vs {e0:.10g}")
    assert err <= 1e-5
    assert abs(e - e0) <= 1e-6 * abs(e0)
    assert abs(w - w0) <= 1e-5 * abs(w0)
    assert np.abs(S - np.asarray(S0).reshape(3, 3)).max() <= 1e-5 * np.abs(S0).max()


def test_dropin_species_pairs_rocksalt():
    say("drop-in rocksalt")
    x, L = E.fcc(4, 5.64)
    x = np.mod(x + np.random.default_rng(3).normal(0.0, 0.05, x.shape), L)
    z = np.where(np.arange(x.shape[0]) % 2 == 0, 11, 17).astype(np.int32)
    terms = [("B1 424.1 B2 1.05 Ro 0.3170 cutoff 7.0", dict(B1=424.1, B2=1.05, Ro=0.3170, cutoff=7.0), (11, 11)),
             ("B1 1256.3 B2 2.0 Ro 0.3321 cutoff 7.0", dict(B1=1256.3, B2=2.0, Ro=0.3321, cutoff=7.0), (17, 17)),
             ("B1 1254.5 B2 0.0 Ro 0.3249 cutoff 7.0", dict(B1=1254.5, B2=0.0, Ro=0.3249, cutoff=7.0), (11, 17))]
    conf = R.RefConf(x, L, z=z)
    cm = R.RefPlugin.dropin("verletlist", "cutoff 7.0 maxneighbors 400 each 5")
    cm.attach(conf)
    pots = [R.RefPlugin.dropin("buckingham", t[0]) for t in terms]
    acc, e, w, S = R.combined_forces(conf, pots, species=[t[2] for t in terms])
    oc = O.Conf(x, L, z=z)
    om = O.CellManager("verletlist", 7.0, maxneighbors=400, each=5)
    e0, w0, _ = O.update_forces(oc, om, [O.Pot("buckingham", s1=t[2][0], s2=t[2][1], **t[1]) for t in terms])
    err = rel_rms(acc, oc.a)
    say(f"  rel force rms {err:.2e}  E {e:.10g} vs {e0:.10g}")
    assert err <= 1e-5
    assert abs(e - e0) <= 1e-6 * abs(e0)


def native_run(x, L, z, v, cm, cm_args, pot, dt, steps, integrator="velocityverlet"):
    s = E.System(x, L, z, v=v)
    s.set_cellmanager(cm, cm_args)
    s.append_potential(*pot)
    if integrator == "leapfrog":
        s.leapfrog_init(dt)
        s.leapfrog(dt, steps)
    elif integrator == "beeman":
        s.beeman_init(dt)
        s.beeman(dt, steps)
    elif integrator == "nosehoover":
        s.nosehoover_init(dt, NH_Q, NH_T)
        s.nosehoover(dt, steps)
    else:
        s.velocityverlet(dt, steps)
    return s.get_state()


# the nosehoover keywords of the drop-in runs below (fmass, t)
NH_Q, NH_T = 1e7, 300.0


def integrator_args(scheme, dt):
    if scheme == "nosehoover":
        return f"dt {dt} fmass {NH_Q} t {NH_T}"
    return f"dt {dt}"


@pytest.mark.parametrize("integrator", ["reference", "dropin"])
def test_dropin_integration_follows_native_path(integrator):
    say("drop-in integration", integrator)
    x, L = E.fcc(5, 3.61)
    x = np.mod(x + np.random.default_rng(8).normal(0.0, 0.02, x.shape), L)
    m = np.full(x.shape[0], E.ELEMENT_MASS[29])
    v = E.maxwell_boltzmann(m, 600.0, seed=4)
    pot = ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22")
    cm_args = "cutoff 7.22 maxneighbors 400 skin 0.3"
    steps, dt = 30, 2.0
    conf = R.RefConf(x, L, z=29)
    conf.set_velocities(v)
    cm = R.RefPlugin.dropin("verletlist", cm_args)
    cm.attach(conf)
    pots = [R.RefPlugin.dropin(*pot)]
    if integrator == "reference":
        itg = R.RefPlugin.reference("velocityverlet", f"dt {dt}")
    else:
        itg = R.RefPlugin.dropin("velocityverlet", f"dt {dt}")
    ep = R.integrate(itg, conf, pots, steps)
    xr, vr, ar = conf.get()
    xn, vn, an = native_run(x, L, 29, v, "verletlist", cm_args, pot, dt, steps)
    dx = np.abs(xr - xn).max()
    dv = np.abs(vr - vn).max()
    say(f"  max |dx| {dx:.3e}  |dv| {dv:.3e}")
    assert dx <= 1e-10 and dv <= 1e-12
    if integrator == "dropin":
        # the drop-in integrator sets the potential-energy tag every step
        assert np.all(np.isfinite(ep)) and ep[-1] != 0.0


@pytest.mark.parametrize("scheme,integrator", [("leapfrog", "reference"), ("leapfrog", "dropin"),
                                                ("beeman", "reference"), ("beeman", "dropin"),
                                                ("nosehoover", "dropin")])
def test_dropin_oldconfig_integrators_follow_native_path(integrator, scheme):
    """The reference leapfrog / beeman plugin over drop-in forces (its set-up
    force call on the old configuration included) and the drop-in plugin of
    the same name both follow the engine's own scheme.  (The reference
    nosehoover reads the session's kboltzmann, which needs a runtime context
    this harness cannot build: only the drop-in runs it here.)"""
    say("drop-in", scheme, integrator)
    x, L = E.fcc(5, 3.61)
    x = np.mod(x + np.random.default_rng(9).normal(0.0, 0.02, x.shape), L)
    m = np.full(x.shape[0], E.ELEMENT_MASS[29])
    v = E.maxwell_boltzmann(m, 600.0, seed=5)
    pot = ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22")
    cm_args = "cutoff 7.22 maxneighbors 400 skin 0.3"
    steps, dt = 30, 2.0
    conf = R.RefConf(x, L, z=29)
    conf.set_velocities(v)
    cm = R.RefPlugin.dropin("verletlist", cm_args)
    cm.attach(conf)
    pots = [R.RefPlugin.dropin(*pot)]
    make = R.RefPlugin.reference if integrator == "reference" else R.RefPlugin.dropin
    itg = make(scheme, integrator_args(scheme, dt))
    ep = R.integrate(itg, conf, pots, steps)
    xr, vr, ar = conf.get()
    xn, vn, an = native_run(x, L, 29, v, "verletlist", cm_args, pot, dt, steps, integrator=scheme)
    dx = np.abs(xr - xn).max()
    dv = np.abs(vr - vn).max()
    say(f"  max |dx| {dx:.3e}  |dv| {dv:.3e}")
    assert dx <= 1e-10 and dv <= 1e-12
    if integrator == "dropin":
        assert np.all(np.isfinite(ep)) and ep[-1] != 0.0


@pytest.mark.parametrize("name,args,kw,z,a", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_dropin_atom_energy_matches_oracle(name, args, kw, z, a):
    """Potential::AtomEnergy through the drop-in (PairPotential::AtomEnergy,
    pairpotential.cc:21-33; MetalPotential::AtomEnergy, metalpotential.cc:47-60)
    for every atom, computed on the device, against the oracle's restatement;
    moving the atoms refreshes them."""
    say("drop-in atom energy", name)
    x, L = E.fcc(5, a)
    x = np.mod(x + np.random.default_rng(12).normal(0.0, 0.03, x.shape), L)
    rc = kw["cutoff"]
    conf = R.RefConf(x, L, z=z)
    cm = R.RefPlugin.dropin("verletlist", f"cutoff {rc} maxneighbors 400 each 5")
    cm.attach(conf)
    pot = R.RefPlugin.dropin(name, args)
    oc = O.Conf(x, L, z=z)
    ocm = O.CellManager("verletlist", rc, maxneighbors=400, each=5)
    err = rel_rms(R.atom_energies(pot, conf), O.atom_energies(oc, ocm, O.Pot(name, **kw)))
    say(f"  rel rms {err:.2e}")
    assert err <= 1e-6
    x2 = np.mod(x + 0.05, L)
    conf.set_positions(x2)
    oc2 = O.Conf(x2, L, z=z)
    err2 = rel_rms(R.atom_energies(pot, conf), O.atom_energies(oc2, ocm, O.Pot(name, **kw)))
    assert err2 <= 1e-6


@pytest.mark.parametrize("scheme", ["velocityverlet", "leapfrog", "beeman", "nosehoover"])
def test_dropin_integrators_honour_fixed_atoms(scheme):
    """Atoms tagged fixedpos ("true") and fixedvel through the reference's
    BasicParticleSet::SetTag: the drop-in integrator's device steps keep the
    fixedpos atoms where they are with their velocities and zero the fixedvel
    atoms' accelerations, following the oracle's TwoStepIntegrator /
    OneStepIntegrator restatement (velocity Verlet) and the engine's own
    tagged run (all four schemes)."""
    say("drop-in fixed atoms", scheme)
    x, L = E.fcc(4, 5.26)
    x = np.mod(x + np.random.default_rng(13).normal(0.0, 0.05, x.shape), L)
    n = x.shape[0]
    v = E.maxwell_boltzmann(np.full(n, E.ELEMENT_MASS[18]), 100.0, seed=13)
    fv = (np.arange(n) % 9) == 2
    fp = (np.arange(n) % 10) == 5
    pot = ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
    cm_args = "cutoff 8.5 maxneighbors 400 each 5 skin 0.5"
    steps, dt = 25, 2.0
    conf = R.RefConf(x, L, z=18)
    conf.set_velocities(v)
    for i in np.nonzero(fv)[0]:
        R.set_atom_tag(conf, i, "fixedvel", "true")
    for i in np.nonzero(fp)[0]:
        R.set_atom_tag(conf, i, "fixedpos", "true")
    cm = R.RefPlugin.dropin("verletlist", cm_args)
    cm.attach(conf)
    pots = [R.RefPlugin.dropin(*pot)]
    itg = R.RefPlugin.dropin(scheme, integrator_args(scheme, dt))
    R.integrate(itg, conf, pots, steps)
    xr, vr, ar = conf.get()
    np.testing.assert_array_equal(xr[fp], x[fp])
    np.testing.assert_array_equal(vr[fp], v[fp])
    assert np.all(ar[fv] == 0.0)
    s = E.System(x, L, 18, v=v)
    s.set_cellmanager("verletlist", cm_args)
    s.append_potential(*pot)
    s.set_fixed(fixedvel=fv, fixedpos=fp)
    if scheme == "leapfrog":
        s.leapfrog_init(dt)
        s.leapfrog(dt, steps)
    elif scheme == "beeman":
        s.beeman_init(dt)
        s.beeman(dt, steps)
    elif scheme == "nosehoover":
        s.nosehoover_init(dt, NH_Q, NH_T)
        s.nosehoover(dt, steps)
    else:
        s.velocityverlet(dt, steps)
    xn, vn, an = s.get_state()
    assert np.abs(xr - xn).max() <= 1e-10 and np.abs(vr - vn).max() <= 1e-12
    if scheme == "velocityverlet":
        oc = O.Conf(x, L, z=18, v=v)
        oc.set_fixed(fixedvel=fv, fixedpos=fp)
        O.md_run(oc, O.CellManager("minimumimage", 8.5), [O.Pot("lennardjones", sigma=3.41, epsilon=0.0103,
                                                                   cutoff=8.5)], dt, steps)
        dx = xr - oc.x
        dx -= L * np.round(dx / L)
        say(f"  vs oracle max |dx| {np.abs(dx).max():.2e} |dv| {np.abs(vr - oc.v).max():.2e}")
        assert np.abs(dx).max() < 1e-7 and np.abs(vr - oc.v).max() < 2e-8

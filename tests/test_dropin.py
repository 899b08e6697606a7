"""CPU checks of the drop-in plugins (lpmd2_amd/plugins): every plugin object
loads through the reference's create(args) entry point (liblpmd/lpmd/plugin.h:42-46),
parses its keywords like the reference plugin of the same name, reports the
same cutoff, and fails loudly when no GPU is visible (no CPU fallback).
Skipped where the plugins are not built (they need the reference headers)."""
import os

import numpy as np
import pytest

import lpmd2_amd as E
from oracle import oracle as O
from oracle import refcheck as R

pytestmark = pytest.mark.skipif(not R.dropin_available(), reason="drop-in plugins not built")

POTS = [
    ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5"),
    ("fastlj", "sigma 3.41 epsilon 0.0103 cutoff 8.5"),
    ("morse", "depth 0.3429 a 1.3588 re 2.866"),
    ("buckingham", "B1 1254.5 B2 2.0 Ro 0.3249 cutoff 7.0"),
    ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22"),
    ("finnissinclair", "c0 1.2371147 c1 -0.3592185 c2 -0.0385607 A 1.828905 B 1.8 c 3.4 d 3.569745"),
    ("gupta", "A 0.0855 r0 2.556 p 10.960 B 1.2240 q 2.278 cutoff 7.22"),
    ("harmonic", "k 3.4 a 3.0"),
    ("lennardjonesMod", "sigma 3.41 epsilon 0.0103 m 10 n 5 cohesive 0.8 cutoff 8.5"),
    ("lennardjonesMod", "sigma 3.41 epsilon 0.0103 cutoff 8.5"),
    ("finnissinclair-ext", "c0 47.98066 c1 -34.09924 c2 5.832293 c3 0.017494 c4 0.020393 A 1.848648 B 1.2 "
                           "c 3.2572 d 4.1472"),
    ("tabulatedpair", "file " + os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                             "table_morse_cu.dat")),
]


METALS = ("suttonchen", "finnissinclair", "gupta", "finnissinclair-ext")


def create_in_child(name, args):
    """create(args) of the drop-in `name` in a child process through the
    reference's loader: (exit status, stderr)."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r); from oracle import refcheck as R; "
            "R.RefPlugin.dropin(%r, %r); print('created')" % (ROOT, name, args))
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stderr


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name,args", POTS, ids=[f"{p[0]}-{k}" for k, p in enumerate(POTS)])
def test_potential_plugins_match_reference_keywords(name, args):
    ours = R.RefPlugin.dropin(name, args)
    ref = R.RefPlugin.reference(name, args)
    # Potential::GetCutoff through the reference's interface
    assert R.lib().ref_pot_cutoff(ours.h) == R.lib().ref_pot_cutoff(ref.h)
    ours.close()


@pytest.mark.parametrize("name,args", POTS[:2] + POTS[4:5],
                         ids=[p[0] for p in POTS[:2] + POTS[4:5]])
def test_bad_arguments_end_the_program_without_unwinding_out_of_create(name, args):
    """A dangling keyword: the reference's create() throws (which the
    reference's own loader cannot survive, the plugin being dlclosed while the
    exception is in flight); the drop-in's create() prints the same message,
    as lpmd::EndWithError does, and exits with status 1."""
    with pytest.raises(RuntimeError) as e:
        R.RefPlugin.reference(name, args + " cutoff")
    rc, err = create_in_child(name, args + " cutoff")
    assert rc == 1, err
    assert "[Error] " + str(e.value) in err
    rc, err = create_in_child(name, args)
    assert rc == 0, err


@pytest.mark.parametrize("name,args", POTS, ids=[f"{p[0]}-{k}" for k, p in enumerate(POTS)])
def test_per_pair_virtuals_match_reference(name, args):
    """The drop-ins are lpmd::PairPotential / lpmd::MetalPotential objects:
    through dynamic_cast (as plugins/localpressure.cc:128-129 uses them) their
    pairEnergy / pairForce, or pairEnergy / rhoij / F / PairForce / ManyBodies,
    give the reference plugin's values bit for bit on 300 random pairs."""
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    ours = R.RefPot(name, args, path=os.path.join(R.DROPIN_DIR, name + ".so"))
    ref = R.RefPot(name, args)
    rc = ref.cutoff if np.isfinite(ref.cutoff) else 7.0
    d = rng.normal(size=(300, 3))
    rij = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.6 * rc, 0.999 * rc, 300)[:, None]
    if name in METALS:
        rhoi, rhoj = rng.uniform(0.5, 40.0, 300), rng.uniform(0.5, 40.0, 300)
        for a, b in zip(ours.metal(rij, rhoi, rhoj), ref.metal(rij, rhoi, rhoj)):
            np.testing.assert_array_equal(a, b)
    else:
        for a, b in zip(ours.pair(rij), ref.pair(rij)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name,args", [("linkedcell", "cutoff 8.5 mode auto"),
                                       ("verletlist", "cutoff 8.5 maxneighbors 200 each 5"),
                                       ("verletlist", "cutoff 8.5 maxneighbors 200 skin 1.0"),
                                       ("minimumimage", "cutoff 8.5"),
                                       ("lcbinary", "cutoff 8.5 mode auto"),
                                       ("velocityverlet", "dt 2.0"),
                                       ("leapfrog", "dt 2.0"),
                                       ("beeman", "dt 2.0"),
                                       ("nosehoover", "dt 2.0 fmass 1e7 t 300.0")])
def test_other_plugins_load(name, args):
    p = R.RefPlugin.dropin(name, args)
    p.close()


@pytest.mark.parametrize("name,fixture,calls_per_step", [("leapfrog", "lf_ar108.npz", 1),
                                                         ("beeman", "bm_ar108.npz", 2)])
def test_integrator_host_path_matches_reference_golden(name, fixture, calls_per_step):
    """With a CPU cell manager the drop-in integrator runs the reference's
    scheme on the host: replayed forces give the reference plugin's
    trajectory bit for bit (no GPU is touched)."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", fixture))
    acc = d["acc"]
    conf = R.RefConf(d["x0"], d["L"])
    conf.set_velocities(d["v0"])
    cm = R.RefCellManager("linkedcell", "cutoff 8.5 nx 4 ny 4 nz 4")
    R.vv_run("dt 2.0", conf, cm, (acc.shape[0] - 1) // calls_per_step, acc.reshape(acc.shape[0], -1),
             path=os.path.join(R.DROPIN_DIR, name + ".so"))
    xf, vf, af = conf.get()
    assert np.array_equal(xf, d["xf"])
    assert np.array_equal(vf, d["vf"])
    assert np.array_equal(af, d["af"])


@pytest.mark.parametrize("held", [False, True])
def test_nosehoover_host_path_matches_oracle(held):
    """The drop-in nosehoover with a CPU cell manager runs NoseHoover over
    TwoStepIntegrator on the host (nosehoover.cc:62-96,
    twostepintegrator.cc:12-68): with the forces of the golden Beeman run
    replayed (a box at 2.4e8 K, where atoms cross the cell each step) and a
    thermostat strong enough to cool it by a fifth (Q 1e12, bath 30 K), the
    trajectory is the oracle's sequential restatement bit for bit, with or
    without fixedpos atoms (which add no friction increment)."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bm_ar108.npz"))
    acc = d["acc"]
    steps = (acc.shape[0] - 1) // 2
    conf = R.RefConf(d["x0"], d["L"])
    conf.set_velocities(d["v0"])
    oc = O.Conf(d["x0"], d["L"], v=d["v0"])
    if held:
        fp = (np.arange(conf.n) % 7) == 3
        for i in np.nonzero(fp)[0]:
            R.set_atom_tag(conf, i, "fixedpos", "true")
        oc.set_fixed(fixedpos=fp)
    cm = R.RefCellManager("linkedcell", "cutoff 8.5 nx 4 ny 4 nz 4")
    R.vv_run("dt 2.0 fmass 1e12 t 30.0", conf, cm, steps, acc.reshape(acc.shape[0], -1),
             path=os.path.join(R.DROPIN_DIR, "nosehoover.so"))
    calls = iter(acc)

    def replay(c):
        c.a[:] = next(calls)
        return 0.0

    O.nosehoover_run(oc, replay, 2.0, steps, fmass=1e12, temp=30.0)
    xf, vf, af = conf.get()
    assert not np.array_equal(vf, d["vf"])  # the friction acted
    assert np.array_equal(xf, oc.x)
    assert np.array_equal(vf, oc.v)
    assert np.array_equal(af, oc.a)


def test_without_gpu_the_plugins_fail_loudly():
    if E.device_count() > 0:
        pytest.skip("a GPU is visible")
    x, L = E.fcc(3, 5.26)
    conf = R.RefConf(x, L, z=18)
    cm = R.RefPlugin.dropin("linkedcell", "cutoff 8.5 mode auto")
    with pytest.raises(RuntimeError, match="no HIP device"):
        cm.attach(conf)

"""CPU checks of the drop-in plugins (lpmd2_amd/plugins): every plugin object
loads through the reference's create(args) entry point (liblpmd/lpmd/plugin.h:42-46),
parses its keywords like the reference plugin of the same name, reports the
same cutoff, and fails loudly when no GPU is visible (no CPU fallback).
Skipped where the plugins are not built (they need the reference headers)."""
import os

import numpy as np
import pytest

import lpmd2_amd as E
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


@pytest.mark.parametrize("name,args", POTS, ids=[f"{p[0]}-{k}" for k, p in enumerate(POTS)])
def test_potential_plugins_match_reference_keywords(name, args):
    ours = R.RefPlugin.dropin(name, args)
    ref = R.RefPlugin.reference(name, args)
    # Potential::GetCutoff through the reference's interface
    assert R.lib().ref_pot_cutoff(ours.h) == R.lib().ref_pot_cutoff(ref.h)
    # a dangling keyword is the same error in both
    msgs = []
    for make in (R.RefPlugin.dropin, R.RefPlugin.reference):
        with pytest.raises(RuntimeError) as e:
            make(name, args + " cutoff")
        msgs.append(str(e.value))
    assert msgs[0] == msgs[1]
    ours.close()


@pytest.mark.parametrize("name,args", [("linkedcell", "cutoff 8.5 mode auto"),
                                       ("verletlist", "cutoff 8.5 maxneighbors 200 each 5"),
                                       ("verletlist", "cutoff 8.5 maxneighbors 200 skin 1.0"),
                                       ("minimumimage", "cutoff 8.5"),
                                       ("lcbinary", "cutoff 8.5 mode auto"),
                                       ("velocityverlet", "dt 2.0"),
                                       ("leapfrog", "dt 2.0"),
                                       ("beeman", "dt 2.0")])
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


def test_without_gpu_the_plugins_fail_loudly():
    if E.device_count() > 0:
        pytest.skip("a GPU is visible")
    x, L = E.fcc(3, 5.26)
    conf = R.RefConf(x, L, z=18)
    cm = R.RefPlugin.dropin("linkedcell", "cutoff 8.5 mode auto")
    with pytest.raises(RuntimeError, match="no HIP device"):
        cm.attach(conf)

"""GPU checks of the batched engine through the C ABI: forces against the FP64
oracle, results identical alone and in a mix, agreement with the plain
single-system path, and an autotuner that leaves trajectories alone."""
import sys

import numpy as np
import pytest

import lpmd2_amd as E
from lpmd2_amd.batch import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LJ = ("lennardjones", "sigma 3.41 epsilon 0.0103 cutoff 8.5")
SC = ("suttonchen", "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22")
GUPTA = ("gupta", "A 0.0855 r0 2.556 p 10.960 B 1.2240 q 2.278 cutoff 7.22")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def system(kind, nx, seed, T):
    if kind == "lj":
        x, L = E.fcc(nx, 5.26)
        z, pot, cm, dt = 18, LJ, "cutoff 8.5 skin 1.0", 2.0
    elif kind == "sc":
        x, L = E.fcc(nx, 3.61)
        z, pot, cm, dt = 29, SC, "cutoff 7.22 skin 0.3", 2.0
    else:
        x, L = E.fcc(nx, 3.61)
        z, pot, cm, dt = 29, GUPTA, "cutoff 7.22 skin 0.3", 2.0
    rng = np.random.default_rng(seed)
    x = np.mod(x + rng.normal(0, 0.02, x.shape), L)
    m = np.full(x.shape[0], E.ELEMENT_MASS[z])
    v = E.maxwell_boltzmann(m, T, seed=seed + 100)
    return dict(x=x, box=L, z=z, potentials=[pot], cm=cm, v=v, dt=dt)


def test_batch_forces_match_oracle():
    say("batch forces")
    specs = [system("lj", 4, 1, 100.0), system("sc", 5, 2, 300.0), system("gupta", 4, 3, 500.0)]
    b = Batch(atom_capacity=4096)
    sids = [b.add(steps=0, **s) for s in specs]
    b.run(1)  # the initial force evaluation
    for sid, s in zip(sids, specs):
        x, v, a, info = b.get(sid)
        assert info["started"] == 1 and info["overflow"] == 0
        name, args = s["potentials"][0]
        kw = dict(zip(args.split()[0::2], (float(q) for q in args.split()[1::2])))
        pot = O.Pot(name, **kw)
        conf = O.Conf(s["x"], s["box"], z=s["z"])
        cm = O.CellManager("verletlist", pot.cutoff, maxneighbors=400, each=5)
        e0, w0, _ = O.update_forces(conf, cm, [pot])
        err = np.linalg.norm(a - conf.a) / np.linalg.norm(conf.a)
        ep, vir, ek = b.energies(sid)
        say(f"  {name}: rel force rms {err:.2e}, E {ep:.9g} vs {e0:.9g}")
        assert err <= 1e-5
        assert abs(ep - e0) <= 1e-6 * abs(e0)
        np.testing.assert_array_equal(x, s["x"])


def test_alone_and_mixed_are_identical():
    say("alone vs mixed")
    a_spec = system("sc", 4, 7, 600.0)
    others = [system("lj", 3, 8, 200.0), system("gupta", 5, 9, 900.0), system("lj", 5, 10, 120.0)]
    b1 = Batch(atom_capacity=2048)
    s1 = b1.add(steps=60, **a_spec)
    b1.run(100)
    x1, v1, a1, i1 = b1.get(s1)
    b2 = Batch(atom_capacity=8192)
    b2.add(steps=80, **others[0])
    s2 = b2.add(steps=60, **a_spec)
    b2.add(steps=30, **others[1])
    b2.add(steps=90, **others[2])
    b2.run(100)
    x2, v2, a2, i2 = b2.get(s2)
    say(f"  builds alone {i1['builds']} mixed {i2['builds']}, done {i1['done']} {i2['done']}")
    assert i1["done"] == i2["done"] == 60
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(a1, a2)


def test_remove_mid_run_keeps_the_others():
    """A system removed between runs leaves the captured steps in place (no
    new plan): the systems still running end exactly as they do alone, and a
    system added afterwards into the freed slices runs as it does alone."""
    say("remove mid-run")
    # 2048 atoms stay, 256 leave: the gone tiles (4) are at most an eighth of
    # the live ones (32), so the removal keeps the captured plan
    a_spec = system("sc", 8, 31, 500.0)
    gone = system("gupta", 4, 32, 900.0)
    c_spec = system("lj", 4, 33, 150.0)
    alone = {}
    for name, spec in (("a", a_spec), ("c", c_spec)):
        b = Batch(atom_capacity=4096)
        s = b.add(steps=50, **spec)
        b.run(60)
        alone[name] = b.get(s)
    b = Batch(atom_capacity=4096)
    sa = b.add(steps=50, **a_spec)
    sg = b.add(steps=200, **gone)
    b.run(12)
    b.remove(sg)
    b.run(13)
    sc = b.add(steps=50, **c_spec)
    b.run(60)
    for name, sid in (("a", sa), ("c", sc)):
        x, v, a, info = b.get(sid)
        x0, v0, a0, i0 = alone[name]
        say(f"  {name}: done {info['done']} alone {i0['done']}, builds {info['builds']} alone {i0['builds']}")
        assert info["done"] == i0["done"] == 50
        np.testing.assert_array_equal(x, x0)
        np.testing.assert_array_equal(v, v0)
        np.testing.assert_array_equal(a, a0)


def test_batch_matches_plain_path():
    say("batch vs plain")
    s = system("lj", 4, 11, 100.0)
    b = Batch(atom_capacity=1024)
    sid = b.add(steps=40, **s)
    b.run(50)
    xb, vb, ab, info = b.get(sid)
    p = E.System(s["x"], s["box"], s["z"], v=s["v"])
    p.set_cellmanager("verletlist", "cutoff 8.5 maxneighbors 400 each 5 skin 1.0")
    p.append_potential(*s["potentials"][0])
    p.velocityverlet(s["dt"], 40)
    xp, vp, ap = p.get_state()
    dx = xb - xp
    dx -= s["box"] * np.round(dx / s["box"])
    say(f"  max |dx| {np.abs(dx).max():.2e}, builds {info['builds']}")
    assert np.abs(dx).max() < 1e-9
    np.testing.assert_allclose(vb, vp, rtol=0, atol=1e-11)


def test_autotune_keeps_state():
    say("autotune")
    specs = [system("lj", 4, 21, 100.0), system("sc", 4, 22, 300.0)]
    b = Batch(atom_capacity=4096)
    sids = [b.add(steps=30, **s) for s in specs]
    b.run(5)
    before = [b.get(s)[:2] for s in sids]
    lanes = b.autotune(trial_steps=5)
    after = [b.get(s)[:2] for s in sids]
    for (x0, v0), (x1, v1) in zip(before, after):
        np.testing.assert_array_equal(x0, x1)
        np.testing.assert_array_equal(v0, v1)
    say(f"  lanes {lanes}")
    assert lanes["lennardjones"] in (1, 2, 4) and lanes["suttonchen"] in (1, 2, 4)
    b.run(40)
    for s in sids:
        assert b.get(s)[3]["done"] == 30


def test_arena_admission_and_reuse():
    say("arena")
    b = Batch(atom_capacity=600, list_words=600 * 160)
    s1 = b.add(steps=5, **system("lj", 3, 31, 100.0))   # 108 atoms
    s2 = b.add(steps=5, **system("lj", 4, 32, 100.0))   # 256 atoms
    with pytest.raises(E.LpmdxError) as e:
        b.add(steps=5, **system("lj", 5, 33, 100.0))    # 500 atoms: no room
    assert e.value.code == E.ENOMEM
    b.remove(s2)
    s3 = b.add(steps=5, **system("lj", 4, 34, 100.0))
    b.run(10)
    assert b.get(s1)[3]["done"] == 5 and b.get(s3)[3]["done"] == 5


def oracle_pot(p):
    """The oracle's potential for a (name, args[, s1, s2]) tuple."""
    name, args = p[0], p[1]
    kw = dict(zip(args.split()[0::2], (float(q) for q in args.split()[1::2])))
    s1, s2 = (p[2], p[3]) if len(p) > 2 else (-1, -1)
    return O.Pot(name, s1, s2, **kw)


def two_species(nx, a, za, zb, seed, noise=0.05):
    """fcc with the two basis halves of species za and zb (every pair sum occurs)."""
    x, L = E.fcc(nx, a)
    rng = np.random.default_rng(seed)
    x = np.mod(x + rng.normal(0.0, noise, x.shape), L)
    z = np.where(rng.random(x.shape[0]) < 0.5, za, zb).astype(np.int32)
    return x, L, z


# Several potentials of one kind on disjoint species pairs: the batch evaluates
# one selected term per pair, and every parameter the kind reads must come from
# that term (the fastlj table, the lennardjonesMod exponents).
MULTI = {
    "fastlj": (lambda: two_species(4, 5.26, 10, 18, 41), [
        ("fastlj", "sigma 2.78 epsilon 0.0031 cutoff 8.0 bins 300", 10, 10),
        ("fastlj", "sigma 3.41 epsilon 0.0103 cutoff 8.5 bins 700", 18, 18),
        ("fastlj", "sigma 3.10 epsilon 0.0057 cutoff 7.5 bins 500", 10, 18)], "cutoff 8.5 skin 0.5"),
    "lennardjonesMod": (lambda: two_species(4, 5.26, 10, 18, 42), [
        ("lennardjonesMod", "sigma 2.78 epsilon 0.0031 m 10 n 5 cohesive 0.8 cutoff 8.0", 10, 10),
        ("lennardjonesMod", "sigma 3.41 epsilon 0.0103 m 12 n 6 cohesive 1.0 cutoff 8.5", 18, 18),
        ("lennardjonesMod", "sigma 3.10 epsilon 0.0057 m 9 n 7 cohesive 0.9 cutoff 8.5", 10, 18)],
        "cutoff 8.5 skin 0.5"),
    "buckingham": (None, None, None),  # the rocksalt NaCl set of the served jobs
}


@pytest.mark.parametrize("kind", sorted(MULTI))
def test_batch_selected_term_matches_oracle(kind):
    from lpmd2_amd import workloads as W

    say("batch selected term", kind)
    if kind == "buckingham":
        x, L, z = W.rocksalt(3, 3, 3, 5.64)
        x = np.mod(x + np.random.default_rng(43).normal(0.0, 0.03, x.shape), L)
        pots = [("buckingham", args, s1, s2) for args, s1, s2 in W.BUCKINGHAM_NACL]
        cm = "cutoff 8.0 skin 0.4"
    else:
        make, pots, cm = MULTI[kind]
        x, L, z = make()
    b = Batch(atom_capacity=2048)
    sid = b.add(x=x, box=L, z=z, potentials=pots, cm=cm, steps=0)
    b.run(1)
    _, _, a, info = b.get(sid)
    assert info["started"] == 1 and info["overflow"] == 0
    conf = O.Conf(x, L, z=z)
    cut = float(dict(zip(cm.split()[0::2], cm.split()[1::2]))["cutoff"])
    ocm = O.CellManager("verletlist", cut, maxneighbors=600, each=5)
    e0, _, _ = O.update_forces(conf, ocm, [oracle_pot(p) for p in pots])
    err = np.linalg.norm(a - conf.a) / np.linalg.norm(conf.a)
    ep, _, _ = b.energies(sid)
    say(f"  {kind}: {x.shape[0]} atoms, rel force rms {err:.2e}, E {ep:.9g} vs {e0:.9g}")
    assert err <= 1e-5
    assert abs(ep - e0) <= 1e-6 * abs(e0)
    # the plain engine on the same system agrees to the same bar
    s = E.System(x, L, z)
    s.set_cellmanager("verletlist", cm + " maxneighbors 600 each 5")
    for p in pots:
        s.append_potential(*p)
    ep2, _, _ = s.update_forces()
    _, _, a2 = s.get_state()
    assert np.linalg.norm(a2 - conf.a) / np.linalg.norm(conf.a) <= 1e-5
    assert abs(ep2 - e0) <= 1e-6 * abs(e0)

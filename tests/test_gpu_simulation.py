"""GPU checks of the drop-in DeviceSimulation (a Simulation whose steps stay
on the device): DoSteps over the drop-in integrators runs on the device in
one call and follows the engine's native trajectory; the step's tags come
from device reductions and equal SetSimulationTags (simulationbuilder.cc:84-124)
of the state read back; host views are refreshed only when read, and a host
change is sent before the next step; a non-device integrator falls back to
the per-step host loop (simulationbuilder.cc:126-136)."""
import sys

import numpy as np
import pytest

import lpmd2_amd as E
import sim_driver as D
from oracle import oracle as O
from oracle import refcheck as R

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not D.available(), reason="drop-in plugins not built")]

SC = "e 1.2382e-2 a 3.61 n 9 m 6 c 39.432 cutoff 7.22"
SC_KW = dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22)
CM = "cutoff 7.22 maxneighbors 400 skin 0.3"
NH_Q, NH_T = 1e7, 300.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(seed=8, T=600.0):
    x, L = E.fcc(5, 3.61)
    x = np.mod(x + np.random.default_rng(seed).normal(0.0, 0.02, x.shape), L)
    m = np.full(x.shape[0], E.ELEMENT_MASS[29])
    v = E.maxwell_boltzmann(m, T, seed=seed)
    return x, L, v


def dropin(name):
    return R.DROPIN_DIR + "/" + name + ".so"


def make_sim(x, L, v, scheme="velocityverlet", dt=2.0, integrator_path=None):
    s = D.Sim(x, L, 29, v=v)
    s.set_cellmanager(dropin("verletlist"), CM)
    s.add_potential(dropin("suttonchen"), SC)
    args = f"dt {dt} fmass {NH_Q} t {NH_T}" if scheme == "nosehoover" else f"dt {dt}"
    s.set_integrator(integrator_path or dropin(scheme), args)
    return s


def native(x, L, v, scheme, dt, steps, s=None):
    if s is None:
        s = E.System(x, L, 29, v=v)
        s.set_cellmanager("verletlist", CM)
        s.append_potential("suttonchen", SC)
        if scheme == "leapfrog":
            s.leapfrog_init(dt)
        elif scheme == "beeman":
            s.beeman_init(dt)
        elif scheme == "nosehoover":
            s.nosehoover_init(dt, NH_Q, NH_T)
    run = {"velocityverlet": s.velocityverlet, "leapfrog": s.leapfrog, "beeman": s.beeman,
           "nosehoover": s.nosehoover}[scheme]
    run(dt, steps)
    return s


def check_tags(sim, x, L, v, energy_here=True):
    """Every tag against SetSimulationTags' formulas on the state read back,
    with the simulation's own potential energy, virial and stress; the
    potential energy and virial against the FP64 oracle at those positions.
    Tags are stored as text with six significant digits (the stream's default
    precision), so a tag agrees to 1e-5 relative.  energy_here: the last force
    evaluation saw these positions (the two-step schemes; leapfrog's step
    evaluates its forces before the drift, OneStepIntegrator::Advance)."""
    conf = O.Conf(x, L, z=29, v=v)
    vol = float(np.prod(L))
    got = {k: sim.tag(k) for k in D.TAGS}
    pf = 160217.65
    virial = got["virial-pressure"] * 3.0 * vol / pf
    stress = np.array([[got["s" + a + b] for b in "xyz"] for a in "xyz"]) * vol / pf
    ref = O.tags(conf, got["potential-energy"], virial, stress)
    for k, i in D.TAGS.items():
        # sums of stored tags take the scale of their parts
        scale = {"pressure": abs(ref[23]) + abs(ref[24]), "total-energy": abs(ref[0]) + abs(ref[5])}.get(k, 0.0)
        assert abs(got[k] - ref[i]) <= 1e-5 * max(abs(ref[i]), scale) + 1e-12, (k, got[k], ref[i])
    if not energy_here:
        return
    e0, w0, _ = O.update_forces(conf, O.CellManager("verletlist", 7.22, maxneighbors=400, each=5),
                                [O.Pot("suttonchen", **SC_KW)])
    assert abs(got["potential-energy"] - e0) <= 1e-5 * abs(e0)
    assert abs(virial - w0) <= 2e-5 * max(abs(w0), abs(e0))


@pytest.mark.parametrize("scheme", ["velocityverlet", "leapfrog", "beeman", "nosehoover"])
def test_do_steps_on_device_follows_native_path(scheme):
    say("device simulation", scheme)
    x, L, v = setup()
    steps, dt = 30, 2.0
    sim = make_sim(x, L, v, scheme, dt)
    sim.do_steps(steps)
    c = sim.counters()
    assert c["device_runs"] == 1 and c["host_steps"] == 0 and c["refreshes"] == 0 and c["step"] == steps
    xs, vs, _ = sim.get()
    assert sim.counters()["refreshes"] == 1
    xn, vn, _ = native(x, L, v, scheme, dt, steps).get_state()
    dx, dv = np.abs(xs - xn).max(), np.abs(vs - vn).max()
    say(f"  max |dx| {dx:.3e}  |dv| {dv:.3e}")
    assert dx <= 1e-10 and dv <= 1e-12
    check_tags(sim, xs, L, vs, energy_here=scheme != "leapfrog")
    sim.close()


def test_host_views_are_lazy_and_host_changes_are_sent():
    """Reading the atoms refreshes them once; stepping on without a host
    change sends nothing; a velocity written through Atoms() is sent before
    the next step, which then follows the native run given the same change."""
    say("device simulation: lazy host views")
    x, L, v = setup(seed=11)
    dt = 2.0
    sim = make_sim(x, L, v)
    sim.do_steps(10)
    sim.do_steps(10)
    c = sim.counters()
    # the first run starts from the state Initialize handed over: no upload
    assert c == dict(refreshes=0, device_runs=2, uploads=0, host_steps=0, step=20)
    xs, vs, _ = sim.get()
    sim.do_steps(5)  # the host read the state but did not change it
    assert sim.counters()["uploads"] == 0
    xs, vs, _ = sim.get()
    assert sim.counters()["refreshes"] == 2
    vnew = vs[7] * 1.5
    sim.set_velocity(7, vnew)
    sim.do_steps(10)
    c = sim.counters()
    assert c["uploads"] == 1 and c["device_runs"] == 4 and c["step"] == 35
    xf, vf, _ = sim.get()
    s = native(x, L, v, "velocityverlet", dt, 25)
    xn, vn, _ = s.get_state()
    assert np.abs(xn - xs).max() <= 1e-10
    vn = vn.copy()
    vn[7] = vnew
    s.set_state(x=xn, v=vn)
    xn, vn, _ = native(None, None, None, "velocityverlet", dt, 10, s=s).get_state()
    dx, dv = np.abs(xf - xn).max(), np.abs(vf - vn).max()
    say(f"  after the host change: max |dx| {dx:.3e}  |dv| {dv:.3e}")
    assert dx <= 1e-10 and dv <= 1e-12
    check_tags(sim, xf, L, vf)
    sim.close()


def test_do_step_calls_equal_one_do_steps():
    say("device simulation: DoStep x 12 against DoSteps(12)")
    x, L, v = setup(seed=13)
    a = make_sim(x, L, v)
    for _ in range(12):
        a.do_step()
    b = make_sim(x, L, v)
    b.do_steps(12)
    assert a.counters()["device_runs"] == 12 and b.counters()["device_runs"] == 1
    xa, va, _ = a.get()
    xb, vb, _ = b.get()
    assert np.abs(xa - xb).max() <= 1e-10 and np.abs(va - vb).max() <= 1e-12
    assert a.tag("step") == b.tag("step") == 12
    a.close()
    b.close()


def test_reference_integrator_runs_the_host_loop():
    """The reference's own velocityverlet plugin is no device integrator: each
    step is SimulationEngine::DoStep on the host over the drop-in forces, and
    the host tags are SetSimulationTags of that state."""
    say("device simulation: host loop")
    x, L, v = setup(seed=17)
    steps, dt = 10, 2.0
    sim = make_sim(x, L, v, integrator_path=R.so("velocityverlet").decode())
    sim.do_steps(steps)
    c = sim.counters()
    assert c["host_steps"] == steps and c["device_runs"] == 0 and c["step"] == steps
    xs, vs, _ = sim.get()
    xn, vn, _ = native(x, L, v, "velocityverlet", dt, steps).get_state()
    assert np.abs(xs - xn).max() <= 1e-10 and np.abs(vs - vn).max() <= 1e-12
    check_tags(sim, xs, L, vs)
    sim.close()


def test_restore_sends_the_restored_state(tmp_path):
    """Dump after a device run, run on, Restore: the next device run starts
    from the restored state (Dump's six significant digits), as a native run
    started from that state does."""
    say("device simulation: Dump / Restore between device runs")
    x, L, v = setup(seed=19)
    dt = 2.0
    sim = make_sim(x, L, v)
    sim.do_steps(10)
    path = str(tmp_path / "state.dump")
    sim.dump(path)
    sim.do_steps(7)
    sim.restore(path)
    assert sim.counters()["step"] == 10
    xr, vr, _ = sim.get()
    sim.do_steps(10)
    assert sim.counters()["uploads"] == 1 and sim.counters()["step"] == 20
    xs, vs, _ = sim.get()
    xn, vn, _ = native(xr, L, vr, "velocityverlet", dt, 10).get_state()
    dx, dv = np.abs(xs - xn).max(), np.abs(vs - vn).max()
    say(f"  max |dx| {dx:.3e}  |dv| {dv:.3e}")
    assert dx <= 1e-10 and dv <= 1e-12
    sim.close()

"""GPU checks of the serving loop through the C ABI: every job's result is the
same when it runs alone and when it runs in a mix of arrivals (bit for bit:
a job's arithmetic does not depend on its neighbours in the batch), jobs are
admitted by HBM budget as memory frees up, jobs that arrive while others run
join them, and bad jobs fail on their own."""
import sys
import time

import numpy as np
import pytest

import lpmd2_amd as E
from lpmd2_amd import workloads as W
from lpmd2_amd.batch import Batch
from lpmd2_amd.server import Server

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_jobs(seed=11, n_jobs=8):
    return [j for _, j in W.served_jobs(seed=seed, n_jobs=n_jobs, min_atoms=300, max_atoms=3000, min_steps=10,
                                         max_steps=80)]


def run_alone(job):
    with Server(1 << 30) as s:
        jid = s.submit_job(job)
        assert s.wait(jid, timeout=120.0)
        return s.result(jid)


def test_served_job_equals_batch_run():
    say("server vs batch")
    job = small_jobs(seed=3, n_jobs=1)[0]
    x, v, a, info = run_alone(job)
    b = Batch(atom_capacity=8192)
    sid = b.add(x=job["x"], box=job["box"], z=job["z"], potentials=job["potentials"], cm=job["cm"], v=job["v"],
                dt=job["dt"], steps=job["steps"])
    b.run(job["steps"] + 1)
    xb, vb, ab, ib = b.get(sid)
    assert ib["done"] == job["steps"] and info["steps"] == job["steps"]
    assert np.array_equal(x, xb) and np.array_equal(v, vb) and np.array_equal(a, ab)
    ep, vir, ek = b.energies(sid)
    assert (ep, vir, ek) == (info["epot"], info["virial"], info["ekin"])


def test_results_same_alone_and_in_mix():
    jobs = small_jobs()
    say("mix of", len(jobs), "jobs:", [j["x"].shape[0] for j in jobs], [j["steps"] for j in jobs])
    with Server(1 << 30, max_round_steps=32) as s:
        ids = []
        for k, j in enumerate(jobs):
            ids.append(s.submit_job(j))
            if k % 3 == 2:
                time.sleep(0.01)  # some arrive while earlier ones run
        assert s.wait(timeout=300.0)
        mixed = [s.result(i) for i in ids]
        st = s.stats()
    say("  stats", st)
    assert st["done"] == len(jobs) and st["failed"] == 0
    assert st["max_resident"] >= 2
    for j, (x, v, a, info) in zip(jobs, mixed):
        xa, va, aa, ia = run_alone(j)
        assert info["steps"] == j["steps"]
        assert np.array_equal(x, xa), np.abs(x - xa).max()
        assert np.array_equal(v, va)
        assert np.array_equal(a, aa)
        assert info["epot"] == ia["epot"] and info["ekin"] == ia["ekin"]


def test_admission_by_budget():
    jobs = small_jobs(seed=5, n_jobs=10)
    need = max(j["x"].shape[0] for j in jobs) * E.batch.bytes_per_atom() * 2 * 1.3 / 0.25
    budget = int(max(need, 4 << 20))
    say("budget", budget)
    with Server(budget, max_round_steps=16) as s:
        ids = [s.submit_job(j) for j in jobs]
        assert s.wait(timeout=300.0)
        st = s.stats()
        say("  stats", st)
        assert st["done"] == len(jobs) and st["failed"] == 0
        # the budget holds a few of these jobs, not all of them
        assert 1 <= st["max_resident"] < len(jobs)
        res = [s.result(i) for i in ids]
    for j, (x, v, a, info) in list(zip(jobs, res))[:3]:
        xa, va, aa, ia = run_alone(j)
        assert np.array_equal(x, xa) and np.array_equal(v, va)


def test_arrivals_join_running_jobs():
    jobs = small_jobs(seed=7, n_jobs=2)
    long_job = dict(jobs[0], steps=5000)
    short_job = dict(jobs[1], steps=10)
    with Server(1 << 30, max_round_steps=64) as s:
        t0 = time.perf_counter()
        a = s.submit_job(long_job)
        time.sleep(0.01)
        b = s.submit_job(short_job)
        t1 = time.perf_counter()
        assert s.wait(timeout=300.0)
        _, _, _, ia = s.result(a)
        _, _, _, ib = s.result(b)
    say("  long", ia, "short", ib)
    # the short job was admitted beside the long one and finished first
    assert ib["latency_s"] + (t1 - t0) < ia["latency_s"]
    assert ib["queued_s"] < 0.5 * ia["latency_s"]


def test_bad_jobs_fail_alone():
    good = small_jobs(seed=9, n_jobs=1)[0]
    with Server(8 << 20) as s:
        bad = s.submit(good["x"], good["box"], good["z"], [("nosuchpotential", "")], cm=good["cm"], steps=5)
        x, L = E.fcc(30, 3.61)  # 108000 atoms: beyond an 8 MiB budget
        big = s.submit(x, L, 29, [("gupta", "A 0.0855 r0 2.556 p 10.96 B 1.224 q 2.278 cutoff 7.22")],
                       cm="cutoff 7.22 skin 0.3", steps=5)
        ok = s.submit_job(dict(good, steps=5))
        assert s.wait(timeout=120.0)
        with pytest.raises(E.LpmdxError) as e1:
            s.result(bad)
        assert e1.value.code == E.EINVAL
        with pytest.raises(E.LpmdxError) as e2:
            s.result(big)
        assert e2.value.code == E.ENOMEM
        x1, v1, a1, info = s.result(ok)
        assert info["steps"] == 5
        st = s.stats()
        assert st["done"] == 1 and st["failed"] == 2

"""Served form (gcs_server_*): each job's result is the same alone and in a mix.

Jobs are admitted, stepped together with whatever else is resident, and
retired as they finish; the composition of the forest changes between rounds.
Bar: a job's final state is bit-identical to the same job run alone through
the server and through a plain engine (instances share no arithmetic).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 77
EPS2 = float(np.float32(0.02) ** 2)
KW = dict(G=1.0, eps2=EPS2, dt=0.005, theta=0.5, leaf_size=8, use_graph=1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_server(gcs, jobs, **cfg):
    cfg.setdefault("gen_threads", 4)
    with gcs.Server(hbm_budget_bytes=4 << 30, keep_final=True, **cfg,
                    precision=gcs.PREC_COMPENSATED, **KW) as s:
        s.submit(jobs, SEED)
        s.run(0.0)
        res = s.results()
        finals = {r["id"]: s.final(r["id"], r["n"]) for r in res}
        return res, finals, s.stats()


MIX = [(0, 2048, 40, 0.0), (1, 6000, 10, 0.0), (2, 4096, 30, 0.0), (3, 12000, 25, 0.0),
       (4, 3000, 50, 0.0), (5, 9002, 7, 0.0)]


@pytest.mark.parametrize("cap", [0, 16000])
def test_job_alone_equals_job_in_mix(gcs, cap):
    say("server mix cap=%d" % cap)
    res, finals, st = run_server(gcs, MIX, steps_per_round=8, max_resident=cap)
    assert sorted(r["id"] for r in res) == [j[0] for j in MIX]
    assert st["particle_steps"] == sum(n * k for _, n, k, _ in MIX)
    for r in res:
        assert r["admit"] >= r["arrival"] and r["finish"] >= r["admit"]
    if cap:
        assert st["resident_cap"] == cap
    for jid, n, k, _ in MIX:
        say("  job %d n=%d steps=%d alone" % (jid, n, k))
        _, alone, _ = run_server(gcs, [(jid, n, k, 0.0)], steps_per_round=8)
        a, b = alone[jid], finals[jid]
        for f in ("x", "y", "z", "vx", "vy", "vz", "mass", "r", "g", "b", "type"):
            assert np.array_equal(a[f], b[f]), (jid, f)


def test_job_equals_plain_engine(gcs):
    jid, n, k = 2, 4096, 30
    say("server job vs engine n=%d steps=%d" % (n, k))
    _, finals, _ = run_server(gcs, [(jid, n, k, 0.0)], steps_per_round=7)
    st = gcs.generate(gcs.random_spec(SEED, jid, n, n))
    p = gcs.params(precision=gcs.PREC_COMPENSATED, **KW)
    with gcs.Engine(n) as e:
        e.load(st)
        e.step(p, k)
        ref = e.instance(0)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(ref[f], finals[jid][f]), f


def test_admission_respects_caps(gcs):
    jobs = [(i, 4000 + 500 * i, 6 + i, 0.0) for i in range(8)]
    say("server caps: max_jobs=2")
    res, _, st = run_server(gcs, jobs, steps_per_round=4, max_jobs=2)
    assert len(res) == len(jobs)
    # at most two resident at any time: every job is admitted no earlier than
    # the second-to-last finish among those admitted before it
    order = sorted(res, key=lambda r: r["admit"])
    for i in range(2, len(order)):
        fin = sorted(r["finish"] for r in order[:i])
        assert order[i]["admit"] >= fin[i - 2]
    assert st["particle_steps"] == sum(n * k for _, n, k, _ in jobs)


def test_device_generated_jobs(gcs):
    """gen_threads = 0: each job is generated on the device at admission.  A
    job in the mix equals the same collision generated on the device and
    stepped by a plain engine, next to another instance."""
    say("server with device generation")
    res, finals, st = run_server(gcs, MIX[:4], steps_per_round=8, gen_threads=0)
    assert sorted(r["id"] for r in res) == [j[0] for j in MIX[:4]]
    assert st["particle_steps"] == sum(n * k for _, n, k, _ in MIX[:4])
    p = gcs.params(precision=gcs.PREC_COMPENSATED, **KW)
    for jid, n, k, _ in MIX[:4]:
        say("  job %d n=%d steps=%d, engine" % (jid, n, k))
        other = gcs.random_spec(SEED + 1, jid, 3000, 3000)
        with gcs.Engine(n + 3000, 2) as e:
            e.add_generated(other)
            before = e.instance(0).copy()
            iid = e.add_generated(gcs.random_spec(SEED, jid, n, n))
            assert e.instance(0).tobytes() == before.tobytes()  # the other is untouched
            e.step(p, k)
            ref = e.instance(iid)
        for f in ("x", "y", "z", "vx", "vy", "vz", "mass", "r", "g", "b", "type"):
            assert np.array_equal(ref[f], finals[jid][f]), (jid, f)

"""Serving (cupss_server_* through the Python Server binding): jobs of
different sizes and models (one or two dynamic fields) arrive over time, are admitted by memory budget
and advance together with continuous batching.  Checked here: each job's
result is bit-identical to advancing its system alone (also for model-H
jobs at the served configuration's sizes, 128^2 to 1024^2), admission keeps
arrival order and the budget, and the general step program replayed from a
hipGraph equals direct launches.
"""
import numpy as np
import pytest

import cupss_amd as C
from cupss_amd.systems import build, spec_cahn_hilliard, spec_gray_scott, spec_kpz, spec_model_h

pytestmark = pytest.mark.gpu


def job_specs():
    # (name, spec, steps); field 0 is a dynamic field in every spec (the
    # Gray-Scott jobs have two)
    return [
        ("modelh64", spec_model_h(64, seed=11), 37),
        ("kpz64", spec_kpz(64, nu=1.0, lam=1.5, seed=5), 50),
        ("ch128", spec_cahn_hilliard((128, 128, 1), seed=4), 23),
        ("modelh128", spec_model_h(128, seed=12), 61),
        ("ew64", spec_kpz(64, nu=0.7, lam=0.0, seed=9), 44),
        ("modelh64b", spec_model_h(64, seed=13), 10),
        ("kpz128", spec_kpz(128, nu=1.3, lam=0.8, seed=6), 31),
        ("gs64", spec_gray_scott(64, seed=15), 29),
        ("gs128", spec_gray_scott(128, seed=16), 41),
    ]


def alone(spec, nsteps):
    s, _ = build(spec)
    s.step(nsteps)
    out = s.get_real(0)
    s.close()
    return out


def serve(jobs, budget, quantum, max_active, gap):
    results = {}
    est = {}
    for name, spec, _ in jobs:
        s, _ = build(spec)
        est[name] = s.device_bytes()
        s.close()
    srv = C.Server(budget, quantum=quantum, max_active=max_active)
    ids = []
    for i, (name, spec, n) in enumerate(jobs):
        def done(sysv, name=name):
            results[name] = sysv.get_real(0)
        ids.append(srv.submit(lambda spec=spec: build(spec)[0], n, arrival=gap * i, bytes_estimate=est[name],
                              done=done))
    srv.run()
    times = [srv.times(j) for j in ids]
    stats = srv.stats()
    srv.close()
    return results, times, stats, est


@pytest.mark.parametrize("quantum,max_active", [(8, 3), (4, 16)])
def test_served_results_equal_alone(quantum, max_active):
    jobs = job_specs()
    results, times, stats, _ = serve(jobs, budget=1 << 40, quantum=quantum, max_active=max_active, gap=0.003)
    print("served %d jobs in %.3f s, peak active %d" % (len(jobs), stats["wall_s"], stats["peak_active"]))
    for name, spec, n in jobs:
        ref = alone(spec, n)
        assert np.array_equal(results[name], ref), name
    for a, b, f in times:
        assert 0 <= a <= b <= f


def test_full_size_model_h_stream_equals_alone():
    """Configuration C5's job mix at its sizes: model-H jobs of 128^2,
    256^2, 512^2 and 1024^2 (and a second 128^2 and 512^2) arriving over
    time, served with continuous batching under a budget that holds only
    part of them at once; each result equals the job's system advanced
    alone, bit for bit."""
    jobs = [("mh128", spec_model_h(128, seed=51), 120), ("mh1024", spec_model_h(1024, seed=52), 30),
            ("mh256", spec_model_h(256, seed=53), 90), ("mh512", spec_model_h(512, seed=54), 60),
            ("mh128b", spec_model_h(128, seed=55), 75), ("mh512b", spec_model_h(512, seed=56), 40)]
    est = []
    for _, spec, _ in jobs:
        sv, _ = build(spec)
        est.append(sv.device_bytes())
        sv.close()
    budget = max(est) + sorted(est)[-2]
    results, times, stats, _ = serve(jobs, budget=budget, quantum=10, max_active=8, gap=0.002)
    print("served %d model-H jobs (128..1024) in %.3f s, peak active %d, peak %d B of %d" %
          (len(jobs), stats["wall_s"], stats["peak_active"], stats["peak_bytes"], budget))
    assert stats["peak_bytes"] <= budget
    for name, spec, n in jobs:
        ref = alone(spec, n)
        assert np.array_equal(results[name], ref), name


def test_admission_keeps_budget_and_arrival_order():
    jobs = job_specs()
    est = []
    for _, spec, _ in jobs:
        s, _ = build(spec)
        est.append(s.device_bytes())
        s.close()
    budget = 2 * max(est)
    results, times, stats, _ = serve(jobs, budget=budget, quantum=6, max_active=64, gap=0.0)
    print("budget %d B, peak %d B, peak active %d" % (budget, stats["peak_bytes"], stats["peak_active"]))
    assert stats["peak_bytes"] <= budget
    admitted = [t[1] for t in times]
    assert all(admitted[i] <= admitted[i + 1] for i in range(len(admitted) - 1))
    assert len(results) == len(jobs)
    expected = sum(spec["dims"][0] * spec["dims"][1] * n for _, spec, n in jobs)
    assert stats["point_steps"] == pytest.approx(expected)


def test_general_program_graph_replay_equals_direct_launches():
    spec = spec_model_h(64, seed=21)
    a, _ = build(spec)
    b, _ = build(spec)
    b.set_graph_steps(6)
    for n in (13, 6, 1, 20):
        a.step(n)
        b.step(n)
        for f in range(len(spec["fields"])):
            assert np.array_equal(a.get_real(f), b.get_real(f)), (n, f)

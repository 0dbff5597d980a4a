"""GPU: parameter changes in the middle of a run (evolver::updateParameter,
src/evolver.cpp:386-394 -> new prefactors of terms and implicit operators).

* The fused paths (2D two-kernel, 3D four-pass) take the new values at the
  next step: after 12 steps, new prefactors, 12 more steps, they agree with
  the general step program given the same changes (relative L2 <= 1e-5),
  and differ from a run without the change.
* A change of one batch member's values (same structure: no re-planning of
  the batch) leaves it bit-identical to the same member advanced alone with
  the same change, and the other members untouched.
* A change of structure (a monomial with other powers) re-plans and still
  agrees with the general program.
"""
import numpy as np
import pytest

import cupss_amd as C
from tests.helpers import build_system, rel_l2, spec_cahn_hilliard, spec_kpz

pytestmark = pytest.mark.gpu


def make(spec, path=0, noise_mode=0):
    g, ids = build_system(C.System, C.pres, spec)
    g.set_path(path)
    g.set_noise_mode(noise_mode)
    return g, ids


def kpz_change(g, ids, lam=1.7, nu=0.6):
    # terms 0 and 1: the two gradient products; the implicit nu*q^2 of h
    g.set_term_prefactors(0, [C.pres(0.5 * lam)])
    g.set_term_prefactors(1, [C.pres(0.5 * lam)])
    g.set_implicits(ids["h"], [C.pres(-nu, 1)])


def ch_change(g, ids, a=-0.8, b=1.4, k=1.2):
    g.set_term_prefactors(0, [C.pres(-b, 1)])
    g.set_implicits(ids["phi"], [C.pres(-a, 1), C.pres(-k, 2)])


CASES = {
    "kpz_2d": (lambda: spec_kpz(64, nu=1.1, lam=0.9, D=0.01, dt=0.005, seed=3,
                                h0=np.random.default_rng(3).uniform(-0.1, 0.1, 64 * 64).astype(np.float32)),
               kpz_change, "h"),
    "ch_2d_512": (lambda: spec_cahn_hilliard((512, 512, 1), seed=4), ch_change, "phi"),
    "ch_3d": (lambda: spec_cahn_hilliard((32, 32, 32), seed=5), ch_change, "phi"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_parameter_update_mid_run(case):
    mk, change, key = CASES[case]
    print("parameter update mid-run:", case, flush=True)
    runs = {}
    for name, path, update in (("fused", 0, True), ("general", 1, True), ("unchanged", 0, False)):
        g, ids = make(mk(), path, noise_mode=1)
        g.step(12)
        if update:
            change(g, ids)
        g.step(12)
        runs[name] = g.get_real(ids[key])
    e = rel_l2(runs["fused"], runs["general"])
    d = rel_l2(runs["fused"], runs["unchanged"])
    print("  fused-vs-general %.3e, effect of the change %.3e" % (e, d), flush=True)
    assert e <= 1e-5
    assert d > 1e-4


def test_batch_member_update_equals_alone():
    specs = [spec_kpz(n, nu=1.0 + 0.1 * i, lam=0.8, D=0.01, dt=0.005, seed=20 + i,
                      h0=np.random.default_rng(i).uniform(-0.1, 0.1, n * n).astype(np.float32))
             for i, n in enumerate((64, 128, 64))]
    batch = C.Batch(graph_steps=4)
    members = []
    for sp in specs:
        g, ids = make(sp, noise_mode=1)
        batch.add(g)
        members.append((g, ids))
    batch.step(10)
    kpz_change(*members[1])
    batch.step(10)
    for i, sp in enumerate(specs):
        a, ia = make(sp, noise_mode=1)
        a.step(10)
        if i == 1:
            kpz_change(a, ia)
        a.step(10)
        assert np.array_equal(a.get_real(ia["h"]), members[i][0].get_real(members[i][1]["h"])), i


def test_structure_change_replans():
    spec = spec_cahn_hilliard((64, 64, 1), seed=6)
    out = []
    for path in (0, 1):
        g, ids = make(spec, path)
        g.step(8)
        # the product term gains a q^4 monomial: another multiplier shape
        g.set_term_prefactors(0, [C.pres(-1.0, 1), C.pres(-0.05, 2)])
        g.step(8)
        out.append(g.get_real(ids["phi"]))
    e = rel_l2(out[0], out[1])
    print("  structure change, fused-vs-general %.3e" % e, flush=True)
    assert e <= 1e-5

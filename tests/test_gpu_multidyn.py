"""GPU: systems with several dynamic fields on the fused 2D step and in the
batched step.  The first dynamic field is phi; the others are stored fields
that the column pass advances with their own semi-implicit Euler step
(setDynamic_k, src/field_kernels.cu:199-227), in the reference's phase
order (src/evolver.cpp:199-226): constraints first, from the dynamic fields
after the previous step, then every dynamic field from the same pre-step
state.

* Gray-Scott reaction-diffusion (u, v: products u v^2 in both, a constant
  and linear terms) and a coupled system that reads every source kind a
  second dynamic field adds -- a derived field of it (gx = iqx psi), a
  constraint solved from both dynamic fields (w (1 + q^2) = psi + 0.3 phi),
  products of a dynamic field with that constraint -- against the oracle's
  FP32 scheme (relative L2 <= 1e-4 after 100 steps, as the one-field
  cases).
* Runs split into pieces (resume of the column-major state and of the
  stored fields) are bit-identical to one run, and a batch member is
  bit-identical to the same system advanced alone.
* Noise on a second dynamic field, drawn in real or in spectral space: the
  fused step equals the general program (same Philox counters; FP32
  rounding only).
* A grid whose row kernel does not exist for the number of real fields
  (more than 3 at sx = 32) takes the general program and still matches the
  oracle.
* Constraints whose products read dynamic fields ("fresh": the reference's
  phase A of step m reads the dynamic fields of step m, so the column pass
  solves them before the dynamic update): Cahn-Hilliard written with its
  chemical potential mu = b phi^3 against the oracle and against the usual
  form, and a two-field system where such a constraint is read by both
  dynamic fields and, one step late, by another constraint.
* A constraint field without terms keeps its value (a coefficient field;
  setNotDynamic_k writes nothing): example 07's inhomogeneous diffusion
  without its callbacks, against the oracle.
* Noisy constraint fields (setNotDynamic_k adds the noise times 1/sqrt(dt)
  before the implicit division): a pure-noise field, a constraint of
  dynamic-field products and a constraint solved for the next step, which
  draws that step's noise; fused against the general program with the
  same Philox counters, in both noise modes.
* Parameter changes in the middle of a run (evolver::updateParameter,
  src/evolver.cpp:386-394) reach the second dynamic field's terms and
  implicit operator: fused against general, and a batch member against the
  same system advanced alone, bit for bit.
"""
import numpy as np
import pytest

import cupss_amd as C
from oracle import oracle as O
from cupss_amd.systems import spec_gray_scott
from tests.helpers import build_system, rel_l2, spec_cahn_hilliard

pytestmark = pytest.mark.gpu


def gpu(spec, path=0, noise_mode=0, profile=False):
    g, ids = build_system(C.System, C.pres, spec)
    g.set_path(path)
    g.set_noise_mode(noise_mode)
    if profile:
        g.profile_kernel(True)
    return g, ids


def spec_coupled(dims=(64, 32, 1), seed=32, noise=None):
    """dt phi + q^2 phi = -0.2 q^2 phi^3 + 0.5 gx phi + 0.2 w
       dt psi + 0.5 q^2 psi = 0.3 iqx phi - 0.1 psi w
       gx = iqx psi;   w (1 + q^2) = psi + 0.3 phi"""
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    terms = [("phi", ["phi", "phi", "phi"], [(-0.2, 1)]), ("phi", ["gx", "phi"], [(0.5,)]), ("phi", ["w"], [(0.2,)]),
             ("psi", ["phi"], [(0.3, 0, 1)]), ("psi", ["psi", "w"], [(-0.1,)]),
             ("gx", ["psi"], [(1.0, 0, 1)]),
             ("w", ["psi"], [(1.0,)]), ("w", ["phi"], [(0.3,)])]
    spec = {"dims": dims, "dt": 0.01,
            "fields": [("phi", True), ("psi", True), ("gx", False), ("w", False)],
            "implicit": {"phi": [(-1.0, 1)], "psi": [(-0.5, 1)], "w": [(1.0,), (1.0, 1)]},
            "terms": terms,
            "init": {"phi": rng.uniform(-0.2, 0.2, n).astype(np.float32),
                     "psi": rng.uniform(-0.2, 0.2, n).astype(np.float32)}}
    if noise is not None:
        spec["noise"] = {"psi": ((noise,), seed + 1)}
    return spec


CASES = {"gray_scott": (spec_gray_scott, ("u", "v")),
         "coupled": (spec_coupled, ("phi", "psi", "gx", "w"))}


def assert_fused_stored_dynamic(g):
    name = g.kernel_time()[2]
    # per-pass column kernel, or the whole-instance kernel of 64^2 / 128^2
    assert name.startswith(("k_fcols<", "k_fsmall<")) and name.endswith("true, true>"), name


@pytest.mark.parametrize("case", sorted(CASES))
def test_multi_dynamic_fused_vs_oracle(case):
    mk, names = CASES[case]
    spec = mk()
    print("several dynamic fields, fused vs oracle:", case, flush=True)
    g, gid = gpu(spec, profile=True)
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(100)
    o.step(100)
    assert_fused_stored_dynamic(g)
    for f in names:
        got, ref = g.get_real(gid[f]), o.get_real(oid[f])
        assert np.all(np.isfinite(got))
        err = rel_l2(got, ref)
        print("  %-3s gpu-vs-f32 %.3e" % (f, err), flush=True)
        assert err <= 1e-4, f


@pytest.mark.parametrize("case", sorted(CASES))
def test_multi_dynamic_split_runs_and_batch_bitwise(case):
    mk, names = CASES[case]
    spec = mk()
    one, ia = gpu(spec)
    one.step(30)
    split, ib = gpu(spec)
    for n in (7, 1, 12, 10):
        split.step(n)
    for f in names:
        assert np.array_equal(one.get_real(ia[f]), split.get_real(ib[f])), f
    # a batch of this system, a one-field system and the other case
    other = spec_coupled((64, 64, 1), seed=40) if case == "gray_scott" else spec_gray_scott(32, seed=41)
    others = [spec_cahn_hilliard((32, 64, 1), seed=42), other]
    batch = C.Batch(graph_steps=4)
    members = []
    for sp in [spec] + others:
        g, ids = gpu(sp)
        batch.add(g)
        members.append((g, ids, sp))
    batch.step(30)
    for g, ids, sp in members:
        alone, ia2 = gpu(sp)
        alone.step(30)
        for f in ids:
            assert np.array_equal(alone.get_real(ia2[f]), g.get_real(ids[f])), f


@pytest.mark.parametrize("noise_mode", [0, 1])
def test_noise_on_second_dynamic_field(noise_mode):
    """Noise on psi, the second dynamic field: the fused step against the
    general program with the same Philox counters (real space: the row pass
    draws psi's cells into psi's own noise rows; spectral: the column pass
    draws its bins)."""
    spec = spec_coupled(noise=0.02)
    runs = {}
    for path in (0, 1):
        g, ids = gpu(spec, path=path, noise_mode=noise_mode, profile=path == 0)
        g.step(40)
        if path == 0:
            assert g.fused_eligible
            assert_fused_stored_dynamic(g)
        runs[path] = (g, ids)
    for f in ("phi", "psi", "w"):
        e = rel_l2(runs[0][0].get_real(runs[0][1][f]), runs[1][0].get_real(runs[1][1][f]))
        print("  noise mode %d, %-3s fused-vs-general %.3e" % (noise_mode, f, e), flush=True)
        assert e <= 1e-5, f


def test_no_row_kernel_for_the_grid_takes_the_general_program():
    spec = spec_coupled((32, 64, 1), seed=44)  # 4 real fields, sx/2 = 16: no wide row kernel
    g, gid = gpu(spec)
    assert not g.fused_eligible
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(50)
    o.step(50)
    for f in ("phi", "psi"):
        err = rel_l2(g.get_real(gid[f]), o.get_real(oid[f]))
        print("  32x64, %-3s general-vs-f32 %.3e" % (f, err), flush=True)
        assert err <= 1e-4, f


def spec_mu_form(dims=(128, 64, 1), seed=45, a=-1.0, b=1.0, k=1.0):
    """dt phi + (a q^2 + k q^4) phi = -q^2 mu;   mu = b phi^3"""
    n = dims[0] * dims[1] * dims[2]
    x0 = np.random.default_rng(seed).uniform(-0.1, 0.1, n).astype(np.float32)
    return {"dims": dims, "dt": 0.01, "fields": [("phi", True), ("mu", False)],
            "implicit": {"phi": [(-a, 1), (-k, 2)]},
            "terms": [("phi", ["mu"], [(-1.0, 1)]), ("mu", ["phi", "phi", "phi"], [(b,)])],
            "init": {"phi": x0}}


def spec_fresh_coupled(dims=(64, 64, 1), seed=46):
    """dt phi + q^2 phi + q^4 phi = -q^2 mu
       dt psi + 0.5 q^2 psi = 0.2 mu + 0.1 s psi
       mu = phi^3 + 0.3 phi psi + 0.1 psi - 0.05   (fresh: products of dynamic fields)
       s = 0.5 mu                                 (reads mu one step late)"""
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    terms = [("phi", ["mu"], [(-1.0, 1)]),
             ("psi", ["mu"], [(0.2,)]), ("psi", ["s", "psi"], [(0.1,)]),
             ("mu", ["phi", "phi", "phi"], [(1.0,)]), ("mu", ["phi", "psi"], [(0.3,)]), ("mu", ["psi"], [(0.1,)]),
             ("mu", [], [(-0.05,)]),
             ("s", ["mu"], [(0.5,)])]
    return {"dims": dims, "dt": 0.01, "fields": [("phi", True), ("psi", True), ("mu", False), ("s", False)],
            "implicit": {"phi": [(1.0, 1), (-1.0, 2)], "psi": [(-0.5, 1)]}, "terms": terms,
            "init": {"phi": rng.uniform(-0.1, 0.1, n).astype(np.float32),
                     "psi": rng.uniform(-0.1, 0.1, n).astype(np.float32)}}


FRESH = {"mu_form": (spec_mu_form, ("phi", "mu")), "fresh_coupled": (spec_fresh_coupled, ("phi", "psi", "mu", "s"))}


@pytest.mark.parametrize("case", sorted(FRESH))
def test_fresh_constraints_fused_vs_oracle(case):
    mk, names = FRESH[case]
    spec = mk()
    print("constraints of the dynamic fields' products, fused vs oracle:", case, flush=True)
    g, gid = gpu(spec, profile=True)
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(100)
    o.step(100)
    assert_fused_stored_dynamic(g)
    for f in names:
        got, ref = g.get_real(gid[f]), o.get_real(oid[f])
        assert np.all(np.isfinite(got))
        err = rel_l2(got, ref)
        print("  %-3s gpu-vs-f32 %.3e" % (f, err), flush=True)
        assert err <= 1e-4, f
    if case == "mu_form":
        # the same scheme as the usual form: dt phi + (a q^2 + k q^4) phi = -b q^2 phi^3
        u, uid = gpu(spec_cahn_hilliard(spec["dims"], seed=45))
        u.set_real(uid["phi"], spec["init"]["phi"])
        u.step(100)
        err = rel_l2(g.get_real(gid["phi"]), u.get_real(uid["phi"]))
        print("  mu form vs usual form %.3e" % err, flush=True)
        assert err <= 1e-5


@pytest.mark.parametrize("case", sorted(FRESH))
def test_fresh_constraints_split_runs_and_batch_bitwise(case):
    mk, names = FRESH[case]
    spec = mk()
    one, ia = gpu(spec)
    one.step(30)
    split, ib = gpu(spec)
    for n in (7, 1, 12, 10):
        split.step(n)
    for f in names:
        assert np.array_equal(one.get_real(ia[f]), split.get_real(ib[f])), f
    batch = C.Batch(graph_steps=4)
    members = []
    for sp in (spec, spec_gray_scott(64, seed=47), spec_cahn_hilliard((64, 64, 1), seed=48)):
        g, ids = gpu(sp)
        batch.add(g)
        members.append((g, ids, sp))
    batch.step(30)
    for g, ids, sp in members:
        alone, ia2 = gpu(sp)
        alone.step(30)
        for f in ids:
            assert np.array_equal(alone.get_real(ia2[f]), g.get_real(ids[f])), f


def spec_inhomogeneous_diffusion(n=128, seed=49):
    """Example 07's system (examples/07_inhomogeneous_diffusion/
    07_inhomogeneous_diffusion.cu:62-81) without the mirror callbacks:
    dt v + 0.5 q^2 v = iqx (x iqxv) + iqy^2 (x v);  iqxv = iqx v;  x fixed."""
    i = np.arange(n, dtype=np.float64)
    x = np.tile(1.0 + np.tanh((i - 0.45 * n) / (0.05 * n)), (n, 1)).astype(np.float32)
    v = np.random.default_rng(seed).uniform(0.0, 1.0, (n, n)).astype(np.float32)
    terms = [("v", ["x", "iqxv"], [(1.0, 0, 1)]), ("v", ["x", "v"], [(1.0, 0, 0, 2)]),
             ("iqxv", ["v"], [(1.0, 0, 1)])]
    return {"dims": (n, n, 1), "dt": 0.01, "fields": [("v", True), ("iqxv", False), ("x", False)],
            "implicit": {"v": [(-0.5, 1)]}, "terms": terms, "init": {"v": v.ravel(), "x": x.ravel()}}


def test_termless_constraint_with_implicit_operator():
    """A constraint without terms but with an implicit operator keeps its
    value and divides it by the operator's sum every step (q != 0), as
    setNotDynamic_k does: x (1 + 0.5 q^2) = (nothing), read by v's products."""
    spec = spec_inhomogeneous_diffusion(64, seed=55)
    spec["implicit"] = dict(spec["implicit"], x=[(1.0,), (0.5, 1)])
    g, gid = gpu(spec, profile=True)
    assert g.fused_eligible
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(30)
    o.step(30)
    for f in ("v", "x"):
        err = rel_l2(g.get_real(gid[f]), o.get_real(oid[f]))
        print("  %-2s gpu-vs-f32 %.3e" % (f, err), flush=True)
        assert err <= 1e-4, f
    assert rel_l2(g.get_real(gid["x"]), spec["init"]["x"]) > 1e-3  # x is smoothed every step


def test_coefficient_field_keeps_its_value():
    spec = spec_inhomogeneous_diffusion()
    g, gid = gpu(spec, profile=True)
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(100)
    o.step(100)
    name = g.kernel_time()[2]
    assert name.startswith(("k_fcols<", "k_fsmall<")) and name.endswith("true, false>"), name
    for f in ("v", "iqxv", "x"):
        err = rel_l2(g.get_real(gid[f]), o.get_real(oid[f]))
        print("  %-4s gpu-vs-f32 %.3e" % (f, err), flush=True)
        assert err <= 1e-4, f
    assert rel_l2(g.get_real(gid["x"]), spec["init"]["x"]) <= 5e-6  # R2C/C2R round trips only


def gray_scott_change(g, ids, F=0.045, k=0.065, Du=0.2, Dv=0.1):
    # terms in spec order: 0 u(-u v^2), 1 u(F), 2 u(-F u), 3 v(u v^2), 4 v(-(F + k) v)
    g.set_term_prefactors(1, [C.pres(F)])
    g.set_term_prefactors(2, [C.pres(-F)])
    g.set_term_prefactors(4, [C.pres(-(F + k))])
    g.set_implicits(ids["u"], [C.pres(-Du, 1)])
    g.set_implicits(ids["v"], [C.pres(-Dv, 1)])


def test_parameter_update_two_dynamic_fields():
    spec = spec_gray_scott(64, seed=50)
    runs = {}
    for name, path, change in (("fused", 0, True), ("general", 1, True), ("unchanged", 0, False)):
        g, ids = gpu(spec, path=path)
        g.step(12)
        if change:
            gray_scott_change(g, ids)
        g.step(12)
        runs[name] = np.concatenate([g.get_real(ids["u"]), g.get_real(ids["v"])])
    e, d = rel_l2(runs["fused"], runs["general"]), rel_l2(runs["fused"], runs["unchanged"])
    print("  fused-vs-general %.3e, effect of the change %.3e" % (e, d), flush=True)
    assert e <= 1e-5 and d > 1e-4
    # a batch member changed in place (same structure) equals the same system alone
    specs = [spec_gray_scott(64, seed=51), spec_gray_scott(128, seed=52), spec_cahn_hilliard((64, 64, 1), seed=53)]
    batch = C.Batch(graph_steps=4)
    members = []
    for sp in specs:
        g, ids = gpu(sp)
        batch.add(g)
        members.append((g, ids))
    batch.step(10)
    gray_scott_change(*members[1])
    batch.step(10)
    for i, sp in enumerate(specs):
        a, ia = gpu(sp)
        a.step(10)
        if i == 1:
            gray_scott_change(a, ia)
        a.step(10)
        for f in ia:
            assert np.array_equal(a.get_real(ia[f]), members[i][0].get_real(members[i][1][f])), (i, f)


def spec_noisy_constraints(dims=(64, 64, 1), seed=54):
    """dt phi + q^2 phi + q^4 phi = -q^2 mu + 0.2 iqx f + 0.3 c
       mu = phi^3 + noise;   f = noise;   c (1 + q^2) = 0.5 mu + noise"""
    n = dims[0] * dims[1] * dims[2]
    x0 = np.random.default_rng(seed).uniform(-0.1, 0.1, n).astype(np.float32)
    terms = [("phi", ["mu"], [(-1.0, 1)]), ("phi", ["f"], [(0.2, 0, 1)]), ("phi", ["c"], [(0.3,)]),
             ("mu", ["phi", "phi", "phi"], [(1.0,)]), ("c", ["mu"], [(0.5,)])]
    return {"dims": dims, "dt": 0.01,
            "fields": [("phi", True), ("mu", False), ("f", False), ("c", False)],
            "implicit": {"phi": [(1.0, 1), (-1.0, 2)], "c": [(1.0,), (1.0, 1)]}, "terms": terms,
            "noise": {"mu": ((0.001,), seed + 1), "f": ((0.002, 1), seed + 2), "c": ((0.001,), seed + 3)},
            "init": {"phi": x0}}


@pytest.mark.parametrize("noise_mode", [0, 1])
def test_noisy_constraints_fused_equal_general(noise_mode):
    spec = spec_noisy_constraints()
    runs = {}
    for path in (0, 1):
        g, ids = gpu(spec, path=path, noise_mode=noise_mode, profile=path == 0)
        if path == 0:
            assert g.fused_eligible
        for n in (7, 1, 12):
            g.step(n)
        if path == 0:
            assert_fused_stored_dynamic(g)
        runs[path] = (g, ids)
    for f in ("phi", "mu", "f", "c"):
        a, b = runs[0][0].get_real(runs[0][1][f]), runs[1][0].get_real(runs[1][1][f])
        e = rel_l2(a, b)
        print("  noise mode %d, %-3s fused-vs-general %.3e" % (noise_mode, f, e), flush=True)
        assert e <= 1e-5, f

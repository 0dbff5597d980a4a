"""GPU: the fused step of 3D systems (natural layout, four passes per step).

* Cahn-Hilliard in 3D on the fused path against the oracle's FP32 scheme
  (relative L2 <= 1e-4 after 100 steps), cubic and rectangular grids.
* The fused path against the general step program of the same system:
  relative L2 <= 1e-5 (same arithmetic per bin, other transform grouping),
  deterministic CH, KPZ with three derived gradient fields and spectral
  noise (both paths draw the same Philox stream per bin), and the derived
  fields written back after a run.
* Split runs and one step per call equal one run bit for bit (prologue
  equivalence, and runs that continue from the previous run's real fields).
* Mass conservation at the full 256^3 size (every CH term carries q^2), and
  the oracle's FP32 scheme at that size over a few steps.
* A noisy 3D system in real-space noise mode runs the general program.
"""
import numpy as np
import pytest

import cupss_amd as C
from oracle import oracle as O
from tests.helpers import build_system, rel_l2, spec_cahn_hilliard

pytestmark = pytest.mark.gpu


def gpu(spec, path=0, noise_mode=0, profile=False):
    g, ids = build_system(C.System, C.pres, spec)
    g.set_path(path)
    g.set_noise_mode(noise_mode)
    if profile:
        g.profile_kernel(True)
    return g, ids


def assert_fused_3d(g):
    name = g.kernel_time()[2]
    assert name.startswith("k_f3z<"), name


def spec_kpz3d(n, nu=1.0, lam=1.2, D=0.01, dt=0.005, seed=21):
    x0 = np.random.default_rng(seed).uniform(-0.05, 0.05, n ** 3).astype(np.float32)
    terms = [("h", ["iqxh", "iqxh"], [(0.5 * lam,)]), ("h", ["iqyh", "iqyh"], [(0.5 * lam,)]),
             ("h", ["iqzh", "iqzh"], [(0.5 * lam,)]),
             ("iqxh", ["h"], [(1.0, 0, 1)]), ("iqyh", ["h"], [(1.0, 0, 0, 1)]), ("iqzh", ["h"], [(1.0, 0, 0, 0, 1)])]
    return {"dims": (n, n, n), "dt": dt,
            "fields": [("h", True), ("iqxh", False), ("iqyh", False), ("iqzh", False)],
            "implicit": {"h": [(-nu, 1)]}, "terms": terms, "noise": {"h": ((2 * D,), seed)}, "init": {"h": x0}}


@pytest.mark.parametrize("dims", [(32, 32, 32), (64, 16, 32), (16, 32, 64)])
def test_cahn_hilliard_3d_fused_vs_oracle(dims):
    spec = spec_cahn_hilliard(dims, seed=5)
    print("fused 3D CH vs oracle", dims, flush=True)
    g, gid = gpu(spec, profile=True)
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(100)
    o.step(100)
    assert_fused_3d(g)
    err = rel_l2(g.get_real(gid["phi"]), o.get_real(oid["phi"]))
    print("  gpu-vs-f32 %.3e" % err, flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("n", [64, 128])
def test_cahn_hilliard_3d_fused_equals_general(n):
    spec = spec_cahn_hilliard((n, n, n), seed=8)
    print("fused vs general 3D CH", n, flush=True)
    a, ia = gpu(spec, path=0, profile=True)
    b, ib = gpu(spec, path=1)
    a.step(40)
    b.step(40)
    assert_fused_3d(a)
    err = rel_l2(a.get_real(ia["phi"]), b.get_real(ib["phi"]))
    print("  fused-vs-general %.3e" % err, flush=True)
    assert err <= 1e-5


def test_kpz_3d_derived_fields_and_spectral_noise():
    spec = spec_kpz3d(32)
    print("fused vs general 3D KPZ, spectral noise", flush=True)
    a, ia = gpu(spec, path=0, noise_mode=1, profile=True)
    b, ib = gpu(spec, path=1, noise_mode=1)
    a.step(30)
    b.step(30)
    assert_fused_3d(a)
    for name in ("h", "iqxh", "iqyh", "iqzh"):
        err = rel_l2(a.get_real(ia[name]), b.get_real(ib[name]))
        print("  %-5s fused-vs-general %.3e" % (name, err), flush=True)
        assert err <= 1e-5, name


def test_split_runs_equal_one_run_3d():
    for spec in (spec_cahn_hilliard((32, 16, 64), seed=4), spec_kpz3d(16)):
        key = "phi" if "phi" in dict(spec["fields"]) else "h"
        a, ia = gpu(spec, noise_mode=1)
        a.step(23)
        b, ib = gpu(spec, noise_mode=1)
        b.step(10)
        b.step(1)
        b.step(12)
        assert np.array_equal(a.get_real(ia[key]), b.get_real(ib[key]))
        # one step per call (the C++ evolver's advanceTime), reading the
        # state in between
        c, ic = gpu(spec, noise_mode=1)
        for i in range(23):
            c.step(1)
            if i in (4, 17):
                c.get_real(ic[key])
        assert np.array_equal(a.get_real(ia[key]), c.get_real(ic[key]))


def test_real_space_noise_3d_runs_general_program():
    spec = spec_kpz3d(16)
    a, ia = gpu(spec, path=0, noise_mode=0)
    b, ib = gpu(spec, path=1, noise_mode=0)
    a.step(12)
    b.step(12)
    assert np.array_equal(a.get_real(ia["h"]), b.get_real(ib["h"]))


def test_cahn_hilliard_256_cubed_mass_and_general():
    """The full 256^3 grid: the mean of phi (the q = 0 bin, untouched by
    every CH term) is conserved, and the fused path stays with the general
    program over 10 steps."""
    n = 256
    spec = spec_cahn_hilliard((n, n, n), seed=2)
    print("256^3 CH: fused and general, 10 steps", flush=True)
    a, ia = gpu(spec, path=0, profile=True)
    m0 = float(np.mean(a.get_real(ia["phi"]).astype(np.float64)))
    a.step(10)
    assert_fused_3d(a)
    x = a.get_real(ia["phi"])
    m1 = float(np.mean(x.astype(np.float64)))
    b, ib = gpu(spec, path=1)
    b.step(10)
    err = rel_l2(x, b.get_real(ib["phi"]))
    print("  mean %.9e -> %.9e, fused-vs-general %.3e" % (m0, m1, err), flush=True)
    assert abs(m1 - m0) <= 1e-6 * max(1.0, float(np.sqrt(np.mean(x.astype(np.float64) ** 2))))
    assert err <= 1e-5


def test_cahn_hilliard_256_cubed_against_oracle():
    """Configuration C3 at its full size: 256^3 Cahn-Hilliard on the fused
    3D path against the oracle's FP32 scheme, 3 steps (about 20 s of oracle
    work on one host core)."""
    spec = spec_cahn_hilliard((256, 256, 256), seed=2)
    print("256^3 CH: fused vs oracle, 3 steps", flush=True)
    g, gid = gpu(spec, profile=True)
    g.step(3)
    assert_fused_3d(g)
    got = g.get_real(gid["phi"])
    g.close()
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    o.step(3)
    err = rel_l2(got, o.get_real(oid["phi"]))
    print("  gpu-vs-f32 %.3e" % err, flush=True)
    assert err <= 1e-5


def spec_mixed_terms(dims, seed=13):
    """One dynamic field with every kind of term the fused steps take:
    a constant (SRC_CONST), a derived field (SRC_DERIVED: psi = iqx*phi),
    phi itself, and two product groups with different multipliers (one of
    them reading a second derived field chi = q^2*phi).
    dt phi + 0.5*q^2*phi = 0.3 + 0.4*psi - 0.2*q^2*phi^3 + 0.1*iqy*(phi*chi)"""
    n = dims[0] * dims[1] * dims[2]
    x0 = np.random.default_rng(seed).uniform(-0.2, 0.2, n).astype(np.float32)
    terms = [("phi", [], [(0.3,)]),
             ("phi", ["psi"], [(0.4,)]),
             ("phi", ["phi", "phi", "phi"], [(-0.2, 1)]),
             ("phi", ["phi", "chi"], [(0.1, 0, 0, 1)]),
             ("psi", ["phi"], [(1.0, 0, 1)]),
             ("chi", ["phi"], [(1.0, 1)])]
    return {"dims": dims, "dt": 0.01, "fields": [("phi", True), ("psi", False), ("chi", False)],
            "implicit": {"phi": [(-0.5, 1)]}, "terms": terms, "init": {"phi": x0}}


@pytest.mark.parametrize("dims", [(64, 32, 1), (16, 16, 32)])
def test_mixed_terms_fused_vs_oracle(dims):
    """The fused 2D and 3D steps on a system with constant, derived and
    grouped product terms, against the oracle's FP32 scheme."""
    spec = spec_mixed_terms(dims)
    g, gid = gpu(spec, profile=True)
    o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
    g.step(50)
    o.step(50)
    name = g.kernel_time()[2]
    assert name.startswith("k_f3z<" if dims[2] > 1 else "k_fcols<"), name
    for f in ("phi", "psi", "chi"):
        err = rel_l2(g.get_real(gid[f]), o.get_real(oid[f]))
        print("  %s %-4s gpu-vs-f32 %.3e" % (dims, f, err), flush=True)
        assert err <= 1e-4, f

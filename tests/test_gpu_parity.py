"""GPU parity: the HIP step program (through the C-ABI) against the oracle's
FP32 restatement of the reference scheme on the same seeded inputs, plus the
reference's own golden fixtures and closed forms.

Bar: relative L2 <= 1e-4 after 100 steps against the FP32
scheme; the FP64 run of the same scheme is reported beside it.
"""
import numpy as np
import pytest

import cupss_amd as C
from oracle import oracle as O
from tests.helpers import (build_system, droplet, f32, load_truth, qgrid, rel_l2, spec_cahn_hilliard,
                           spec_diffusion, spec_kpz, spec_model_h)

pytestmark = pytest.mark.gpu

TOL = 1e-4


def gpu(spec, path=0):
    g, ids = build_system(C.System, C.pres, spec)
    g.set_path(path)
    return g, ids


def orc(spec, precision="f32"):
    sp = dict(spec)
    sp["precision"] = precision
    return build_system(O.System, O.pres, sp)


def compare(spec, steps, fields=None, tol=TOL, report=None, path=0, expect_fused=None):
    print("parity", spec["dims"], steps, "path", "general" if path else "auto", flush=True)
    g, gid = gpu(spec, path)
    if expect_fused is not None:
        assert g.fused_eligible == expect_fused
    o, oid = orc(spec)
    o64, _ = orc(spec, "f64")
    g.step(steps)
    o.step(steps)
    o64.step(steps)
    errs = {}
    for name in (fields or gid.keys()):
        a = g.get_real(gid[name])
        b = o.get_real(oid[name])
        c = o64.get_real(oid[name])
        errs[name] = (rel_l2(a, b), rel_l2(a, c), rel_l2(b, c))
    for name, e in errs.items():
        print("  %-8s gpu-vs-f32 %.3e  gpu-vs-f64 %.3e  f32-vs-f64 %.3e" % ((name,) + e), flush=True)
    if report is not None:
        report.update(errs)
    for name, (e32, e64, _) in errs.items():
        assert e32 <= tol, (name, errs)
    return errs


def test_operator_fixtures_1d_3d():
    """tests/tests.cpp:191-422 on the GPU path (iqy/iqz checked properly)."""
    for dim in (1, 3):
        dims = (16, 16 if dim == 3 else 1, 16 if dim == 3 else 1)
        names = ["lapphi", "iqxphi"] + (["iqyphi", "iqzphi"] if dim == 3 else []) + ["invqphi"]
        ops = {"lapphi": (-1.0, 1), "iqxphi": (1.0, 0, 1), "iqyphi": (1.0, 0, 0, 1),
               "iqzphi": (1.0, 0, 0, 0, 1), "invqphi": (1.0, 0, 0, 0, 0, 1)}
        spec = {"dims": dims, "dt": 1.0, "fields": [("phi", True)] + [(n, False) for n in names],
                "implicit": {"phi": [(-1.0, 1)]},
                "terms": [("phi", ["iqxphi", "iqxphi"], [(1.0,)])] + [(n, ["phi"], [ops[n]]) for n in names],
                "init": {"phi": droplet(16, dims[1], dims[2], 0, 1, 2, 4, 8, 8 if dim == 3 else 0, 0)}}
        g, ids = gpu(spec)
        g.step(1)
        for n in names:
            got = g.get_real(ids[n])
            for dev in ("cpu", "gpu"):
                truth = load_truth("%s_%s_%dd" % (n, dev, dim), dim)
                assert np.max(np.abs(got - truth)) < 1e-4, (n, dev)


def test_diffusion_config1_closed_form_and_oracle():
    spec = spec_diffusion(256)
    errs = compare(spec, 100)
    # the HIP path has no per-step FFT round trip: it is closer to FP64
    assert errs["phi"][1] < 1e-5
    # full config: 1000 steps against the exact per-mode decay
    g, ids = gpu(spec)
    g.step(1000)
    x0 = spec["init"]["phi"].astype(np.float64).reshape(256, 256)
    qx, qy, _ = qgrid(256, 256)
    q2 = (qx ** 2 + qy ** 2)[0]
    dt, D = f32(0.01), 1.0
    exact = np.fft.ifft2(np.fft.fft2(x0) / (1 + dt * D * q2) ** 1000).real
    err = rel_l2(g.get_real(ids["phi"]), exact)
    print("  1000 steps vs exact decay: rel L2 %.3e" % err)
    # 1000 float divisions per mode: ~1000 * 2^-24 relative drift per mode
    assert err < 1e-4


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("n", [64, 128])
def test_cahn_hilliard_2d(n, path):
    compare(spec_cahn_hilliard((n, n, 1)), 100, path=path, expect_fused=True)


@pytest.mark.parametrize("path", [0, 1])
def test_cahn_hilliard_2d_1024_parity(path):
    """2D Cahn-Hilliard at 1024^2, 100 steps."""
    compare(spec_cahn_hilliard((1024, 1024, 1)), 100, path=path)


@pytest.mark.parametrize("dims", [(256, 32, 1), (32, 512, 1), (2048, 16, 1)])
def test_cahn_hilliard_2d_rectangular_fused(dims):
    compare(spec_cahn_hilliard(dims), 60, expect_fused=True)


def test_cahn_hilliard_3d():
    compare(spec_cahn_hilliard((32, 32, 32)), 100)


def test_cahn_hilliard_1d_and_anisotropic():
    compare(spec_cahn_hilliard((256, 1, 1)), 100)
    compare(spec_cahn_hilliard((64, 16, 1)), 100)
    compare(spec_cahn_hilliard((16, 8, 32)), 100)


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("lam", [0.0, 1.5])
def test_kpz_and_edwards_wilkinson_with_noise(lam, path):
    compare(spec_kpz(64, nu=0.8, lam=lam, D=0.01, dt=0.005), 100, path=path, expect_fused=True)


@pytest.mark.parametrize("n", [128, 512])
def test_kpz_fused_sizes(n):
    compare(spec_kpz(n, nu=1.2, lam=1.0, D=0.01, dt=0.005), 50, expect_fused=True)


@pytest.mark.parametrize("path", [0, 1])
def test_reinitialise_keeps_stale_dealiased_copy(path):
    """The reference re-runs prepareProblem without clearing real_dealiased
    (examples/06_kpz/06_kpz.cpp:39-44): the first step after a new initial
    condition reads the previous dealiased field."""
    spec = spec_cahn_hilliard((64, 64, 1))
    x1 = np.random.default_rng(9).uniform(-0.2, 0.2, 64 * 64).astype(np.float32)
    g, gid = gpu(spec, path)
    o, oid = orc(spec)
    g.step(7)
    o.step(7)
    g.set_real(gid["phi"], x1)
    o.set_real(oid["phi"], x1)
    g.prepare()
    o.prepare()
    g.step(13)
    o.step(13)
    err = rel_l2(g.get_real(gid["phi"]), o.get_real(oid["phi"]))
    print("  re-initialised CH: gpu-vs-f32 %.3e" % err)
    assert err < TOL


@pytest.mark.parametrize("path", [0, 1])
def test_model_h(path):
    """Stored constraint fields (sigma, v, w lag phi by one and two steps)
    on the fused path, and the general program."""
    compare(spec_model_h(64), 50, path=path, expect_fused=True if path == 0 else None)


def test_model_h_1024_against_oracle():
    """Model H at the served configuration's largest size, 1024^2 (fused
    path: stored stresses and velocities), 6 steps against the oracle's FP32
    scheme (about 13 s of oracle work on one host core): phi, the velocity
    and the stresses, whose one- and two-step lags are all in effect."""
    spec = spec_model_h(1024, seed=17)
    print("parity model H 1024^2, 6 steps", flush=True)
    g, gid = gpu(spec)
    assert g.fused_eligible
    g.step(6)
    got = {k: g.get_real(gid[k]) for k in ("phi", "vx", "vy", "sigxy")}
    g.close()
    o, oid = orc(spec)
    o.step(6)
    for k, a in got.items():
        err = rel_l2(a, o.get_real(oid[k]))
        print("  %-6s gpu-vs-f32 %.3e" % (k, err), flush=True)
        assert err <= TOL, k


def test_constraint_division_noise_and_nyquist():
    x0 = (np.random.default_rng(3).standard_normal(256) + 0.5).astype(np.float32)
    spec = {"dims": (16, 16, 1), "dt": 0.05,
            "fields": [("phi", True), ("psi", False), ("chi", False), ("g", False), ("nz", False)],
            "implicit": {"psi": [(1.0,), (1.0, 1)], "chi": [(2.0,)]},
            "terms": [("phi", [], [(0.0,)]), ("psi", ["phi"], [(1.0,)]), ("chi", ["phi"], [(1.0,)]),
                      ("g", ["phi"], [(1.0, 0, 1, 1), (2.0, 0, 2, 2)]), ("nz", [], [(0.0,)])],
            "noise": {"phi": ((0.6, 1), 5), "nz": ((0.2,), 6)},
            "init": {"phi": x0}}
    compare(spec, 20)


def test_product_cutoffs_1d():
    for p in (2, 3, 4):
        x0 = (0.5 * np.cos(2 * np.pi * 5 * np.arange(64) / 64)).astype(np.float32)
        spec = {"dims": (64, 1, 1), "dt": 0.01, "fields": [("phi", True)],
                "terms": [("phi", ["phi"] * p, [(1.0,)])], "init": {"phi": x0}}
        compare(spec, 30)

"""Pins the oracle (CPU restatement of the reference step) before it is used
as the checker for the HIP path.

Each test is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a wrong cut-off, a transposed axis) fails it:
  * FFT against an O(n^2) double DFT (1D/2D/3D, power-of-two and not);
  * Philox4x32-10 against the published Random123 known-answer vectors;
  * spectral operators (-q^2, iqx, iqy, iqz, 1/q) and the droplet initial
    condition against the reference's own golden files
    (tests/base_truths, reproduced in tests/golden/base_truths);
  * the implicit Euler update against the exact per-mode diffusion decay;
  * products, the dealias cut-off N/(p+1) and the step-0 zero dealiased
    field against the closed form of cos(k x)^p;
  * noise amplitude (dynamic and constraint) against sqrt(dt*A/dV)*xi;
  * the constraint-field division (q=0 bin left undivided) against hand
    derived values;
  * mass conservation of Cahn-Hilliard.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import PI_F, droplet, f32, load_truth, qgrid, rel_l2

P = O.pres


# --------------------------------------------------------------- FFT / DFT

@pytest.mark.parametrize("dims", [(16, 1, 1), (8, 4, 1), (4, 8, 2), (12, 1, 1), (6, 4, 1), (8, 8, 8)])
@pytest.mark.parametrize("direction", [-1, 1])
def test_fft_matches_dft(dims, direction):
    rng = np.random.default_rng(1)
    n = dims[0] * dims[1] * dims[2]
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = O.dft(x, dims, direction)
    assert rel_l2(O.fft(x, dims, direction, "f64"), ref) < 1e-13
    assert rel_l2(O.fft(x, dims, direction, "f32"), ref) < 2e-6
    # sign convention: forward is exp(-2 pi i k m / n), like numpy
    if dims[2] == 1:
        npref = (np.fft.fftn if direction < 0 else lambda a: np.fft.ifftn(a) * a.size)(
            x.reshape(dims[1], dims[0]))
        assert rel_l2(ref, npref.ravel()) < 1e-12


# ------------------------------------------------------------------ Philox

def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normal_stream_statistics():
    z = O.normals(12345, 7, 3, 1 << 20).astype(np.float64)
    assert abs(z.mean()) < 5e-3
    assert abs(z.var() - 1.0) < 5e-3
    # fourth moment of a Gaussian is 3
    assert abs((z ** 4).mean() - 3.0) < 0.05
    # distinct steps / streams / seeds give distinct streams
    assert not np.array_equal(O.normals(12345, 8, 3, 64), O.normals(12345, 7, 3, 64))
    assert not np.array_equal(O.normals(12345, 7, 4, 64), O.normals(12345, 7, 3, 64))
    assert not np.array_equal(O.normals(12346, 7, 3, 64), O.normals(12345, 7, 3, 64))
    # lag-1 correlation inside a Box-Muller pair and across groups is small
    assert abs(np.corrcoef(z[0::2], z[1::2])[0, 1]) < 5e-3


# ------------------------------------------------- reference golden files

@pytest.mark.parametrize("dim", [1, 2, 3])
def test_droplet_initial_condition_matches_fixture(dim):
    # tests/tests.cpp:64-123: initializeDroplet(f, 0, 1, 16/8, 4, 16/2, 16/2 (2D/3D), 0)
    sx, sy, sz = 16, (16 if dim > 1 else 1), (16 if dim > 2 else 1)
    x = droplet(sx, sy, sz, 0, 1, 2, 4, 8, 8 if dim > 1 else 0, 0)
    assert np.max(np.abs(x - load_truth("phi_%dd" % dim, dim))) < 1e-5


def _operator_system(dim, precision):
    """The system of tests/tests.cpp:191-422 with the parser's pres written
    out by hand (src/parser.cpp:547-589, 628-767)."""
    s = O.System(16, 16 if dim == 3 else 1, 16 if dim == 3 else 1, 1.0, 1.0, 1.0, 1.0, precision)
    phi = s.add_field(True)
    names = ["lapphi", "iqxphi"] + (["iqyphi", "iqzphi"] if dim == 3 else []) + ["invqphi"]
    ids = {n: s.add_field(False) for n in names}
    # "dt phi + q^2*phi = iqxphi^2": implicit q^2 negated (parser.cpp:697)
    s.add_implicit(phi, P(-1.0, q2n=1))
    s.add_term(phi, [ids["iqxphi"], ids["iqxphi"]], [P(1.0)])
    s.add_term(ids["lapphi"], [phi], [P(-1.0, q2n=1)])        # "lapphi = -q^2*phi"
    s.add_term(ids["iqxphi"], [phi], [P(1.0, iqx=1)])          # "iqxphi = iqx*phi"
    if dim == 3:
        s.add_term(ids["iqyphi"], [phi], [P(1.0, iqy=1)])
        s.add_term(ids["iqzphi"], [phi], [P(1.0, iqz=1)])
    s.add_term(ids["invqphi"], [phi], [P(1.0, invq=1)])        # "invqphi = 1/q*phi"
    s.set_real(phi, droplet(16, s.dims[1], s.dims[2], 0, 1, 2, 4, 8, 8 if dim == 3 else 0, 0))
    s.prepare()
    s.step(1)
    return s, ids


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("dim", [1, 3])
def test_operators_match_reference_fixtures(dim, precision):
    s, ids = _operator_system(dim, precision)
    for name, f in ids.items():
        got = s.get_real(f)
        for dev in ("cpu", "gpu"):
            truth = load_truth("%s_%s_%dd" % (name, dev, dim), dim)
            # the reference tests use EXPECT_NEAR(..., 1e-4)
            assert np.max(np.abs(got - truth)) < 1e-4, (name, dev)


def test_operator_fixtures_are_not_interchangeable():
    """The check above must be able to fail: iqx and iqy results differ."""
    s, ids = _operator_system(3, "f64")
    iqy = s.get_real(ids["iqyphi"])
    assert np.max(np.abs(iqy - load_truth("iqxphi_gpu_3d", 3))) > 1e-2
    assert np.max(np.abs(iqy - load_truth("iqyphi_gpu_3d", 3))) < 1e-4


# ------------------------------------------------------ closed-form checks

@pytest.mark.parametrize("dims", [(64, 1, 1), (32, 16, 1), (8, 8, 8)])
def test_diffusion_exact_mode_decay(dims):
    sx, sy, sz = dims
    D, dt, nsteps = f32(1.3), f32(0.05), 25
    s = O.System(sx, sy, sz, 1.0, 1.0, 1.0, dt, "f64")
    phi = s.add_field(True)
    s.add_implicit(phi, P(-D, q2n=1))          # "dt phi + D*q^2*phi = 0"
    s.add_term(phi, [], [P(0.0)])              # the "0" right-hand side
    x0 = np.random.default_rng(2).standard_normal(s.n).astype(np.float32)
    s.set_real(phi, x0)
    s.prepare()
    s.step(nsteps)
    qx, qy, qz = qgrid(sx, sy, sz)
    q2 = (qx ** 2 + qy ** 2 + qz ** 2).ravel()
    k0 = np.fft.fftn(x0.astype(np.float64).reshape(sz, sy, sx)).ravel()
    expected = np.fft.ifftn((k0 / (1 + dt * D * q2) ** nsteps).reshape(sz, sy, sx)).real.ravel()
    # the projection onto real fields is exact here (q^2 is even)
    assert rel_l2(s.get_real(phi), expected) < 1e-12
    s32 = O.System(sx, sy, sz, 1.0, 1.0, 1.0, dt, "f32")
    phi = s32.add_field(True)
    s32.add_implicit(phi, P(-D, q2n=1))
    s32.add_term(phi, [], [P(0.0)])
    s32.set_real(phi, x0)
    s32.prepare()
    s32.step(nsteps)
    # 25 steps of FFT round trips in float
    assert rel_l2(s32.get_real(phi), expected) < 1e-5


@pytest.mark.parametrize("p,m,kept", [(3, 8, True), (3, 9, False), (2, 10, True), (2, 11, False)])
def test_product_dealias_cutoff_and_step0(p, m, kept):
    """dt phi = phi^p in 1D, N=32: cutoff keeps |n| <= N/(p+1) (dealias_k,
    field_kernels.cu:241-254); real_dealiased is zero on step 0."""
    N, dt, A = 32, f32(0.01), 0.7
    s = O.System(N, 1, 1, 1.0, 1.0, 1.0, dt, "f64")
    phi = s.add_field(True)
    s.add_term(phi, [phi] * p, [P(1.0)])
    x = np.arange(N)
    x0 = (A * np.cos(2 * np.pi * m * x / N)).astype(np.float32)
    s.set_real(phi, x0)
    s.prepare()
    assert s.aliasing_order(phi) == p
    s.step(1)
    # step 0: the product reads the zero dealiased field -> no change
    assert np.max(np.abs(s.get_real(phi) - x0)) < 1e-12
    d = s.get_dealiased(phi)
    # x0 is float32: its rounding noise (~1e-8) spreads over all modes
    if kept:
        assert np.max(np.abs(d - x0)) < 1e-6
    else:
        assert np.max(np.abs(d)) < 1e-6
    s.step(1)
    c = x0.astype(np.float64)
    expected = c + dt * (c ** p if kept else 0.0)
    # cos^p has modes up to p*m < N so the product is not aliased here ...
    if kept and p * m >= N // 2:
        # ... unless p*m wraps: fold the frequencies like the DFT does
        expected = c + dt * np.fft.ifft(np.fft.fft(c ** p)).real
    assert np.max(np.abs(s.get_real(phi) - expected)) < 1e-6


def test_noise_amplitude_dynamic_and_constraint():
    sx, sy, dt, Dn, dx = 16, 8, f32(0.02), f32(0.3), 0.5
    seed = 99
    s = O.System(sx, sy, 1, dx, dx, 1.0, dt, "f64")
    phi = s.add_field(True)
    psi = s.add_field(False)
    chi = s.add_field(True)
    s.add_term(phi, [], [P(0.0)])              # "dt phi = 0"
    s.add_term(psi, [], [P(0.0)])              # "psi = 0"
    s.add_term(chi, [], [P(0.0)])              # "dt chi = 0"
    s.set_noise(phi, P(2 * Dn), seed)          # addNoise("phi", "2*D")
    s.set_noise(psi, P(2 * Dn), seed + 1)
    s.set_noise(chi, P(2 * Dn, q2n=1), seed + 2)   # conserved noise "2*D*q^2"
    s.prepare()
    s.step(1)
    dV = dx * dx * 1.0
    amp = np.sqrt(dt * 2 * Dn / dV)
    xi_phi = O.normals(seed, 0, phi, s.n).astype(np.float64)
    xi_psi = O.normals(seed + 1, 0, psi, s.n).astype(np.float64)
    xi_chi = O.normals(seed + 2, 0, chi, s.n).astype(np.float64)
    assert rel_l2(s.get_real(phi), amp * xi_phi) < 1e-6
    # constraint fields carry dt^(-1/2) (setNotDynamic_k, field_kernels.cu:153-166)
    assert rel_l2(s.get_real(psi), amp / np.sqrt(dt) * xi_psi) < 1e-6
    qx, qy, qz = qgrid(sx, sy, 1, dx, dx)
    absq = np.sqrt(qx ** 2 + qy ** 2).ravel()
    exp_chi = np.fft.ifft2((amp * absq * np.fft.fft2(xi_chi.reshape(sy, sx)).ravel()).reshape(sy, sx)).real
    assert rel_l2(s.get_real(chi), exp_chi.ravel()) < 1e-6
    # the second step draws a fresh deviate (counter includes the step)
    s.step(1)
    xi1 = O.normals(seed, 1, phi, s.n).astype(np.float64)
    assert rel_l2(s.get_real(phi), amp * (xi_phi + xi1)) < 1e-6


@pytest.mark.parametrize("lhs", ["psi*(1+q^2)", "2*psi"])
def test_constraint_division_leaves_q0_undivided(lhs):
    sx, sy = 16, 16
    s = O.System(sx, sy, 1, 1.0, 1.0, 1.0, 0.1, "f64")
    phi = s.add_field(True)
    psi = s.add_field(False)
    s.add_term(phi, [], [P(0.0)])
    if lhs == "psi*(1+q^2)":
        s.add_implicit(psi, P(1.0))
        s.add_implicit(psi, P(1.0, q2n=1))
    else:
        s.add_implicit(psi, P(2.0))
    s.add_term(psi, [phi], [P(1.0)])           # "... = phi"
    x0 = (np.random.default_rng(3).standard_normal(s.n) + 0.5).astype(np.float32)
    s.set_real(phi, x0)
    s.prepare()
    s.step(1)
    qx, qy, _ = qgrid(sx, sy)
    q2 = (qx ** 2 + qy ** 2).ravel()
    den = (1 + q2) if lhs == "psi*(1+q^2)" else np.full(s.n, 2.0)
    den[0] = 1.0                               # idx 0 is not divided (field_kernels.cu:167)
    k0 = np.fft.fft2(x0.astype(np.float64).reshape(sy, sx)).ravel()
    expected = np.fft.ifft2((k0 / den).reshape(sy, sx)).real.ravel()
    assert rel_l2(s.get_real(psi), expected) < 1e-12
    if lhs == "2*psi":
        assert abs(s.get_real(psi).mean() - x0.mean()) < 1e-6


def test_nyquist_projection_of_odd_multipliers():
    """The real-part projection removes iqx at the Nyquist index: the result
    equals the R2C rule 'odd power of q_d at index N_d/2 -> 0'."""
    sx, sy = 8, 8
    s = O.System(sx, sy, 1, 1.0, 1.0, 1.0, 0.1, "f64")
    phi = s.add_field(True)
    g = s.add_field(False)
    s.add_term(phi, [], [P(0.0)])
    s.add_term(g, [phi], [P(1.0, iqx=1, iqy=1), P(2.0, iqx=2, iqy=2)])
    x0 = np.random.default_rng(4).standard_normal(s.n).astype(np.float32)
    s.set_real(phi, x0)
    s.prepare()
    s.step(1)
    qx, qy, _ = qgrid(sx, sy)
    qx, qy = qx[0], qy[0]
    m = -(qx * qy) + 2 * (qx * qy) ** 2       # (iqx)(iqy) = -qx qy ; (iqx)^2 (iqy)^2 = qx^2 qy^2
    nyx = np.zeros_like(qx, bool)
    nyx[:, sx // 2] = True
    nyy = np.zeros_like(qx, bool)
    nyy[sy // 2, :] = True
    odd = (nyx ^ nyy)                          # total odd power over Nyquist axes
    m_r2c = np.where(odd, 2 * (qx * qy) ** 2, m)
    k0 = np.fft.fft2(x0.astype(np.float64).reshape(sy, sx))
    expected = np.fft.ifft2(m_r2c * k0)
    assert np.max(np.abs(expected.imag)) < 1e-10
    assert rel_l2(s.get_real(g), expected.real.ravel()) < 1e-12


def test_cahn_hilliard_conserves_mass_and_cpu_bug_switch_differs():
    sx = sy = 32
    def run(bug):
        O.lib().orc_set_ref_cpu_dealias_bug(1 if bug else 0)
        try:
            s = O.System(sx, sy, 1, 1.0, 1.0, 1.0, 0.01, "f64")
            phi = s.add_field(True)
            s.add_implicit(phi, P(1.0, q2n=1))     # "dt phi + (a*q^2 + k*q^4)*phi = -b*q^2*phi^3", a=-1
            s.add_implicit(phi, P(-1.0, q2n=2))
            s.add_term(phi, [phi] * 3, [P(-1.0, q2n=1)])
            x0 = (0.1 * np.random.default_rng(5).uniform(-1, 1, s.n) + 0.05).astype(np.float32)
            s.set_real(phi, x0)
            s.prepare()
            s.step(50)
            return x0, s.get_real(phi)
        finally:
            O.lib().orc_set_ref_cpu_dealias_bug(0)
    x0, good = run(False)
    assert abs(good.mean() - x0.astype(np.float64).mean()) < 1e-12
    _, bad = run(True)
    assert rel_l2(bad, good) > 1e-4

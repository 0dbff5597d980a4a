"""Pins the oracle where the reference's fixtures leave it open: constraint
chains with the two-phase lag, products of two different fields, the model-H
Stokes velocity, and the field callbacks.

Each check is a closed form or an equation of the physics, evaluated here
with numpy in double precision from the oracle's state, never the oracle's
own code path retyped:
  * a two-constraint chain psi1 = c1(q) phi, psi2 = c2 psi1 feeding back into
    a diffusing phi obeys a per-mode delay recurrence fixed by the phase order
    A-D of src/evolver.cpp:199-226 (constraint terms from the pre-step state,
    so psi2 lags phi by two steps);
  * a product of two different fields, cos(2 pi p x/N) * (cos(2 pi r y/N) +
    B cos(2 pi r2 y/N)), goes through each factor's N/3 dealias mask
    (term_kernels.cu:48-70, field_kernels.cu:229-256): the r2 part is cut;
  * model H (examples/04_model_h/modelh_explicit_velocity.cpp:14-39): the
    velocity is incompressible, q.v = 0, and solves Stokes,
    eta q^2 v = (I - q q^T / q^2) (i q . sigma) from the previous step's
    stress; the stress is the dealiased product of the previous gradients;
    the gradients are i q phi of the previous step; phi takes the
    semi-implicit Euler step of the model-H equation;
  * callbacks (src/field.cpp:48-58, 70-86): a Fourier low-pass keeps the
    spectrum inside its band, identity callbacks change nothing, a mirror
    callback leaves mirror-symmetric fields.
Each structural check is shown to fail with a mistake planted in the oracle
(dynamic phases before constraint phases; products of undealiased factors).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import f32, lowpass, qgrid, rel_l2

P = O.pres


@pytest.fixture
def plant():
    """Sets a planted mistake for one test and always clears it."""
    def set_(code):
        O.set_plant(code)
    yield set_
    O.set_plant(O.PLANT_NONE)


def _spec2(a, sx, sy):
    return np.fft.fft2(np.asarray(a, np.float64).reshape(sy, sx))


def _real2(k):
    return np.fft.ifft2(k).real


def _mask(sx, sy, p):
    """dealias_k's keep-mask for aliasing order p: |n_d| <= N_d/(p+1)."""
    nx = np.arange(sx)
    nx = np.where(nx > sx // 2, nx - sx, nx)
    ny = np.arange(sy)
    ny = np.where(ny > sy // 2, ny - sy, ny)
    return (np.abs(nx)[None, :] <= sx // (p + 1)) & (np.abs(ny)[:, None] <= sy // (p + 1))


def _interior(sx, sy):
    """Bins off the Nyquist row and column (where the real projection
    removes odd powers of q)."""
    m = np.ones((sy, sx), bool)
    m[:, sx // 2] = False
    m[sy // 2, :] = False
    return m


# ------------------------------------------------------ constraint chain lag

def _chain(nsteps, sx=16, sy=16, D=1.0, dt=0.05, c2=0.5, g=1.0):
    s = O.System(sx, sy, 1, 1.0, 1.0, 1.0, f32(dt), "f64")
    phi = s.add_field(True)
    psi1 = s.add_field(False)
    psi2 = s.add_field(False)
    s.add_implicit(phi, P(-D, q2n=1))               # "dt phi + D*q^2*phi = -g*psi2"
    s.add_term(phi, [psi2], [P(-g)])
    s.add_term(psi1, [phi], [P(-1.0, q2n=1)])       # "psi1 = -q^2*phi"
    s.add_term(psi2, [psi1], [P(c2)])               # "psi2 = c2*psi1"
    x0 = np.random.default_rng(11).standard_normal(s.n).astype(np.float32)
    s.set_real(phi, x0)
    s.prepare()
    s.step(nsteps)
    return s, x0, (phi, psi1, psi2)


def _chain_expected(x0, nsteps, sx=16, sy=16, D=1.0, dt=0.05, c2=0.5, g=1.0):
    """Per mode: psi1 after step m = c1 phi_m, psi2 after step m = c2 psi1
    before step m = c1 c2 phi_{m-1}; phi_{m+1} = (phi_m - dt g c1 c2
    phi_{m-1}) / (1 + dt D q^2), phi_{-1} = 0, with c1 = -q^2."""
    qx, qy, _ = qgrid(sx, sy)
    q2 = (qx ** 2 + qy ** 2)[0]
    dt = f32(dt)
    c1 = -q2
    hist = [np.zeros((sy, sx), complex), _spec2(x0, sx, sy)]      # phi_{-1}, phi_0
    for _ in range(nsteps):
        hist.append((hist[-1] - dt * g * c1 * c2 * hist[-2]) / (1 + dt * D * q2))
    phi_n, phi_n1, phi_n2 = hist[-1], hist[-2], hist[-3]
    return _real2(phi_n), _real2(c1 * phi_n1), _real2(c1 * c2 * phi_n2)


@pytest.mark.parametrize("nsteps", [1, 2, 7])
def test_constraint_chain_two_step_lag(nsteps):
    s, x0, (phi, psi1, psi2) = _chain(nsteps)
    e_phi, e_psi1, e_psi2 = _chain_expected(x0, nsteps)
    assert rel_l2(s.get_real(phi), e_phi.ravel()) < 1e-12
    assert rel_l2(s.get_real(psi1), e_psi1.ravel()) < 1e-12
    if nsteps >= 2:
        assert rel_l2(s.get_real(psi2), e_psi2.ravel()) < 1e-12
    else:
        assert np.max(np.abs(s.get_real(psi2))) == 0.0   # psi1 was zero before step 0


def test_constraint_chain_check_fails_with_phases_swapped(plant):
    plant(O.PLANT_PHASES)
    s, x0, (phi, psi1, psi2) = _chain(7)
    e_phi, e_psi1, _ = _chain_expected(x0, 7)
    assert rel_l2(s.get_real(phi), e_phi.ravel()) > 1e-3
    assert rel_l2(s.get_real(psi1), e_psi1.ravel()) > 1e-3


# ------------------------------------------------ product of two different fields

def _two_field_product(N=32, p=4, r=7, r2=13, B=0.8, c=0.6, dt=0.1, nsteps=3):
    s = O.System(N, N, 1, 1.0, 1.0, 1.0, f32(dt), "f64")
    u = s.add_field(True)
    f = s.add_field(False)
    g = s.add_field(False)
    s.add_term(u, [f, g], [P(c)])                    # "dt u = c*f*g"
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    fx = np.cos(2 * np.pi * p * i / N)
    gy = np.cos(2 * np.pi * r * j / N) + B * np.cos(2 * np.pi * r2 * j / N)
    u0 = np.random.default_rng(12).uniform(-1, 1, (N, N))
    s.set_real(u, u0.astype(np.float32).ravel())
    s.set_real(f, fx.astype(np.float32).ravel())
    s.set_real(g, gy.astype(np.float32).ravel())
    s.prepare()
    s.step(nsteps)
    return s, (u, f, g), u0.astype(np.float32).astype(np.float64), fx, gy, (N, p, r, r2, B, c, f32(dt), nsteps)


def test_two_field_product_through_dealias_mask():
    s, (u, f, g), u0, fx, gy, (N, p, r, r2, B, c, dt, nsteps) = _two_field_product()
    assert s.aliasing_order(f) == 2 and s.aliasing_order(g) == 2
    assert r <= N // 3 < r2 and p <= N // 3
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    # f and g are constraint fields without terms: they keep their values,
    # and their dealiased copies are fresh in phase C of every step; the
    # r2 mode of g is above N/3 and is cut from the product
    expected = u0 + nsteps * dt * c * np.cos(2 * np.pi * p * i / N) * np.cos(2 * np.pi * r * j / N)
    # inputs are float32-rounded (~3e-8 relative), which bounds the match
    assert np.max(np.abs(s.get_real(u) - expected.ravel())) < 1e-6
    assert np.max(np.abs(s.get_dealiased(g) - np.cos(2 * np.pi * r * j / N).ravel())) < 1e-6


def test_two_field_product_check_fails_undealiased(plant):
    plant(O.PLANT_UNDEALIAS)
    s, (u, f, g), u0, fx, gy, (N, p, r, r2, B, c, dt, nsteps) = _two_field_product()
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    expected = u0 + nsteps * dt * c * np.cos(2 * np.pi * p * i / N) * np.cos(2 * np.pi * r * j / N)
    assert np.max(np.abs(s.get_real(u) - expected.ravel())) > 1e-2


# ------------------------------------------------------------------ model H

def _model_h(N=32, steps_before=3, a=-1.0, b=1.0, k=4.0, eta=1.0, ka=-4.0, dt=0.01):
    from cupss_amd.systems import build, spec_model_h
    spec = spec_model_h(N, a=a, b=b, k=k, eta=eta, ka=ka, dt=dt)
    spec["precision"] = "f64"
    s, ids = build(spec, O.System, O.pres)
    s.step(steps_before)
    prev = {n: (s.get_real(f).reshape(N, N), s.get_comp(f).reshape(N, N)) for n, f in ids.items()}
    s.step(1)
    cur = {n: (s.get_real(f).reshape(N, N), s.get_comp(f).reshape(N, N)) for n, f in ids.items()}
    return prev, cur, dict(N=N, a=a, b=b, k=k, eta=eta, ka=ka, dt=f32(dt))


def _modelh_checks(prev, cur, par):
    """Relative errors of each model-H relation (step m-1 -> m)."""
    N = par["N"]
    qx, qy, _ = qgrid(N, N)
    qx, qy = qx[0], qy[0]
    q2 = qx ** 2 + qy ** 2
    inner = _interior(N, N)
    err = {}
    # gradients from the previous phi (constraint phase A reads the pre-step state)
    phik = _spec2(prev["phi"][0], N, N)
    err["iqxphi"] = rel_l2(cur["iqxphi"][0], _real2(1j * qx * phik))
    err["iqyphi"] = rel_l2(cur["iqyphi"][0], _real2(1j * qy * phik))
    # stress from the previous gradients, each through its N/3 mask
    m3 = _mask(N, N, 2)
    gx = _real2(m3 * _spec2(prev["iqxphi"][0], N, N))
    gy = _real2(m3 * _spec2(prev["iqyphi"][0], N, N))
    ka = f32(par["ka"])
    err["sigxx"] = rel_l2(cur["sigxx"][0], -0.5 * ka * gx * gx + 0.5 * ka * gy * gy)
    err["sigxy"] = rel_l2(cur["sigxy"][0], -ka * gx * gy)
    # Stokes velocity from the previous stress (sigma_yy = -sigma_xx)
    sxx, sxy = prev["sigxx"][1], prev["sigxy"][1]
    fx = 1j * qx * sxx + 1j * qy * sxy
    fy = 1j * qx * sxy - 1j * qy * sxx
    q2s = np.where(q2 > 0, q2, 1.0)
    qf = (qx * fx + qy * fy) / q2s
    vx = np.where(q2 > 0, (fx - qx * qf) / (par["eta"] * q2s), 0)
    vy = np.where(q2 > 0, (fy - qy * qf) / (par["eta"] * q2s), 0)
    vxk, vyk = cur["vx"][1], cur["vy"][1]
    err["vx"] = rel_l2(vxk[inner], vx[inner])
    err["vy"] = rel_l2(vyk[inner], vy[inner])
    div = np.abs(qx * vxk + qy * vyk)[inner]
    err["div"] = float(div.max() / (np.sqrt(q2) * np.sqrt(np.abs(vxk) ** 2 + np.abs(vyk) ** 2))[inner].max())
    # phi: semi-implicit Euler of dt phi + (a q^2 + k q^4) phi = -b q^2 phi^3 - v.grad phi
    m4 = _mask(N, N, 3)
    phid = _real2(m4 * phik)
    vxd = _real2(m3 * vxk)
    vyd = _real2(m3 * vyk)
    gxd = _real2(m3 * _spec2(cur["iqxphi"][0], N, N))
    gyd = _real2(m3 * _spec2(cur["iqyphi"][0], N, N))
    rhs = -f32(par["b"]) * q2 * _spec2(phid ** 3, N, N) - _spec2(vxd * gxd + vyd * gyd, N, N)
    dt = par["dt"]
    den = 1 + dt * (f32(par["a"]) * q2 + f32(par["k"]) * q2 * q2)
    err["phi"] = rel_l2(cur["phi"][0], _real2((phik + dt * rhs) / den))
    return err


def test_model_h_stokes_incompressible_and_lags():
    prev, cur, par = _model_h()
    err = _modelh_checks(prev, cur, par)
    for name, e in err.items():
        assert e < 1e-10, (name, e)


def test_model_h_checks_fail_with_planted_mistakes(plant):
    plant(O.PLANT_PHASES)
    err = _modelh_checks(*_model_h())
    assert err["iqxphi"] > 1e-4 and err["iqyphi"] > 1e-4
    plant(O.PLANT_UNDEALIAS)
    err = _modelh_checks(*_model_h())
    assert err["sigxx"] > 1e-4 and err["sigxy"] > 1e-4 and err["phi"] > 1e-6
    # Stokes and incompressibility hold whatever feeds the stress
    assert err["vx"] < 1e-10 and err["div"] < 1e-10


# ---------------------------------------------------------------- callbacks

def _ch(N, hooks=None, steps=20, precision="f64", x0=None):
    s = O.System(N, N, 1, 1.0, 1.0, 1.0, f32(0.01), precision)
    phi = s.add_field(True)
    s.add_implicit(phi, P(1.0, q2n=1))          # a = -1
    s.add_implicit(phi, P(-1.0, q2n=2))         # k = 1
    s.add_term(phi, [phi] * 3, [P(-1.0, q2n=1)])
    if x0 is None:
        x0 = np.random.default_rng(13).uniform(-0.1, 0.1, N * N).astype(np.float32)
    s.set_real(phi, x0)
    if hooks:
        s.set_hooks(phi, **hooks)
    s.prepare()
    s.step(steps)
    return s, phi


def test_fourier_callback_keeps_band_and_mass():
    N, cut = 32, 5
    s, phi = _ch(N, {"fourier": lowpass(cut)})
    c = s.get_comp(phi).reshape(N, N)
    n = np.arange(N)
    n = np.abs(np.where(n > N // 2, n - N, n))
    keep = (n[None, :] <= cut) & (n[:, None] <= cut)
    assert np.max(np.abs(c[~keep])) < 1e-12 * np.max(np.abs(c))
    ref, rphi = _ch(N)
    assert abs(c[0, 0] - ref.get_comp(rphi)[0]) < 1e-10     # the q = 0 bin is kept: mass conserved
    assert rel_l2(s.get_real(phi), ref.get_real(rphi)) > 1e-3


def test_identity_callbacks_change_nothing():
    N = 16
    s, phi = _ch(N, {"fourier": lambda z: z, "real": lambda z: None}, precision="f32")
    ref, rphi = _ch(N, precision="f32")
    assert np.array_equal(s.get_real(phi), ref.get_real(rphi))


def test_mirror_callback_leaves_symmetric_field():
    N = 16

    def mirror(z):
        out = z.copy()
        _, sy, sx = z.shape
        for j in range(sy):
            for i in range(sx):
                mi = i if i <= sx // 2 else sx - i
                mj = j if j <= sy // 2 else sy - j
                out[0, j, i] = z[0, mj, mi].real
        return out

    s, phi = _ch(N, {"real": mirror}, steps=5)
    r = s.get_real(phi).reshape(N, N)
    assert np.max(np.abs(r - r[:, (-np.arange(N)) % N])) < 1e-12
    assert np.max(np.abs(r - r[(-np.arange(N)) % N, :])) < 1e-12

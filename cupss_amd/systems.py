"""System specs of the benchmark and test workloads, and a builder.

A spec is what the reference's parser makes of an equation set
(src/parser.cpp:628-767, createField / createTerm / addNoise calls),
written out by hand: fields (name, dynamic), implicit prefactors per field,
terms (field, factor fields, prefactors) and noise, with pres given as
tuples (pre, q2n, iqx, iqy, iqz, invq) (inc/cupss/defines.h:31-39), plus
the initial real fields.  `build` turns a spec into a prepared system of any
class with the System interface of this package.
"""
import numpy as np


def build(spec, cls=None, pres_fn=None):
    """Prepared system of `cls` (default: the HIP System) and {field name: id}."""
    if cls is None:
        from cupss_amd import System as cls, pres as pres_fn  # noqa: N813
    sx, sy, sz = spec["dims"]
    dx = spec.get("dx", 1.0)
    kw = {}
    if "precision" in spec and cls.__module__.startswith("oracle"):
        kw["precision"] = spec["precision"]
    s = cls(sx, sy, sz, dx, dx if sy > 1 else 1.0, dx if sz > 1 else 1.0, spec["dt"], **kw)
    ids = {}
    for name, dyn in spec["fields"]:
        ids[name] = s.add_field(dyn)
    for name, plist in spec.get("implicit", {}).items():
        for p in plist:
            s.add_implicit(ids[name], pres_fn(*p))
    for name, factors, plist in spec.get("terms", []):
        s.add_term(ids[name], [ids[f] for f in factors], [pres_fn(*p) for p in plist])
    for name, (p, seed) in spec.get("noise", {}).items():
        s.set_noise(ids[name], pres_fn(*p), seed)
    for name, x in spec.get("init", {}).items():
        s.set_real(ids[name], x)
    s.prepare()
    return s, ids


def spec_diffusion(n=256, steps_dt=0.01, D=1.0, seed=0, sigma=10.0):
    """2D diffusion (reference example 01): dt phi + D*q^2*phi = 0, Gaussian of width 10 + U(+-1e-3)."""
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    c = n // 2
    x0 = np.exp(-((i - c) ** 2 + (j - c) ** 2) / (2 * sigma ** 2))
    x0 = x0 + np.random.default_rng(seed).uniform(-1e-3, 1e-3, x0.shape)
    return {"dims": (n, n, 1), "dt": steps_dt, "fields": [("phi", True)],
            "implicit": {"phi": [(-D, 1)]}, "terms": [("phi", [], [(0.0,)])],
            "init": {"phi": x0.astype(np.float32).ravel()}}


def spec_cahn_hilliard(dims, a=-1.0, b=1.0, k=1.0, dt=0.01, seed=1):
    """Cahn-Hilliard (model B), 2D or 3D: dt phi + (a*q^2 + k*q^4)*phi = -b*q^2*phi^3."""
    n = dims[0] * dims[1] * dims[2]
    x0 = np.random.default_rng(seed).uniform(-0.1, 0.1, n).astype(np.float32)
    return {"dims": dims, "dt": dt, "fields": [("phi", True)],
            "implicit": {"phi": [(-a, 1), (-k, 2)]},
            "terms": [("phi", ["phi", "phi", "phi"], [(-b, 1)])],
            "init": {"phi": x0}}


def spec_kpz(n, nu=1.0, lam=1.0, D=0.01, dt=0.005, seed=7, h0=None):
    """KPZ / Edwards-Wilkinson instance: dt h + nu*q^2*h = 0.5*lambda*iqxh^2 + 0.5*lambda*iqyh^2,
    iqxh = iqx*h, iqyh = iqy*h, addNoise(h, "2*D") (lambda = 0: Edwards-Wilkinson)."""
    terms = [("iqxh", ["h"], [(1.0, 0, 1)]), ("iqyh", ["h"], [(1.0, 0, 0, 1)])]
    if lam != 0.0:
        terms = [("h", ["iqxh", "iqxh"], [(0.5 * lam,)]), ("h", ["iqyh", "iqyh"], [(0.5 * lam,)])] + terms
    else:
        terms = [("h", [], [(0.0,)])] + terms
    init = {"h": np.zeros(n * n, np.float32) if h0 is None else h0}
    return {"dims": (n, n, 1), "dt": dt, "fields": [("h", True), ("iqxh", False), ("iqyh", False)],
            "implicit": {"h": [(-nu, 1)]}, "terms": terms, "noise": {"h": ((2 * D,), seed)}, "init": init}


def spec_model_h(n, a=-1.0, b=1.0, k=4.0, eta=1.0, friction=0.0, ka=-4.0, dt=0.01, seed=3):
    """Model H (Cahn-Hilliard advected by the Stokes velocity of the phi stress):
    examples/04_model_h/modelh_explicit_velocity.cpp:14-39."""
    x0 = np.random.default_rng(seed).uniform(-0.1, 0.1, n * n).astype(np.float32)
    terms = [
        ("phi", ["phi", "phi", "phi"], [(-b, 1)]),
        ("phi", ["iqxphi", "vx"], [(-1.0,)]),
        ("phi", ["iqyphi", "vy"], [(-1.0,)]),
        ("iqxphi", ["phi"], [(1.0, 0, 1)]),
        ("iqyphi", ["phi"], [(1.0, 0, 0, 1)]),
        ("sigxx", ["iqxphi", "iqxphi"], [(-0.5 * ka,)]),
        ("sigxx", ["iqyphi", "iqyphi"], [(0.5 * ka,)]),
        ("sigxy", ["iqxphi", "iqyphi"], [(-ka,)]),
        ("vx", ["sigxx"], [(1.0, 0, 1), (1.0, 0, 3, 0, 0, 2), (-1.0, 0, 1, 2, 0, 2)]),
        ("vx", ["sigxy"], [(1.0, 0, 0, 1), (1.0, 0, 2, 1, 0, 2), (1.0, 0, 2, 1, 0, 2)]),
        ("vy", ["sigxy"], [(1.0, 0, 1), (1.0, 0, 1, 2, 0, 2), (1.0, 0, 1, 2, 0, 2)]),
        ("vy", ["sigxx"], [(-1.0, 0, 0, 1), (-1.0, 0, 0, 3, 0, 2), (1.0, 0, 2, 1, 0, 2)]),
        ("w", ["vy"], [(0.5, 0, 1)]),
        ("w", ["vx"], [(-0.5, 0, 0, 1)]),
    ]
    fields = [("phi", True), ("iqxphi", False), ("iqyphi", False), ("sigxx", False), ("sigxy", False),
              ("vx", False), ("vy", False), ("w", False)]
    imp = {"phi": [(-a, 1), (-k, 2)], "vx": [(friction,), (eta, 1)], "vy": [(friction,), (eta, 1)]}
    return {"dims": (n, n, 1), "dt": dt, "fields": fields, "implicit": imp, "terms": terms,
            "init": {"phi": x0}}


def spec_gray_scott(n=64, seed=31, F=0.035, k=0.065, dt=0.5):
    """Gray-Scott reaction-diffusion, two dynamic fields:
    dt u + 0.16 q^2 u = -u v^2 + F (1 - u);  dt v + 0.08 q^2 v = u v^2 - (F + k) v,
    from u = 1, v = 0 with a square of u = 1/2, v = 1/4 in the middle plus
    U(-0.02, 0.02) on both."""
    rng = np.random.default_rng(seed)
    u = np.ones((n, n), np.float32)
    v = np.zeros((n, n), np.float32)
    c = slice(n // 2 - n // 8, n // 2 + n // 8)
    u[c, c] = 0.5
    v[c, c] = 0.25
    u += rng.uniform(-0.02, 0.02, (n, n)).astype(np.float32)
    v += rng.uniform(-0.02, 0.02, (n, n)).astype(np.float32)
    terms = [("u", ["u", "v", "v"], [(-1.0,)]), ("u", [], [(F,)]), ("u", ["u"], [(-F,)]),
             ("v", ["u", "v", "v"], [(1.0,)]), ("v", ["v"], [(-(F + k),)])]
    return {"dims": (n, n, 1), "dt": dt, "fields": [("u", True), ("v", True)],
            "implicit": {"u": [(-0.16, 1)], "v": [(-0.08, 1)]}, "terms": terms,
            "init": {"u": u.ravel(), "v": v.ravel()}}

"""Shared test helpers: reference fixture loader, q grid, initial conditions.

The q grid follows the reference convention q_d(i) = (i < (s+1)/2 ? i : i-s)
* 2*PI/(d*s) with PI = 3.1415926535f (src/field_init.cpp:207-213,
inc/cupss/defines.h:54).
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI_F = float(np.float32(3.1415926535))


def load_truth(name, dim):
    """Value column of a reference CSV fixture (tests/tests.cpp:10-62)."""
    vals = []
    with open(os.path.join(GOLDEN, "base_truths", name)) as fh:
        for line in fh:
            line = line.strip()
            if not line:
                continue
            vals.append(float(line.split(",")[dim]))
    return np.array(vals)


def qgrid(sx, sy=1, sz=1, dx=1.0, dy=1.0, dz=1.0):
    """(qx, qy, qz) arrays of shape (sz, sy, sx) in float64."""
    def ax(s, d):
        n = np.arange(s)
        n = np.where(n < (s + 1) // 2, n, n - s).astype(np.float64)
        # stepq is computed in float: 2.0f*PI/(d*(float)s)
        step = np.float32(np.float32(2.0) * np.float32(PI_F)) / (np.float32(d) * np.float32(s))
        return n * float(np.float32(step))
    qz, qy, qx = np.meshgrid(ax(sz, dz), ax(sy, dy), ax(sx, dx), indexing="ij")
    return qx, qy, qz


def droplet(sx, sy, sz, v1, v2, radius, xi, px, py, pz):
    """evolver::initializeDroplet (src/initializers.cpp:142-177)."""
    k, j, i = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    xd = (i - px % sx).astype(np.float32)
    yd = (j - py % sy).astype(np.float32)
    zd = (k - pz % sz).astype(np.float32)
    rc = np.sqrt(xd * xd + yd * yd + zd * zd).astype(np.float32)
    val = v1 + (v2 - v1) * 0.5 * (1.0 + np.tanh((rc.astype(np.float64) - radius) / (np.sqrt(2.0) * xi)))
    return val.astype(np.float32).ravel()


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = a.astype(np.complex128)
        b = b.astype(np.complex128)
    else:
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def lowpass(cut):
    """Oracle Fourier callback: zero every mode with |n_x| > cut or
    |n_y| > cut (signed mode numbers, n > N/2 counted as n - N)."""
    def fn(z):
        _, sy, sx = z.shape
        nx = np.arange(sx)
        nx = np.where(nx > sx // 2, nx - sx, nx)
        ny = np.arange(sy)
        ny = np.where(ny > sy // 2, ny - sy, ny)
        keep = (np.abs(nx)[None, None, :] <= cut) & (np.abs(ny)[None, :, None] <= cut)
        return np.where(keep, z, 0)
    return fn


def f32(x):
    """The float32 value of a parameter, as the reference stores it."""
    return float(np.float32(x))


# ------------------------------------------------------------------ systems
#
# A system spec: the parser output of the reference for an equation set,
# written by hand (src/parser.cpp:628-767), so the same spec builds both the
# oracle and the HIP system.  Pres are tuples (pre, q2n, iqx, iqy, iqz, invq).

from cupss_amd.systems import (spec_cahn_hilliard, spec_diffusion, spec_kpz,  # noqa: E402,F401
                               spec_model_h)
from cupss_amd.systems import build as _build  # noqa: E402


def build_system(cls, pres_fn, spec):
    return _build(spec, cls, pres_fn)

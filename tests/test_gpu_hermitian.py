"""GPU: the kx = 0 and kx = sx/2 planes of every field's half spectrum stay
Hermitian, as the reference's IFFT / real part / FFT round trip of every
field every step keeps them (src/field.cpp:59-89).  Both partners of a bin
on these planes are stored; their anti-Hermitian part is invisible to C2R
and, fed by rounding, used to grow in linearly unstable bands (x12 per
1000 steps in the spinodal band of Cahn-Hilliard at dt = 0.01, until it
leaked into the real fields after ~10^4 steps).  After a long run the
planes are exactly Hermitian on every path (2D fused and general, 3D fused
and general, a model-H system with stored constraint fields), and a long
Cahn-Hilliard run stays bounded.
"""
import numpy as np
import pytest

import cupss_amd as C
from tests.helpers import build_system, spec_cahn_hilliard, spec_model_h

pytestmark = pytest.mark.gpu


def anti_hermitian(h, dims):
    """max |a - conj(partner)| / 2 over the kx = 0 and kx = sx/2 planes."""
    sx, sy, sz = dims
    h = h.reshape(sz, sy, sx // 2 + 1)
    worst = 0.0
    for kx in (0, sx // 2):
        p = h[:, :, kx]
        q = np.conj(p[(-np.arange(sz)) % sz][:, (-np.arange(sy)) % sy])
        worst = max(worst, float(np.abs(p - q).max()) / 2)
    return worst


CASES = [("ch_2d", lambda: spec_cahn_hilliard((128, 128, 1), seed=9), "phi", 3000),
         ("ch_3d", lambda: spec_cahn_hilliard((32, 32, 32), seed=9), "phi", 2000),
         ("model_h", lambda: spec_model_h(64, dt=0.05, seed=9), "phi", 2000)]


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_edge_planes_stay_hermitian(case, path):
    name, mk, key, steps = next(c for c in CASES if c[0] == case)
    spec = mk()
    g, ids = build_system(C.System, C.pres, spec)
    g.set_path(path)
    g.step(steps)
    worst = 0.0
    for f, fid in ids.items():
        h = g.get_spectrum(fid)
        a = anti_hermitian(h, spec["dims"])
        scale = max(1.0, float(np.abs(h).max()))
        print("  %s path %d field %-7s anti-Hermitian %.3e of max %.3e" % (case, path, f, a, scale), flush=True)
        worst = max(worst, a / scale)
    assert worst <= 1e-6
    x = g.get_real(ids[key])
    assert np.all(np.isfinite(x)) and float(np.abs(x).max()) < 1.5


def test_long_cahn_hilliard_stays_bounded():
    spec = spec_cahn_hilliard((256, 256, 1), seed=3)
    g, ids = build_system(C.System, C.pres, spec)
    g.step(15000)
    x = g.get_real(ids["phi"])
    print("  15000 steps: max|phi| %.4f" % float(np.abs(x).max()), flush=True)
    assert np.all(np.isfinite(x)) and float(np.abs(x).max()) < 1.2

"""ctypes wrapper of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  See oracle.h for what the
oracle restates and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


class Pres(ctypes.Structure):
    """Mirror of the reference `pres` (inc/cupss/defines.h:31-39)."""

    _fields_ = [("preFactor", ctypes.c_float), ("q2n", ctypes.c_int), ("iqx", ctypes.c_int),
                ("iqy", ctypes.c_int), ("iqz", ctypes.c_int), ("invq", ctypes.c_int)]


# void hook(void *user, void *data, int sx, int sy, int sz)
HOOK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int)

PLANT_NONE, PLANT_PHASES, PLANT_UNDEALIAS = 0, 1, 2


def set_plant(code):
    """Plant a known mistake in the oracle (tests that show a check can
    fail): PLANT_PHASES runs the dynamic phases before the constraint ones,
    PLANT_UNDEALIAS makes products read undealiased factors."""
    lib().orc_set_plant(int(code))


def pres(pre=1.0, q2n=0, iqx=0, iqy=0, iqz=0, invq=0):
    return Pres(float(pre), int(q2n), int(iqx), int(iqy), int(iqz), int(invq))


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        for pfx, real in (("orc_f32_", ctypes.c_float), ("orc_f64_", ctypes.c_double)):
            rp = ctypes.POINTER(real)
            getattr(L, pfx + "create").restype = ctypes.c_void_p
            getattr(L, pfx + "create").argtypes = [ctypes.c_int] * 3 + [ctypes.c_float] * 4
            getattr(L, pfx + "destroy").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "add_field").argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, pfx + "add_implicit").argtypes = [ctypes.c_void_p, ctypes.c_int, Pres]
            getattr(L, pfx + "add_term").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                     ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                                     ctypes.POINTER(Pres)]
            getattr(L, pfx + "set_noise").argtypes = [ctypes.c_void_p, ctypes.c_int, Pres, ctypes.c_uint64]
            getattr(L, pfx + "set_real").argtypes = [ctypes.c_void_p, ctypes.c_int,
                                                     ctypes.POINTER(ctypes.c_float)]
            getattr(L, pfx + "prepare").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "step").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "steps").argtypes = [ctypes.c_void_p, ctypes.c_int]
            for g in ("get_real", "get_comp", "get_dealiased"):
                getattr(L, pfx + g).argtypes = [ctypes.c_void_p, ctypes.c_int, rp]
            getattr(L, pfx + "aliasing_order").argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, pfx + "fft").argtypes = [rp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_dft_f64.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)] + \
            [ctypes.c_int] * 4
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.orc_normal.restype = ctypes.c_float
        L.orc_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_normal_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.orc_set_ref_cpu_dealias_bug.argtypes = [ctypes.c_int]
        L.orc_set_plant.argtypes = [ctypes.c_int]
        for pfx in ("orc_f32_", "orc_f64_"):
            getattr(L, pfx + "set_hooks").argtypes = [ctypes.c_void_p, ctypes.c_int, HOOK, HOOK, ctypes.c_void_p]
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def normals(seed, step, stream, n):
    out = np.empty(n, np.float32)
    lib().orc_normal_fill(seed, step, stream, _ptr(out, ctypes.c_float), n)
    return out


def fft(x, dims, direction, precision="f64"):
    """Unnormalised C2C over dims=(sx, sy, sz); x complex array of sx*sy*sz."""
    real = np.float32 if precision == "f32" else np.float64
    ct = ctypes.c_float if precision == "f32" else ctypes.c_double
    buf = np.empty(2 * x.size, real)
    buf[0::2] = x.real.ravel()
    buf[1::2] = x.imag.ravel()
    getattr(lib(), "orc_%s_fft" % precision)(_ptr(buf, ct), dims[0], dims[1], dims[2], direction)
    return buf[0::2] + 1j * buf[1::2]


def dft(x, dims, direction):
    buf = np.empty(2 * x.size, np.float64)
    buf[0::2] = x.real.ravel()
    buf[1::2] = x.imag.ravel()
    out = np.empty_like(buf)
    lib().orc_dft_f64(_ptr(buf, ctypes.c_double), _ptr(out, ctypes.c_double), dims[0], dims[1], dims[2],
                      direction)
    return out[0::2] + 1j * out[1::2]


class System:
    """One oracle system: fields, terms (products with summed prefactors),
    implicit operators and noise, stepped with the reference's phase order."""

    def __init__(self, sx, sy=1, sz=1, dx=1.0, dy=1.0, dz=1.0, dt=0.01, precision="f32"):
        self.p = "orc_%s_" % precision
        self.real = np.float32 if precision == "f32" else np.float64
        self.ct = ctypes.c_float if precision == "f32" else ctypes.c_double
        self.L = lib()
        self.dims = (sx, sy, sz)
        self.n = sx * sy * sz
        self.h = getattr(self.L, self.p + "create")(sx, sy, sz, dx, dy, dz, dt)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            getattr(self.L, self.p + "destroy")(h)
            self.h = None

    def _f(self, name):
        return getattr(self.L, self.p + name)

    def add_field(self, dynamic):
        return self._f("add_field")(self.h, int(bool(dynamic)))

    def add_implicit(self, f, p):
        self._f("add_implicit")(self.h, f, p)

    def add_term(self, f, factors, prefs):
        fac = (ctypes.c_int * max(1, len(factors)))(*factors)
        pr = (Pres * len(prefs))(*prefs)
        r = self._f("add_term")(self.h, f, len(factors), fac, len(prefs), pr)
        assert r >= 0
        return r

    def set_noise(self, f, amp, seed):
        self._f("set_noise")(self.h, f, amp, seed)

    def set_real(self, f, x):
        x = np.ascontiguousarray(np.asarray(x, np.float32).ravel())
        assert x.size == self.n
        self._f("set_real")(self.h, f, _ptr(x, ctypes.c_float))

    def set_hooks(self, f, fourier=None, real=None):
        """Field callbacks (src/field.cpp:48-58, 70-86).  Each is a Python
        function taking a complex128 view-copy of the N values (sz, sy, sx)
        and returning the new values (or None to leave them): `fourier` sees
        the updated full spectrum before dealiasing, `real` the normalised
        real array (imaginary parts zero), then the dealiased copy."""
        dims = self.dims
        real_t = self.real

        def wrap(fn):
            if fn is None:
                return HOOK()

            def cb(_user, data, sx, sy, sz):
                n = sx * sy * sz
                buf = np.ctypeslib.as_array(ctypes.cast(data, ctypes.POINTER(self.ct)), shape=(2 * n,))
                z = (buf[0::2] + 1j * buf[1::2]).reshape(dims[2], dims[1], dims[0])
                out = fn(z)
                if out is not None:
                    out = np.asarray(out).reshape(-1)
                    buf[0::2] = out.real.astype(real_t)
                    buf[1::2] = out.imag.astype(real_t)
            return HOOK(cb)

        hooks = (wrap(fourier), wrap(real))
        self._hooks = getattr(self, "_hooks", {})
        self._hooks[f] = hooks  # keep the callbacks alive as long as the system
        self._f("set_hooks")(self.h, f, hooks[0], hooks[1], None)

    def prepare(self):
        self._f("prepare")(self.h)

    def step(self, n=1):
        self._f("steps")(self.h, n)

    def get_real(self, f):
        out = np.empty(self.n, self.real)
        self._f("get_real")(self.h, f, _ptr(out, self.ct))
        return out

    def get_dealiased(self, f):
        out = np.empty(self.n, self.real)
        self._f("get_dealiased")(self.h, f, _ptr(out, self.ct))
        return out

    def get_comp(self, f):
        out = np.empty(2 * self.n, self.real)
        self._f("get_comp")(self.h, f, _ptr(out, self.ct))
        return out[0::2] + 1j * out[1::2]

    def aliasing_order(self, f):
        return self._f("aliasing_order")(self.h, f)

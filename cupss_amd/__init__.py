"""cupss_amd -- Python binding of libcupss's C-ABI (include/cupss_hip.h).

The library is the MI355X HIP implementation of cuPSS's per-timestep
pseudospectral update; this module only loads it with ctypes and mirrors the
reference's system-building calls (evolver::createField / createTerm /
addNoise / advanceTime, src/evolver.cpp) for tests, benchmarks and scripts.
There is no CPU fallback: if libcupss.so is missing or no GPU is present,
the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcupss.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


class CupssError(RuntimeError):
    pass


class Pres(ctypes.Structure):
    """cupss_pres: same layout as the reference `pres` (inc/cupss/defines.h:31-39)."""

    _fields_ = [("preFactor", ctypes.c_float), ("q2n", ctypes.c_int), ("iqx", ctypes.c_int),
                ("iqy", ctypes.c_int), ("iqz", ctypes.c_int), ("invq", ctypes.c_int)]


def pres(pre=1.0, q2n=0, iqx=0, iqy=0, iqz=0, invq=0):
    return Pres(float(pre), int(q2n), int(iqx), int(iqy), int(iqz), int(invq))


_lib = None

_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_FP = ctypes.POINTER(ctypes.c_float)
_SIGS = {
    "cupss_last_error": (ctypes.c_char_p, []),
    "cupss_abi_version": (_I, []),
    "cupss_system_create": (_I, [_I, _I, _I, _F, _F, _F, _F, ctypes.POINTER(_VP)]),
    "cupss_system_destroy": (_I, [_VP]),
    "cupss_field_create": (_I, [_VP, _I, ctypes.POINTER(_I)]),
    "cupss_field_add_implicit": (_I, [_VP, _I, Pres]),
    "cupss_field_set_implicits": (_I, [_VP, _I, ctypes.POINTER(Pres), _I]),
    "cupss_term_create": (_I, [_VP, _I, ctypes.POINTER(_I), _I, ctypes.POINTER(Pres), _I, ctypes.POINTER(_I)]),
    "cupss_term_set_prefactors": (_I, [_VP, _I, ctypes.POINTER(Pres), _I]),
    "cupss_field_set_noise": (_I, [_VP, _I, Pres, ctypes.c_uint64]),
    "cupss_field_set_hooks": (_I, [_VP, _I, _VP, _VP, _VP]),
    "cupss_field_set_real": (_I, [_VP, _I, _FP]),
    "cupss_field_get_real": (_I, [_VP, _I, _FP]),
    "cupss_field_get_dealiased": (_I, [_VP, _I, _FP]),
    "cupss_field_get_spectrum": (_I, [_VP, _I, _FP]),
    "cupss_field_set_spectrum": (_I, [_VP, _I, _FP]),
    "cupss_system_half_bins": (ctypes.c_long, [_VP]),
    "cupss_system_prepare": (_I, [_VP]),
    "cupss_field_aliasing_order": (_I, [_VP, _I]),
    "cupss_system_advance": (_I, [_VP, _I]),
    "cupss_system_time_advance": (_I, [_VP, _I, _FP]),
    "cupss_system_synchronize": (_I, [_VP]),
    "cupss_system_get_step": (_I, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    "cupss_system_set_step": (_I, [_VP, ctypes.c_uint64]),
    "cupss_system_set_stream": (_I, [_VP, _VP]),
    "cupss_system_set_dt": (_I, [_VP, _F]),
    "cupss_system_set_path": (_I, [_VP, _I]),
    "cupss_system_fused_eligible": (_I, [_VP]),
    "cupss_system_set_noise_mode": (_I, [_VP, _I]),
    "cupss_batch_create": (_I, [_VP, ctypes.POINTER(_VP)]),
    "cupss_batch_destroy": (_I, [_VP]),
    "cupss_batch_add": (_I, [_VP, _VP]),
    "cupss_batch_size": (_I, [_VP]),
    "cupss_batch_set_graph_steps": (_I, [_VP, _I]),
    "cupss_batch_set_autotune": (_I, [_VP, _I]),
    "cupss_batch_schedule": (_I, [_VP, ctypes.c_char_p, _I]),
    "cupss_batch_advance": (_I, [_VP, _I]),
    "cupss_batch_time_advance": (_I, [_VP, _I, _FP]),
    "cupss_batch_synchronize": (_I, [_VP]),
    "cupss_batch_remove": (_I, [_VP, _VP]),
    "cupss_system_set_graph_steps": (_I, [_VP, _I]),
    "cupss_system_device_bytes": (_I, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
    "cupss_system_profile_kernel": (_I, [_VP, _I]),
    "cupss_system_kernel_time": (_I, [_VP, _FP, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, _I]),
    "cupss_batch_profile_kernel": (_I, [_VP, _I]),
    "cupss_batch_kernel_time": (_I, [_VP, _FP, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, _I]),
    "cupss_server_create": (_I, [ctypes.c_size_t, _I, _I, ctypes.POINTER(_VP)]),
    "cupss_server_destroy": (_I, [_VP]),
    "cupss_server_submit": (_I, [_VP, ctypes.c_double, _I, ctypes.c_size_t, _VP, _VP, _VP, ctypes.POINTER(_I)]),
    "cupss_server_run": (_I, [_VP]),
    "cupss_server_job_times": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
    "cupss_server_stats": (_I, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_I)]),
    "cupss_field_spectrum_device": (_I, [_VP, _I, ctypes.POINTER(_VP)]),
    "cupss_r2c": (_I, [_I, _I, _I, _VP, _VP, _VP]),
    "cupss_c2r": (_I, [_I, _I, _I, _VP, _I, _F, _VP, _VP, _VP]),
    "cupss_c2r_noalloc": (_I, [_I, _I, _I, _VP, _I, _F, _VP, _VP]),
    "cupss_normals": (_I, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, _VP, ctypes.c_long, _VP]),
    "cupss_device_alloc": (_I, [ctypes.c_size_t, ctypes.POINTER(_VP)]),
    "cupss_device_free": (_I, [_VP]),
    "cupss_memcpy_h2d": (_I, [_VP, _VP, ctypes.c_size_t]),
    "cupss_memcpy_d2h": (_I, [_VP, _VP, ctypes.c_size_t]),
    "cupss_memcpy_d2d": (_I, [_VP, _VP, ctypes.c_size_t]),
    "cupss_device_synchronize": (_I, []),
    "cupss_device_count": (_I, [ctypes.POINTER(_I)]),
}


def lib():
    """Load libcupss.so (built by cupss_amd/build.py).  Raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CupssError("libcupss.so is not built: run `python -m cupss_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise CupssError("libcupss error %d: %s" % (rc, lib().cupss_last_error().decode()))
    return rc


def _fptr(a):
    return a.ctypes.data_as(_FP)


def gpu_count():
    n = ctypes.c_int(0)
    rc = lib().cupss_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """A raw device allocation made by the library (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(lib().cupss_device_alloc(self.nbytes, ctypes.byref(self.ptr)))

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().cupss_memcpy_h2d(self.ptr, arr.ctypes.data_as(_VP), arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        check(lib().cupss_memcpy_d2h(out.ctypes.data_as(_VP), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().cupss_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class System:
    """One system on the GPU: the device side of one reference evolver."""

    def __init__(self, sx, sy=1, sz=1, dx=1.0, dy=1.0, dz=1.0, dt=0.01):
        self.L = lib()
        self.h = ctypes.c_void_p()
        check(self.L.cupss_system_create(sx, sy, sz, dx, dy, dz, dt, ctypes.byref(self.h)))
        self.dims = (sx, sy, sz)
        self.n = sx * sy * sz
        self.nh = (sx // 2 + 1) * sy * sz

    owned = True

    @classmethod
    def borrow(cls, handle, dims):
        """A view of a system owned elsewhere (e.g. by a server); never destroys it."""
        s = cls.__new__(cls)
        s.L = lib()
        s.h = ctypes.c_void_p(handle)
        s.owned = False
        s.dims = tuple(dims)
        s.n = dims[0] * dims[1] * dims[2]
        s.nh = (dims[0] // 2 + 1) * dims[1] * dims[2]
        return s

    def detach(self):
        """Hand the system over (e.g. to a server, which destroys it); returns the raw handle."""
        h = self.h.value
        self.h = ctypes.c_void_p()
        return h

    def close(self):
        if getattr(self, "h", None) and self.h.value and self.owned:
            self.L.cupss_system_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph_steps(self, k):
        check(self.L.cupss_system_set_graph_steps(self.h, k))

    def profile_kernel(self, on=True):
        check(self.L.cupss_system_profile_kernel(self.h, int(on)))

    def kernel_time(self):
        """(average ms, real points per launch, name) of the timed column launches."""
        ms, pts, name = ctypes.c_float(), ctypes.c_double(), ctypes.create_string_buffer(64)
        check(self.L.cupss_system_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(pts), name, 64))
        return ms.value, pts.value, name.value.decode()

    def device_bytes(self):
        b = ctypes.c_size_t()
        check(self.L.cupss_system_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def add_field(self, dynamic):
        fid = ctypes.c_int()
        check(self.L.cupss_field_create(self.h, int(bool(dynamic)), ctypes.byref(fid)))
        return fid.value

    def add_implicit(self, f, p):
        check(self.L.cupss_field_add_implicit(self.h, f, p))

    def set_implicits(self, f, prefs):
        arr = (Pres * max(1, len(prefs)))(*prefs)
        check(self.L.cupss_field_set_implicits(self.h, f, arr, len(prefs)))

    def add_term(self, f, factors, prefs):
        fac = (ctypes.c_int * max(1, len(factors)))(*factors)
        arr = (Pres * len(prefs))(*prefs)
        tid = ctypes.c_int()
        check(self.L.cupss_term_create(self.h, f, fac, len(factors), arr, len(prefs), ctypes.byref(tid)))
        return tid.value

    def set_term_prefactors(self, t, prefs):
        arr = (Pres * len(prefs))(*prefs)
        check(self.L.cupss_term_set_prefactors(self.h, t, arr, len(prefs)))

    def set_noise(self, f, amp, seed):
        check(self.L.cupss_field_set_noise(self.h, f, amp, seed))

    def set_real(self, f, x):
        x = np.ascontiguousarray(np.asarray(x, np.float32).ravel())
        assert x.size == self.n
        check(self.L.cupss_field_set_real(self.h, f, _fptr(x)))

    def prepare(self):
        check(self.L.cupss_system_prepare(self.h))

    def step(self, n=1):
        check(self.L.cupss_system_advance(self.h, n))

    def time_steps(self, n):
        ms = ctypes.c_float()
        check(self.L.cupss_system_time_advance(self.h, n, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        check(self.L.cupss_system_synchronize(self.h))

    def get_real(self, f):
        out = np.empty(self.n, np.float32)
        check(self.L.cupss_field_get_real(self.h, f, _fptr(out)))
        return out

    def get_dealiased(self, f):
        out = np.empty(self.n, np.float32)
        check(self.L.cupss_field_get_dealiased(self.h, f, _fptr(out)))
        return out

    def get_spectrum(self, f):
        out = np.empty(2 * self.nh, np.float32)
        check(self.L.cupss_field_get_spectrum(self.h, f, _fptr(out)))
        return out[0::2] + 1j * out[1::2]

    def set_path(self, mode):
        """0 = fused two-kernel step when eligible, 1 = general step program."""
        check(self.L.cupss_system_set_path(self.h, int(mode)))

    def set_noise_mode(self, mode):
        """0 = real-space Philox + R2C (the reference's draw), 1 = spectral-space draw."""
        check(self.L.cupss_system_set_noise_mode(self.h, int(mode)))

    @property
    def fused_eligible(self):
        return bool(self.L.cupss_system_fused_eligible(self.h))

    def aliasing_order(self, f):
        return self.L.cupss_field_aliasing_order(self.h, f)

    @property
    def step_count(self):
        v = ctypes.c_uint64()
        check(self.L.cupss_system_get_step(self.h, ctypes.byref(v)))
        return v.value


class Batch:
    """Many systems advanced together by the batched fused step."""

    def __init__(self, graph_steps=0, autotune=True):
        self.L = lib()
        self.h = ctypes.c_void_p()
        check(self.L.cupss_batch_create(None, ctypes.byref(self.h)))
        self.members = []
        if graph_steps:
            check(self.L.cupss_batch_set_graph_steps(self.h, graph_steps))
        if not autotune:
            check(self.L.cupss_batch_set_autotune(self.h, 0))

    def profile_kernel(self, on=True):
        check(self.L.cupss_batch_profile_kernel(self.h, int(on)))

    def kernel_time(self):
        """(average ms, real points per launch, name) of the timed column launches."""
        ms, pts, name = ctypes.c_float(), ctypes.c_double(), ctypes.create_string_buffer(64)
        check(self.L.cupss_batch_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(pts), name, 64))
        return ms.value, pts.value, name.value.decode()

    def schedule(self):
        """The schedule in use: size classes, tile shapes, chunks."""
        buf = ctypes.create_string_buffer(4096)
        check(self.L.cupss_batch_schedule(self.h, buf, len(buf)))
        return buf.value.decode()

    def add(self, system):
        check(self.L.cupss_batch_add(self.h, system.h))
        self.members.append(system)

    def remove(self, system):
        check(self.L.cupss_batch_remove(self.h, system.h))
        self.members.remove(system)

    def step(self, n=1):
        check(self.L.cupss_batch_advance(self.h, n))

    def time_steps(self, n):
        ms = ctypes.c_float()
        check(self.L.cupss_batch_time_advance(self.h, n, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        check(self.L.cupss_batch_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.cupss_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_BUILD = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p))
_DONE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)


class Server:
    """Serving loop with continuous batching (cupss_server_* of the C-ABI).

    submit(build, nsteps, arrival, bytes_estimate, done): `build()` returns a
    prepared System (built only when the job is admitted); `done(system)`
    gets a borrowed view of the finished system before it is destroyed.
    """

    def __init__(self, budget_bytes, quantum=100, max_active=64):
        self.L = lib()
        self.h = ctypes.c_void_p()
        check(self.L.cupss_server_create(int(budget_bytes), quantum, max_active, ctypes.byref(self.h)))
        self.jobs = []
        self.errors = []
        self._build_cb = _BUILD(self._on_build)
        self._done_cb = _DONE(self._on_done)

    def _on_build(self, user, job, out):
        try:
            s = self.jobs[job]["build"]()
            self.jobs[job]["dims"] = s.dims
            out[0] = s.detach()
            return 0
        except Exception as e:  # reported by run()
            self.errors.append(e)
            return -1

    def _on_done(self, user, job, handle):
        try:
            cb = self.jobs[job]["done"]
            if cb is not None:
                cb(System.borrow(handle, self.jobs[job]["dims"]))
        except Exception as e:
            self.errors.append(e)

    def submit(self, build, nsteps, arrival=0.0, bytes_estimate=0, done=None):
        jid = ctypes.c_int()
        self.jobs.append({"build": build, "done": done, "nsteps": nsteps})
        check(self.L.cupss_server_submit(self.h, float(arrival), int(nsteps), int(bytes_estimate),
                                         ctypes.cast(self._build_cb, _VP), ctypes.cast(self._done_cb, _VP),
                                         None, ctypes.byref(jid)))
        return jid.value

    def run(self):
        rc = self.L.cupss_server_run(self.h)
        if self.errors:
            raise self.errors[0]
        check(rc)

    def times(self, job):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.L.cupss_server_job_times(self.h, job, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def stats(self):
        w, p, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_size_t(), ctypes.c_int()
        check(self.L.cupss_server_stats(self.h, ctypes.byref(w), ctypes.byref(p), ctypes.byref(b), ctypes.byref(n)))
        return {"wall_s": w.value, "point_steps": p.value, "peak_bytes": b.value, "peak_active": n.value}

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.cupss_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

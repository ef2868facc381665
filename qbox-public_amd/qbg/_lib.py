"""ctypes binding of libqbg.so (the C ABI declared in include/qbg.h).

The shared library is built in-tree by ``__graft_entry__.build()`` into
``qbox-public_amd/lib/libqbg.so``.  Every entry point returns a status code;
a non-zero code raises :class:`QbgError` with the context's error text.  There
is no CPU fallback: if the library or a HIP device is missing, loading fails
loudly.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent.parent / "lib" / "libqbg.so"

QBG_OK = 0
QBG_E_INVAL = -1
QBG_E_NOMEM = -2
QBG_E_HIP = -3
QBG_E_NOT_SPD = -4
QBG_E_STATE = -5
PREC = {"fp64": 0, "fp32": 1}

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

# name -> argtypes (restype int unless noted)
SIGNATURES = {
    "qbg_init": [C.c_int, C.POINTER(_vp)],
    "qbg_finalize": [_vp],
    "qbg_last_error": [_vp],
    "qbg_sync": [_vp],
    "qbg_malloc": [_vp, C.c_size_t, C.POINTER(_vp)],
    "qbg_free": [_vp, _vp],
    "qbg_memcpy_h2d": [_vp, _vp, _vp, C.c_size_t],
    "qbg_memcpy_d2h": [_vp, _vp, _vp, C.c_size_t],
    "qbg_memset": [_vp, _vp, C.c_int, C.c_size_t],
    "qbg_timer_start": [_vp],
    "qbg_timer_stop": [_vp, C.POINTER(C.c_float)],
    "qbg_grid_create": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip,
                        C.POINTER(_vp)],
    "qbg_grid_destroy": [_vp],
    "qbg_fft_backward": [_vp, _vp, _vp, C.c_int],
    "qbg_fft_backward_pair": [_vp, _vp, _vp, _vp, C.c_int],
    "qbg_fft_forward": [_vp, _vp, _vp, C.c_int],
    "qbg_fft_forward_pair": [_vp, _vp, _vp, _vp, C.c_int],
    "qbg_compute_density": [_vp, _vp, C.c_int, C.c_int, _dp, C.c_double, C.c_double, _vp,
                            C.c_int],
    "qbg_rs_mul_add": [_vp, _vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int],
}

_lib = None


class QbgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qbg error {code}: {msg}")
        self.code = code


def load():
    """Load libqbg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(f"libqbg.so not built at {LIBPATH}; run __graft_entry__.build()")
        lib = C.CDLL(str(LIBPATH))
        for name, args in SIGNATURES.items():
            f = getattr(lib, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = C.c_char_p if name == "qbg_last_error" else C.c_int
        _lib = lib
    return _lib


class Context:
    """qbg_ctx: one device and one stream."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _vp()
        rc = lib.qbg_init(device, C.byref(h))
        if rc != QBG_OK:
            raise QbgError(rc, "qbg_init failed (no HIP device?)")
        self.h = h
        self.lib = lib

    def check(self, rc):
        if rc != QBG_OK:
            raise QbgError(rc, self.lib.qbg_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.qbg_sync(self.h))

    def close(self):
        if self.h:
            self.lib.qbg_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, shape, dtype=np.float64):
        return DArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def timer_start(self):
        self.check(self.lib.qbg_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_float()
        self.check(self.lib.qbg_timer_stop(self.h, C.byref(ms)))
        return ms.value


class DArray:
    """A device allocation with a numpy shape/dtype for transfers."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        ctx.check(ctx.lib.qbg_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p

    @property
    def p(self):
        return self.ptr

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        self.ctx.check(self.ctx.lib.qbg_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.qbg_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                                   self.nbytes))
        return out

    def zero(self):
        self.ctx.check(self.ctx.lib.qbg_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def offset(self, nbytes):
        return _vp(self.ptr.value + nbytes)

    def free(self):
        if self.ptr:
            self.ctx.lib.qbg_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Grid:
    """qbg_grid: FourierTransform + BasisMapping built from Qbox's maps."""

    def __init__(self, ctx: Context, np0, np1, np2, ip, im, zvec_to_val):
        self.ctx = ctx
        ip = np.ascontiguousarray(ip, dtype=np.int32)
        im = np.ascontiguousarray(im, dtype=np.int32)
        z2v = np.ascontiguousarray(zvec_to_val, dtype=np.int32)
        h = _vp()
        ctx.check(ctx.lib.qbg_grid_create(ctx.h, np0, np1, np2, ip.size, z2v.size,
                                          ip.ctypes.data_as(_ip), im.ctypes.data_as(_ip),
                                          z2v.ctypes.data_as(_ip), C.byref(h)))
        self.h = h
        self.np = (np0, np1, np2)
        self.ng = ip.size
        self.N = np0 * np1 * np2

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.qbg_grid_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def fft_backward(self, c: DArray, f: DArray, prec="fp64"):
        self.ctx.check(self.ctx.lib.qbg_fft_backward(self.h, c.p, f.p, PREC[prec]))

    def fft_backward_pair(self, c1, c2, f: DArray, prec="fp64"):
        self.ctx.check(self.ctx.lib.qbg_fft_backward_pair(self.h, c1, c2, f.p, PREC[prec]))

    def fft_forward(self, f: DArray, c: DArray, prec="fp64"):
        self.ctx.check(self.ctx.lib.qbg_fft_forward(self.h, f.p, c.p, PREC[prec]))

    def fft_forward_pair(self, f: DArray, c1, c2, prec="fp64"):
        self.ctx.check(self.ctx.lib.qbg_fft_forward_pair(self.h, f.p, c1, c2, PREC[prec]))

    def compute_density(self, c: DArray, ldc, nst, occ, weight, omega, rho: DArray, prec="fp64"):
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        self.ctx.check(self.ctx.lib.qbg_compute_density(self.h, c.p, ldc, nst,
                                                        occ.ctypes.data_as(_dp), weight, omega,
                                                        rho.p, PREC[prec]))

    def rs_mul_add(self, v: DArray, c: DArray, ldc, nst, hpsi: DArray, ldh, prec="fp64"):
        self.ctx.check(self.ctx.lib.qbg_rs_mul_add(self.h, v.p, c.p, ldc, nst, hpsi.p, ldh,
                                                   PREC[prec]))

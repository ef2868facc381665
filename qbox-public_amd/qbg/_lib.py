"""ctypes binding of libqbg.so (the C ABI declared in include/qbg.h).

The shared library is built in-tree by ``__graft_entry__.build()`` into
``qbox-public_amd/lib/libqbg.so``.  Every entry point returns a status code;
a non-zero code raises :class:`QbgError` with the context's error text.  There
is no CPU fallback: if the library or a HIP device is missing, loading fails
loudly.
"""
from __future__ import annotations

import ctypes as C
import weakref
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent.parent / "lib" / "libqbg.so"

QBG_OK = 0
QBG_E_INVAL = -1
QBG_E_NOMEM = -2
QBG_E_HIP = -3
QBG_E_NOT_SPD = -4
QBG_E_STATE = -5
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x2": 3}

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
    "qbg_grid_info": [_vp, _ip],
    "qbg_grid_tuning": [_vp, C.c_int, _ip],
    "qbg_grid_plan": [_vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "qbg_fft_splits": [C.c_int, _ip, C.c_int],
    "qbg_fft_backward": [_vp, _vp, _vp, C.c_int],
    "qbg_fft_backward_pair": [_vp, _vp, _vp, _vp, C.c_int],
    "qbg_fft_forward": [_vp, _vp, _vp, C.c_int],
    "qbg_fft_forward_pair": [_vp, _vp, _vp, _vp, C.c_int],
    "qbg_compute_density": [_vp, _vp, C.c_int, C.c_int, _dp, C.c_double, C.c_double, _vp,
                            C.c_int],
    "qbg_rs_mul_add": [_vp, _vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int],
    "qbg_stream_copy": [_vp, _vp, _vp, C.c_size_t],
    "qbg_prof_enable": [_vp, C.c_int],
    "qbg_prof_reset": [_vp],
    "qbg_prof_read": [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "qbg_gram": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "qbg_gram_complex": [_vp, _vp, C.c_int, C.c_int, C.c_int],
    "qbg_gram_noalloc": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "qbg_gram_noalloc_bytes": [],
    "qbg_ef_create": [_vp, _dp, _dp, C.c_double, C.c_int, _vp, _ip, _dp, C.c_int, _dp,
                      C.POINTER(_vp)],
    "qbg_ef_create_k": [_vp, _dp, _dp, C.c_double, _dp, C.c_int, _vp, _ip, _dp, C.c_int, _dp,
                        C.POINTER(_vp)],
    "qbg_ef_create_kset": [_vp, _dp, _dp, C.c_double, C.c_int, _dp, _dp, C.c_int, _vp, _ip, _dp,
                           C.c_int, _dp, C.POINTER(_vp)],
    "qbg_ef_kset_basis": [_vp, C.c_int, _ip, _ip, _dp, _dp],
    "qbg_ef_destroy": [_vp],
    "qbg_ef_sizes": [_vp, _ip],
    "qbg_ef_basis": [_vp, C.c_int, _ip, _dp, _dp, _ip, _ip, _ip],
    "qbg_ef_set_positions": [_vp, _dp],
    "qbg_ef_cell_moved": [_vp, _dp],
    "qbg_nl_create": [_vp, C.c_int, C.POINTER(_vp)],
    "qbg_nl_destroy": [_vp],
    "qbg_nl_update_twnl": [_vp],
    "qbg_nl_energy": [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _dp, C.c_int, _dp,
                      C.c_int, _dp],
    "qbg_ef_update_density": [_vp, _vp, C.c_int, C.c_int],
    "qbg_ef_update_vhxc": [_vp, C.c_int],
    "qbg_ef_energy": [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                      _dp, _dp, _dp],
    "qbg_ef_get": [_vp, C.c_int, _dp],
    "qbg_ef_set_nspin": [_vp, C.c_int, C.c_int],
    "qbg_wfstep_create_bands": [_vp, C.c_int, C.c_double, C.c_int, C.POINTER(_vp)],
    "qbg_ef_wf_grid": [_vp, _ip, C.POINTER(_vp)],
    "qbg_ef_band_grid": [_vp, C.POINTER(_vp)],
    "qbg_ef_gemm_plan": [_vp, _ip],
    "qbg_ef_set_vr": [_vp, _dp],
    "qbg_ef_evaluate_async": [_vp, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int],
    "qbg_ef_energy_async": [_vp, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int],
    "qbg_ef_results": [_vp, _vp, _vp, _vp],
    "qbg_wfstep_create": [_vp, C.c_int, C.c_double, C.POINTER(_vp)],
    "qbg_wfstep_destroy": [_vp],
    "qbg_wfstep_preprocess": [_vp],
    "qbg_wfstep_update": [_vp, _vp, C.c_int, _vp, C.c_int],
    "qbg_wfstep_results": [_vp, _dp],
    "qbg_mixer_create": [_vp, C.c_double, C.c_double, C.c_int, C.c_double, C.POINTER(_vp)],
    "qbg_mixer_destroy": [_vp],
    "qbg_mixer_restart": [_vp],
    "qbg_mixer_update": [_vp],
    "qbg_polar": [_vp, _vp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int)],
    "qbg_stream_join": [_vp, _vp],
    "qbg_graph_begin": [_vp],
    "qbg_graph_end": [_vp, C.POINTER(_vp)],
    "qbg_graph_launch": [_vp],
    "qbg_graph_destroy": [_vp],
}

KERNEL_KINDS = ["zbwd", "xbwd", "yvmul", "ydens", "xfwd", "zfwd", "yother", "nl", "gram",
                "other", "vgrid", "gemm", "wfstep", "mix"]
ENERGY_NAMES = ["ekin", "eps", "enl", "ehart", "ehart_e", "ehart_ep", "ehart_p", "exc", "dxc",
                "esr", "eself", "ecoul", "etotal", "total_charge"]


class SpeciesDesc(C.Structure):
    """qbg_species_desc (include/qbg.h)."""
    _fields_ = [("zval", C.c_int), ("lmax", C.c_int), ("nmesh", C.c_int),
                ("deltar", C.c_double), ("rcps", C.c_double),
                ("vloc", _dp), ("proj", _dp), ("dl", _dp),
                ("type", C.c_int), ("llocal", C.c_int), ("vps", _dp), ("phi", _dp),
                ("nlcc", _dp), ("nquad", C.c_int), ("rquad", C.c_double)]


def species_desc(sp, keep):
    """qbg_species_desc of a configs.SpeciesDef (semilocal form with explicit
    projectors) or of a species.NcppSpecies (semilocal potentials and
    pseudo-wavefunctions, turned into Kleinman-Bylander projectors by the
    library)."""
    core = sp.core_density()
    if core is not None:
        core = np.ascontiguousarray(core, dtype=np.float64)
        keep.append(core)
    if getattr(sp, "kind", "") == "ncpp":
        vps = np.ascontiguousarray(sp.vps, dtype=np.float64)
        phi = np.ascontiguousarray(sp.phi, dtype=np.float64)
        keep += [vps, phi]
        return SpeciesDesc(sp.zval, sp.lmax, sp.nmesh, sp.deltar, sp.rcps, None, None, None,
                           1, sp.llocal, _pd(vps), _pd(phi), _pd(core), sp.nquad, sp.rquad)
    vloc, proj, dl = sp.radial()
    keep += [vloc, proj, dl]
    return SpeciesDesc(sp.zval, sp.lmax, sp.nmesh, sp.deltar, sp.rcps, _pd(vloc), _pd(proj),
                       _pd(dl), 0, 0, None, None, _pd(core), 0, 0.0)

_lib = None


class QbgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qbg error {code}: {msg}")
        self.code = code


RESTYPES = {"qbg_last_error": C.c_char_p, "qbg_gram_noalloc_bytes": C.c_size_t}


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
            f.restype = RESTYPES.get(name, C.c_int)
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
        # objects created on this context; released before the context is
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        self._children.add(obj)
        return obj

    def check(self, rc):
        if rc != QBG_OK:
            raise QbgError(rc, self.lib.qbg_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.qbg_sync(self.h))

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.release()
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


class Graph:
    """qbg_graph: the calls made on a context inside `with Graph(ctx) as g:`
    are captured into one HIP graph; g.launch() replays them."""

    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        self.h = None

    def __enter__(self):
        self.ctx.check(self.ctx.lib.qbg_graph_begin(self.ctx.h))
        return self

    def __exit__(self, et, ev, tb):
        h = _vp()
        rc = self.ctx.lib.qbg_graph_end(self.ctx.h, C.byref(h))
        if et is None:
            self.ctx.check(rc)
            self.h = h
            self.ctx.adopt(self)
        elif rc == QBG_OK and h:
            self.ctx.lib.qbg_graph_destroy(h)
        return False

    def launch(self):
        self.ctx.check(self.ctx.lib.qbg_graph_launch(self.h))

    def release(self):
        if self.h and self.ctx.h:
            self.ctx.lib.qbg_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


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
        ctx.adopt(self)

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

    def release(self):
        if self.ptr and self.ctx.h:
            self.ctx.lib.qbg_free(self.ctx.h, self.ptr)
        self.ptr = None

    free = release

    def __del__(self):
        try:
            self.release()
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
        ctx.adopt(self)

    def info(self) -> dict:
        """qbg_grid_info: sizes, ntrans0 and the line splits that serve the grid."""
        out = np.zeros(13, dtype=np.int32)
        self.ctx.check(self.ctx.lib.qbg_grid_info(self.h, out.ctypes.data_as(_ip)))
        return dict(np=tuple(int(x) for x in out[:3]), ng=int(out[3]), nvec=int(out[4]),
                    ntrans0=int(out[5]), four_step=bool(out[6]),
                    splits=[(int(out[7 + 2 * a]), int(out[8 + 2 * a])) for a in range(3)])

    @property
    def four_step(self) -> bool:
        return self.info()["four_step"]

    def tuning(self, prec="fp64") -> dict:
        """qbg_grid_tuning: the launch choice of the batched passes for prec."""
        out = np.zeros(6, dtype=np.int32)
        self.ctx.check(self.ctx.lib.qbg_grid_tuning(self.h, PREC[prec], out.ctypes.data_as(_ip)))
        return dict(tuned=bool(out[0]), x=(int(out[1]), int(out[2])),
                    y=(int(out[3]), int(out[4])), yppb=int(out[5]))

    def plan(self, prec="fp64", x_n1=0, y_n1=0, yppb=0):
        """qbg_grid_plan: fix the x/y splits (by n1, 0 = default) and pairs per block."""
        self.ctx.check(self.ctx.lib.qbg_grid_plan(self.h, PREC[prec], x_n1, y_n1, yppb))

    def release(self):
        if self.h and self.ctx.h:
            self.ctx.lib.qbg_grid_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
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


class _BorrowedGrid(Grid):
    """A grid owned by an energy functional (qbg_ef_wf_grid)."""

    def __init__(self, ctx: Context, h, dims):
        self.ctx = ctx
        self.h = h
        self.np = tuple(dims)
        self.ng = None
        self.N = dims[0] * dims[1] * dims[2]

    def release(self):
        self.h = None


def _pd(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Profiler:
    """Per-kernel HIP-event timing of a context (qbg_prof_*)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def enable(self, on=True):
        self.ctx.check(self.ctx.lib.qbg_prof_enable(self.ctx.h, int(on)))

    def reset(self):
        self.ctx.check(self.ctx.lib.qbg_prof_reset(self.ctx.h))

    def read(self):
        out = {}
        for k, name in enumerate(KERNEL_KINDS):
            ms, n = C.c_double(), C.c_int()
            self.ctx.check(self.ctx.lib.qbg_prof_read(self.ctx.h, k, C.byref(ms), C.byref(n)))
            if n.value:
                out[name] = (ms.value, n.value)
        return out


def stream_join(before: Context, after: Context):
    """qbg_stream_join: work on `after` waits for the work enqueued on
    `before`; inside a capture on `before`, `after` becomes a branch of the
    graph (join it back before the capture ends)."""
    after.check(after.lib.qbg_stream_join(before.h, after.h))


def fft_splits(n) -> list:
    """qbg_fft_splits: n1 of every compiled split of line length n (default first)."""
    out = np.zeros(8, dtype=np.int32)
    k = load().qbg_fft_splits(int(n), out.ctypes.data_as(_ip), 8)
    if k < 0:
        raise QbgError(k, "qbg_fft_splits")
    return [int(x) for x in out[:k]]


def gram(ctx: Context, c: DArray, ldc, ngw, nst, prec="fp64"):
    """SlaterDet::gram on device coefficients (in place)."""
    ctx.check(ctx.lib.qbg_gram(ctx.h, c.p, ldc, ngw, nst, PREC[prec]))


def gram_complex(ctx: Context, c: DArray, ldc, ngw, nst):
    """SlaterDet::gram of the complex (k != 0) basis on device coefficients
    (in place, FP64)."""
    ctx.check(ctx.lib.qbg_gram_complex(ctx.h, c.p, ldc, ngw, nst))


def gram_noalloc(ctx: Context, c: DArray, ldc, ngw, nst, prec="fp64"):
    """qbg_gram_noalloc: SlaterDet::gram with every temporary in the
    library's fixed per-device array (no device allocation in the call)."""
    ctx.check(ctx.lib.qbg_gram_noalloc(ctx.h, c.p, ldc, ngw, nst, PREC[prec]))


class EnergyFunctional:
    """qbg_ef: one system (Gamma, or one k-point off Gamma) resident on the
    device."""

    def __init__(self, ctx: Context, sysdef, refcell=None, kpoint=None, kpoints=None,
                 weights=None):
        """refcell (3x3, rows a0 a1 a2): Wavefunction's reference cell, which
        fixes the basis membership and the grids (src/Basis.cpp:314-325);
        sysdef.refcell when not given.  kpoint (3, units of the reciprocal
        vectors; sysdef.kpoint when not given): k != 0 makes the wavefunction
        basis the complex full sphere (qbg_ef_create_k).  kpoints (nkp x 3)
        and weights (nkp): a k-point set (qbg_ef_create_kset); k-point ik's
        bands are then the rows [ik nst, (ik+1) nst) of the coefficient
        array, whose width is ngw, the largest basis (kset_pack)."""
        self.ctx = ctx
        self.sysdef = sysdef
        if refcell is None:
            refcell = getattr(sysdef, "refcell", None)
        self.refcell = None if refcell is None else np.asarray(refcell, dtype=np.float64)
        descs = (SpeciesDesc * len(sysdef.species))()
        self._keep = [descs]
        for i, sp in enumerate(sysdef.species):
            descs[i] = species_desc(sp, self._keep)
        cell = np.ascontiguousarray(sysdef.cell, dtype=np.float64).ravel()
        na = np.ascontiguousarray(sysdef.na, dtype=np.int32)
        tau = np.ascontiguousarray(sysdef.tau, dtype=np.float64).ravel()
        occ = np.ascontiguousarray(sysdef.occ, dtype=np.float64)
        ref = None
        if refcell is not None:
            ref = np.ascontiguousarray(refcell, dtype=np.float64).ravel()
        self._keep += [cell, na, tau, occ, ref]
        h = _vp()
        if kpoint is None:
            kpoint = getattr(sysdef, "kpoint", None)
        self.kpoint = None
        if kpoint is not None:
            self.kpoint = np.ascontiguousarray(kpoint, dtype=np.float64).ravel()
        # the complex (full-sphere) basis of a k-point off Gamma
        self.complex = self.kpoint is not None and bool(np.any(self.kpoint != 0.0))
        self.nkp = 1
        if kpoints is not None:
            kps = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
            wts = np.ascontiguousarray(weights, dtype=np.float64).ravel()
            self.nkp = kps.shape[0]
            self.complex = True
            self._keep += [kps, wts]
            ctx.check(ctx.lib.qbg_ef_create_kset(ctx.h, _pd(cell),
                                                 _pd(ref) if ref is not None else None,
                                                 sysdef.ecut, self.nkp, _pd(kps), _pd(wts),
                                                 len(sysdef.species), C.cast(descs, _vp),
                                                 na.ctypes.data_as(_ip), _pd(tau), sysdef.nst,
                                                 _pd(occ), C.byref(h)))
        elif self.kpoint is not None:
            self._keep.append(self.kpoint)
            ctx.check(ctx.lib.qbg_ef_create_k(ctx.h, _pd(cell),
                                              _pd(ref) if ref is not None else None,
                                              sysdef.ecut, _pd(self.kpoint), len(sysdef.species),
                                              C.cast(descs, _vp), na.ctypes.data_as(_ip),
                                              _pd(tau), sysdef.nst, _pd(occ), C.byref(h)))
        else:
            ctx.check(ctx.lib.qbg_ef_create(ctx.h, _pd(cell), _pd(ref) if ref is not None else None,
                                            sysdef.ecut, len(sysdef.species),
                                            C.cast(descs, _vp), na.ctypes.data_as(_ip), _pd(tau),
                                            sysdef.nst, _pd(occ), C.byref(h)))
        self.h = h
        s = np.zeros(10, dtype=np.int32)
        ctx.check(ctx.lib.qbg_ef_sizes(h, s.ctypes.data_as(_ip)))
        (self.ngw, self.ngv, self.np0, self.np1, self.np2, self.nvec, self.nvec_v, self.nat,
         self.nproj, self.nst) = [int(x) for x in s]
        self.N = self.np0 * self.np1 * self.np2
        self.kngw = []
        for ik in range(self.nkp):
            n = C.c_int(0)
            ctx.check(ctx.lib.qbg_ef_kset_basis(h, ik, C.byref(n), None, None, None))
            self.kngw.append(int(n.value))
        self.ngw = max(self.kngw)            # the coefficients' width
        ctx.adopt(self)
        if getattr(sysdef, "nspin", 1) == 2:
            self.set_nspin(2, sysdef.nst_up)

    def release(self):
        if self.h and self.ctx.h:
            self.ctx.lib.qbg_ef_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def basis(self, which=0):
        """which 0: wavefunction basis, 1: density basis, 2: the wavefunction
        basis's k + G (g2 = |k+G|^2, gx = k+G).  The band-grid maps (ip, im,
        zvec_to_val) of a k != 0 functional are None."""
        n = self.ngv if which == 1 else self.kngw[0]
        nvec = self.nvec_v if which == 1 else self.nvec
        maps = which == 1 or not self.complex
        idx = np.zeros(3 * n, dtype=np.int32)
        g2 = np.zeros(n)
        gx = np.zeros(3 * n)
        ip = np.zeros(n, dtype=np.int32)
        im = np.zeros(n, dtype=np.int32)
        z2v = np.zeros(nvec, dtype=np.int32)
        self.ctx.check(self.ctx.lib.qbg_ef_basis(
            self.h, which, idx.ctypes.data_as(_ip), _pd(g2), _pd(gx),
            ip.ctypes.data_as(_ip) if maps else None, im.ctypes.data_as(_ip) if maps else None,
            z2v.ctypes.data_as(_ip) if maps else None))
        return dict(idx=idx.reshape(n, 3), g2=g2, gx=gx.reshape(3, n), ip=ip if maps else None,
                    im=im if maps else None, zvec_to_val=z2v if maps else None)

    def kset_basis(self, ik):
        """k-point ik: idx (ngw x 3), |k+G|^2, k+G (3 x ngw)."""
        n = self.kngw[ik]
        idx = np.zeros(3 * n, dtype=np.int32)
        kpg2 = np.zeros(n)
        kpgx = np.zeros(3 * n)
        self.ctx.check(self.ctx.lib.qbg_ef_kset_basis(self.h, ik, None, idx.ctypes.data_as(_ip),
                                                      _pd(kpg2), _pd(kpgx)))
        return dict(idx=idx.reshape(n, 3), kpg2=kpg2, kpgx=kpgx.reshape(3, n))

    def kset_pack(self, cs):
        """per-k-point (nst, ngw_k) arrays -> one (nkp nst, ngw) array, zero padded"""
        out = np.zeros((self.nkp * self.nst, self.ngw), dtype=np.complex128)
        for ik, c in enumerate(cs):
            out[ik * self.nst:(ik + 1) * self.nst, :self.kngw[ik]] = c
        return out

    def kset_unpack(self, a):
        return [np.ascontiguousarray(a[ik * self.nst:(ik + 1) * self.nst, :self.kngw[ik]])
                for ik in range(self.nkp)]

    def set_positions(self, tau):
        tau = np.ascontiguousarray(tau, dtype=np.float64).ravel()
        self.ctx.check(self.ctx.lib.qbg_ef_set_positions(self.h, _pd(tau)))

    def cell_moved(self, cell):
        """qbg_ef_cell_moved: Wavefunction::resize + EnergyFunctional::cell_moved
        for the new cell (rows a0 a1 a2) over the creation's reference cell;
        Cartesian positions unchanged (move them with set_positions)."""
        import dataclasses
        cell = np.ascontiguousarray(cell, dtype=np.float64)
        self.ctx.check(self.ctx.lib.qbg_ef_cell_moved(self.h, _pd(cell.ravel())))
        self.sysdef = dataclasses.replace(self.sysdef, cell=cell.copy())

    def update_density(self, c: DArray, prec="fp64"):
        self.ctx.check(self.ctx.lib.qbg_ef_update_density(self.h, c.p, self.ngw, PREC[prec]))

    def update_vhxc(self, stress=False):
        self.ctx.check(self.ctx.lib.qbg_ef_update_vhxc(self.h, int(stress)))

    def energy(self, c: DArray, hpsi: DArray = None, forces=True, stress=False, prec="fp64"):
        e = np.zeros(14)
        f = np.zeros(3 * self.nat)
        sig = np.zeros(42)
        self.ctx.check(self.ctx.lib.qbg_ef_energy(
            self.h, c.p, self.ngw, int(hpsi is not None), hpsi.p if hpsi is not None else None,
            self.ngw, int(forces), int(stress), PREC[prec], _pd(e), _pd(f), _pd(sig)))
        out = dict(zip(ENERGY_NAMES, e.tolist()))
        out["fion"] = f.reshape(-1, 3)
        out["sigma"] = sig.reshape(7, 6)
        return out

    def evaluate_async(self, c: DArray, hpsi: DArray = None, forces=True, stress=False,
                       prec="fp64"):
        """qbg_ef_evaluate_async: enqueue one evaluation (capturable)."""
        flags = (1 if hpsi is not None else 0) | (2 if forces else 0) | (4 if stress else 0)
        self.ctx.check(self.ctx.lib.qbg_ef_evaluate_async(
            self.h, c.p, self.ngw, hpsi.p if hpsi is not None else None, self.ngw, flags,
            PREC[prec]))

    def energy_async(self, c: DArray, hpsi: DArray = None, forces=False, stress=False,
                     prec="fp64", same_c=False):
        """qbg_ef_energy_async: enqueue EnergyFunctional::energy alone.
        same_c: c is unchanged since the last update_density (QBG_EF_SAME_C),
        so H*psi reuses the backward transforms update_density kept."""
        flags = (1 if hpsi is not None else 0) | (2 if forces else 0) | (4 if stress else 0)
        flags |= 8 if same_c else 0
        self.ctx.check(self.ctx.lib.qbg_ef_energy_async(
            self.h, c.p, self.ngw, hpsi.p if hpsi is not None else None, self.ngw, flags,
            PREC[prec]))

    def results(self):
        """qbg_ef_results: wait and return the last evaluation's results."""
        e = np.zeros(14)
        f = np.zeros(3 * self.nat)
        sig = np.zeros(42)
        self.ctx.check(self.ctx.lib.qbg_ef_results(self.h, _pd(e), _pd(f), _pd(sig)))
        out = dict(zip(ENERGY_NAMES, e.tolist()))
        out["fion"] = f.reshape(-1, 3)
        out["sigma"] = sig.reshape(7, 6)
        return out

    def evaluate(self, c: DArray, hpsi: DArray = None, forces=True, stress=False, prec="fp64"):
        """density -> update_vhxc -> energy (src/BOSampleStepper.cpp:1207-1217)
        as one call (qbg_ef_evaluate_async, then qbg_ef_results)."""
        self.evaluate_async(c, hpsi, forces, stress, prec)
        return self.results()

    def get(self, what):
        n = {0: self.N, 1: self.N, 2: self.N, 3: 2 * self.ngv, 4: 2 * self.ngv,
             5: 2 * self.ngv, 6: 2 * len(self.sysdef.species) * self.ngv,
             7: 3 * self.nat, 8: self.N, 9: self.N, 10: self.N, 11: self.N,
             12: 2 * self.ngv}[what]
        out = np.zeros(n)
        self.ctx.check(self.ctx.lib.qbg_ef_get(self.h, what, _pd(out)))
        if what in (3, 4, 5, 6, 12):
            return out.view(np.complex128)
        return out

    def wf_grid(self):
        """qbg_ef_wf_grid: the wavefunction basis on its own grid (sample I/O)."""
        if getattr(self, "_wfg", None) is None:
            h = _vp()
            dims = np.zeros(3, dtype=np.int32)
            self.ctx.check(self.ctx.lib.qbg_ef_wf_grid(self.h, dims.ctypes.data_as(_ip),
                                                       C.byref(h)))
            self._wfg = _BorrowedGrid(self.ctx, h, [int(x) for x in dims])
        return self._wfg

    def band_grid(self):
        """qbg_ef_band_grid: the grid of the band transforms (owned by the functional)."""
        h = _vp()
        self.ctx.check(self.ctx.lib.qbg_ef_band_grid(self.h, C.byref(h)))
        return _BorrowedGrid(self.ctx, h, [self.np0, self.np1, self.np2])

    def gemm_plan(self) -> dict:
        """qbg_ef_gemm_plan: the tuned launch plan of the BF16 projection product."""
        out = np.zeros(3, dtype=np.int32)
        self.ctx.check(self.ctx.lib.qbg_ef_gemm_plan(self.h, out.ctypes.data_as(_ip)))
        return dict(tuned=bool(out[0]), edge=int(out[1]), split=int(out[2]))

    def set_nspin(self, nspin, nst_up=0):
        """qbg_ef_set_nspin: bands [0, nst_up) spin up, the rest spin down."""
        self.ctx.check(self.ctx.lib.qbg_ef_set_nspin(self.h, int(nspin), int(nst_up)))

    def set_vr(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.ctx.check(self.ctx.lib.qbg_ef_set_vr(self.h, _pd(v)))


class NonLocalPotential:
    """qbg_nl: the NonLocalPotential of one spin's SlaterDet of an energy
    functional (src/NonLocalPotential.h:61-68)."""

    def __init__(self, ef: EnergyFunctional, ispin=0):
        self.ef = ef
        self.ctx = ef.ctx
        h = _vp()
        self.ctx.check(self.ctx.lib.qbg_nl_create(ef.h, int(ispin), C.byref(h)))
        self.h = h
        self.ctx.adopt(self)

    def update_twnl(self):
        self.ctx.check(self.ctx.lib.qbg_nl_update_twnl(self.h))

    def energy(self, c: DArray, dsd: DArray = None, forces=False, stress=False, prec="fp64",
               ldc=None, ldd=None):
        """NonLocalPotential::energy: returns (enl, fion (nat x 3), sigma_enl[6]);
        H*psi_NL is added into dsd when given."""
        f = np.zeros(3 * self.ef.nat)
        sig = np.zeros(6)
        enl = C.c_double(0.0)
        ldc = ldc or self.ef.ngw
        ldd = ldd or self.ef.ngw
        self.ctx.check(self.ctx.lib.qbg_nl_energy(
            self.h, c.p if isinstance(c, DArray) else c, ldc, int(dsd is not None),
            (dsd.p if isinstance(dsd, DArray) else dsd) if dsd is not None else None, ldd,
            int(forces), _pd(f), int(stress), _pd(sig), PREC[prec], C.byref(enl)))
        return enl.value, f.reshape(-1, 3), sig

    def release(self):
        if self.h and self.ctx.h:
            self.ctx.lib.qbg_nl_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class WfStepper:
    """qbg_wfstep: the SD, PSD and PSDA wavefunction steppers
    (src/SDWavefunctionStepper.cpp, src/PSDWavefunctionStepper.cpp,
    src/PSDAWavefunctionStepper.cpp).  param: dt2bye for SD, ecutprec
    (Hartree, 0 = adaptive) for PSD and PSDA."""
    KINDS = {"sd": 0, "psd": 1, "psda": 2}

    def __init__(self, ef: EnergyFunctional, kind, param=0.0, nst=None):
        """nst: a band set of that size (one spin's bands with nspin 2);
        None: all bands of ef."""
        self.ctx = ef.ctx
        self.ef = ef
        k = self.KINDS[kind] if isinstance(kind, str) else int(kind)
        h = _vp()
        if nst is None:
            self.ctx.check(self.ctx.lib.qbg_wfstep_create(ef.h, k, float(param), C.byref(h)))
        else:
            self.ctx.check(self.ctx.lib.qbg_wfstep_create_bands(ef.h, k, float(param), int(nst),
                                                                C.byref(h)))
        self.h = h
        self.ctx.adopt(self)

    def preprocess(self):
        self.ctx.check(self.ctx.lib.qbg_wfstep_preprocess(self.h))

    def update(self, c: DArray, dc: DArray, band0=0):
        """Enqueue WavefunctionStepper::update(dwf): dc holds H c on entry;
        band0: first column of this stepper's band set in c and dc."""
        off = 16 * self.ef.ngw * band0
        cp = c.offset(off) if band0 else c.p
        dp = dc.offset(off) if band0 else dc.p
        self.ctx.check(self.ctx.lib.qbg_wfstep_update(self.h, cp, self.ef.ngw, dp, self.ef.ngw))

    def results(self):
        out = np.zeros(4)
        self.ctx.check(self.ctx.lib.qbg_wfstep_results(self.h, _pd(out)))
        return dict(eigenvalue_sum=out[0], theta=out[1], polar_iterations=int(out[2]),
                    status=int(out[3]))

    def release(self):
        if self.h and self.ctx.h:
            self.ctx.lib.qbg_wfstep_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Mixer:
    """qbg_mixer: BOSampleStepper's charge mixing (Kerker weights, Anderson
    acceleration over ndim previous densities; ndim = 0 is simple mixing).
    With a spin-polarized functional both spins' densities are mixed as one
    vector; the functional's nspin must not change afterwards."""

    def __init__(self, ef: EnergyFunctional, alpha=0.5, rc_kerker=10.0, ndim=3, rc1=3.0):
        self.ctx = ef.ctx
        h = _vp()
        self.ctx.check(self.ctx.lib.qbg_mixer_create(ef.h, float(alpha), float(rc_kerker),
                                                     int(ndim), float(rc1), C.byref(h)))
        self.h = h
        self.ctx.adopt(self)

    def restart(self):
        self.ctx.check(self.ctx.lib.qbg_mixer_restart(self.h))

    def update(self):
        self.ctx.check(self.ctx.lib.qbg_mixer_update(self.h))

    def release(self):
        if self.h and self.ctx.h:
            self.ctx.lib.qbg_mixer_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def polar(ctx: Context, x, tol=1e-6, maxiter=20):
    """DoubleMatrix::polar of a square host matrix on the device; returns
    (orthogonal factor, iterations)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    d = ctx.to_device(np.ascontiguousarray(x.T))          # column major
    it = C.c_int(0)
    ctx.check(ctx.lib.qbg_polar(ctx.h, d.p, n, float(tol), int(maxiter), C.byref(it)))
    u = d.download().reshape(n, n).T.copy()
    d.release()
    return u, it.value

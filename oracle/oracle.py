"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REFLIB = HERE / "_ref" / "libqbref.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SpeciesDesc(C.Structure):
    _fields_ = [("zval", C.c_int), ("lmax", C.c_int), ("nmesh", C.c_int),
                ("deltar", C.c_double), ("rcps", C.c_double),
                ("vloc", _dp), ("proj", _dp), ("dl", _dp),
                ("type", C.c_int), ("llocal", C.c_int), ("vps", _dp), ("phi", _dp),
                ("nlcc", _dp), ("nquad", C.c_int), ("rquad", C.c_double)]


def species_desc(sp, keep):
    """SpeciesDesc of a configs.SpeciesDef (SLPP form) or an NCPP species
    (semilocal potentials vps, pseudo-wavefunctions phi, llocal)."""
    core = sp.core_density()
    if core is not None:
        core = np.ascontiguousarray(core, dtype=float)
        keep.append(core)
    if getattr(sp, "kind", "") == "ncpp":
        vps = np.ascontiguousarray(sp.vps, dtype=float)
        phi = np.ascontiguousarray(sp.phi, dtype=float)
        keep += [vps, phi]
        return SpeciesDesc(sp.zval, sp.lmax, sp.nmesh, sp.deltar, sp.rcps, None, None, None,
                           1, sp.llocal, _ptr(vps), _ptr(phi), _ptr(core), sp.nquad,
                           sp.rquad)
    vloc, proj, dl = sp.radial()
    keep += [vloc, proj, dl]
    return SpeciesDesc(sp.zval, sp.lmax, sp.nmesh, sp.deltar, sp.rcps,
                       _ptr(vloc), _ptr(proj), _ptr(dl), 0, 0, None, None, _ptr(core), 0, 0.0)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle library not built: {LIB}")
        _lib = C.CDLL(str(LIB))
        _lib.qbo_create_k.restype = C.c_void_p
        _lib.qbo_create_k.argtypes = [_dp, _dp, C.c_double, C.c_int, C.POINTER(SpeciesDesc),
                                      _ip, _dp, C.c_int, _dp, _ip, _dp, C.c_int]
        _lib.qbo_create_kset.restype = C.c_void_p
        _lib.qbo_create_kset.argtypes = [_dp, _dp, C.c_double, C.c_int, C.POINTER(SpeciesDesc),
                                         _ip, _dp, C.c_int, _dp, _ip, C.c_int, _dp, _dp]
        _lib.qbo_kset_basis.restype = C.c_int
        _lib.qbo_kset_basis.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp]
        _lib.qbo_kpg.restype = C.c_int
        _lib.qbo_kpg.argtypes = [C.c_void_p, _dp, _dp]
        _lib.qbo_create.restype = C.c_void_p
        _lib.qbo_create.argtypes = [_dp, _dp, C.c_double, C.c_int, C.POINTER(SpeciesDesc),
                                    _ip, _dp, C.c_int, _dp, _ip]
        _lib.qbo_last_error.restype = C.c_char_p
        _lib.qbo_ncpp_kb.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        for name in ["qbo_destroy", "qbo_sizes", "qbo_basis", "qbo_rods", "qbo_fft_backward",
                     "qbo_fft_backward_pair", "qbo_fft_forward", "qbo_fft_forward_pair",
                     "qbo_compute_density", "qbo_rs_mul_add", "qbo_update_density",
                     "qbo_update_vhxc", "qbo_get", "qbo_set_vr", "qbo_species_eval"]:
            getattr(_lib, name).argtypes = None
        _lib.qbo_gram.argtypes = [_dp, C.c_int, C.c_int]
        _lib.qbo_gram_complex.argtypes = [_dp, C.c_int, C.c_int]
        _lib.qbo_species_core.restype = C.c_double
        _lib.qbo_species_core.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        _lib.qbo_sd_dot.restype = C.c_double
        _lib.qbo_sd_dot.argtypes = [_dp, _dp, C.c_int, C.c_int]
        _lib.qbo_prec_ekin.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        _lib.qbo_prec_diag.restype = C.c_double
        _lib.qbo_prec_diag.argtypes = [C.c_double, C.c_double, C.c_double]
        _lib.qbo_residual.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        _lib.qbo_polar.argtypes = [_dp, C.c_int, C.c_double, C.c_int]
        _lib.qbo_ortho_align.argtypes = [_dp, _dp, C.c_int, C.c_int]
        _lib.qbo_wf_update.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, C.c_int, C.c_int,
                                       _dp, _dp, _ip]
        _lib.qbo_anderson_create.restype = C.c_void_p
        _lib.qbo_anderson_create.argtypes = [C.c_long, C.c_int]
        _lib.qbo_anderson_update.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        _lib.qbo_anderson_destroy.argtypes = [C.c_void_p]
        _lib.qbo_mixer_create.restype = C.c_void_p
        _lib.qbo_mixer_create.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int,
                                          C.c_double]
        _lib.qbo_mixer_restart.argtypes = [C.c_void_p]
        _lib.qbo_mixer_update.argtypes = [C.c_void_p, C.c_void_p]
        _lib.qbo_mixer_destroy.argtypes = [C.c_void_p]
        _lib.qbo_update_rhor.argtypes = [C.c_void_p]
        _lib.qbo_energy.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_int,
                                    _dp, _dp, _dp]
        _lib.qbo_nlm.argtypes = [C.c_void_p, C.c_int]
        _lib.qbo_nop.argtypes = [C.c_void_p, C.c_int]
        _lib.qbo_nl_energy.restype = C.c_double
        _lib.qbo_nl_energy.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_int,
                                       _dp, _dp]
    return _lib


ENERGY_NAMES = ["ekin", "eps", "enl", "ehart", "ehart_e", "ehart_ep", "ehart_p",
                "exc", "dxc", "esr", "eself", "ecoul", "etotal", "total_charge"]


class Oracle:
    """One Qbox-like system on the CPU (EnergyFunctional + ChargeDensity)."""

    def __init__(self, sysdef, refcell=None, grid=None, kpoint=None, force_complex=False,
                 kpoints=None, weights=None):
        """refcell (3x3, rows a0 a1 a2): the reference cell that fixes the
        basis membership and grid sizes (src/Basis.cpp:314-325), so that
        strained copies of a system share one basis.  grid (np0, np1, np2):
        the density grid, as FourierTransform(basis, np0, np1, np2) takes it
        (src/FourierTransform.h:105), instead of ChargeDensity's rule; it must
        hold the density basis.  kpoint (3, in units of the reciprocal
        vectors; default the system's own `kpoint` attribute or 0): k != 0
        gives the complex (full-sphere) wavefunction basis, with one weight-1
        k-point; force_complex builds that basis at k = 0 as well.  kpoints
        (nkp x 3) and weights (nkp, summing to 1): a k-point set, every
        k-point on the complex basis; the coefficients are then a list of
        (nst, ngw_k) arrays, one per k-point."""
        L = lib()
        self.sysdef = sysdef
        if refcell is None:
            refcell = getattr(sysdef, "refcell", None)
        self._keep = []
        descs = (SpeciesDesc * len(sysdef.species))()
        for i, sp in enumerate(sysdef.species):
            descs[i] = species_desc(sp, self._keep)
        cell = np.ascontiguousarray(sysdef.cell, dtype=float).ravel()
        na = np.array(sysdef.na, dtype=np.int32)
        tau = np.ascontiguousarray(sysdef.tau, dtype=float).ravel()
        occ = sysdef.occ.astype(float)
        self._keep += [cell, na, tau, occ, descs]
        ref = None
        if refcell is not None:
            ref = np.ascontiguousarray(refcell, dtype=float).ravel()
            self._keep.append(ref)
        g3 = None
        if grid is not None:
            g3 = np.array(grid, dtype=np.int32)
            self._keep.append(g3)
        if kpoint is None:
            kpoint = getattr(sysdef, "kpoint", None)
        kp = None
        if kpoint is not None:
            kp = np.ascontiguousarray(kpoint, dtype=float).ravel()
            self._keep.append(kp)
        self.nkp = 0
        if kpoints is not None:
            kps = np.ascontiguousarray(kpoints, dtype=float).reshape(-1, 3)
            wts = np.ascontiguousarray(weights, dtype=float).ravel()
            self.nkp = kps.shape[0]
            self._keep += [kps, wts]
            h = L.qbo_create_kset(_ptr(cell), _ptr(ref) if ref is not None else None, sysdef.ecut,
                                  len(sysdef.species), descs, _ptr(na, _ip), _ptr(tau),
                                  sysdef.nst, _ptr(occ),
                                  _ptr(g3, _ip) if g3 is not None else None, self.nkp,
                                  _ptr(kps), _ptr(wts))
        else:
            h = L.qbo_create_k(_ptr(cell), _ptr(ref) if ref is not None else None, sysdef.ecut,
                               len(sysdef.species), descs,
                               _ptr(na, _ip), _ptr(tau), sysdef.nst, _ptr(occ),
                               _ptr(g3, _ip) if g3 is not None else None,
                               _ptr(kp) if kp is not None else None, 1 if force_complex else 0)
        if not h:
            raise RuntimeError(L.qbo_last_error().decode())
        self.h = C.c_void_p(h)
        s = np.zeros(12, dtype=np.int32)
        L.qbo_sizes(self.h, _ptr(s, _ip))
        (self.ngw, self.ngv, self.np0, self.np1, self.np2, self.nvec, self.ntrans0,
         self.nvec_v, self.ntrans0_v, self.nrods, self.nrods_v, self.ndft) = [int(x) for x in s]
        self.N = self.np0 * self.np1 * self.np2
        self.nst = sysdef.nst
        if getattr(sysdef, "nspin", 1) == 2:
            if L.qbo_set_nspin(self.h, 2, int(sysdef.nst_up)) != 0:
                raise RuntimeError("bad nspin / nst_up")

    def __del__(self):
        try:
            lib().qbo_destroy(self.h)
        except Exception:
            pass

    # ------------------------------------------------------------- bases
    def basis(self, which=0):
        n = self.ngw if which == 0 else self.ngv
        nvec = self.nvec if which == 0 else self.nvec_v
        idx = np.zeros(3 * n, dtype=np.int32)
        g2 = np.zeros(n)
        gx = np.zeros(3 * n)
        ip = np.zeros(n, dtype=np.int32)
        im = np.zeros(n, dtype=np.int32)
        z2v = np.zeros(nvec, dtype=np.int32)
        npb = np.zeros(9, dtype=np.int32)
        lib().qbo_basis(self.h, which, _ptr(idx, _ip), _ptr(g2), _ptr(gx), _ptr(ip, _ip),
                        _ptr(im, _ip), _ptr(z2v, _ip), _ptr(npb, _ip))
        return dict(idx=idx.reshape(n, 3), g2=g2, gx=gx.reshape(3, n), ip=ip, im=im,
                    zvec_to_val=z2v, np=npb[:3].copy(), idxmin=npb[3:6].copy(),
                    idxmax=npb[6:9].copy())

    def kpg(self):
        """|k+G|^2 (ngw) and k+G (3 x ngw) of the wavefunction basis, and
        whether the basis is real."""
        kpg2 = np.zeros(self.ngw)
        kpgx = np.zeros(3 * self.ngw)
        real = lib().qbo_kpg(self.h, _ptr(kpg2), _ptr(kpgx))
        return kpg2, kpgx.reshape(3, self.ngw), bool(real)

    def rods(self, which=0):
        n = self.nrods if which == 0 else self.nrods_v
        a = [np.zeros(n, dtype=np.int32) for _ in range(4)]
        lib().qbo_rods(self.h, which, *[_ptr(x, _ip) for x in a])
        return a

    # ---------------------------------------------------------- transforms
    def fft_backward(self, c, which=0):
        f = np.zeros(self.N, dtype=np.complex128)
        c = np.ascontiguousarray(c, dtype=np.complex128)
        lib().qbo_fft_backward(self.h, which, _ptr(c), _ptr(f))
        return f

    def fft_backward_pair(self, c1, c2):
        f = np.zeros(self.N, dtype=np.complex128)
        c1 = np.ascontiguousarray(c1, dtype=np.complex128)
        c2 = np.ascontiguousarray(c2, dtype=np.complex128)
        lib().qbo_fft_backward_pair(self.h, _ptr(c1), _ptr(c2), _ptr(f))
        return f

    def fft_forward(self, f, which=0):
        n = self.ngw if which == 0 else self.ngv
        f = np.array(f, dtype=np.complex128)
        c = np.zeros(n, dtype=np.complex128)
        lib().qbo_fft_forward(self.h, which, _ptr(f), _ptr(c))
        return c

    def fft_forward_pair(self, f):
        f = np.array(f, dtype=np.complex128)
        c1 = np.zeros(self.ngw, dtype=np.complex128)
        c2 = np.zeros(self.ngw, dtype=np.complex128)
        lib().qbo_fft_forward_pair(self.h, _ptr(f), _ptr(c1), _ptr(c2))
        return c1, c2

    # ------------------------------------------------------------ physics
    def compute_density(self, c):
        c = np.ascontiguousarray(c, dtype=np.complex128)
        rho = np.zeros(self.N)
        lib().qbo_compute_density(self.h, _ptr(c), _ptr(rho))
        return rho

    def rs_mul_add(self, v, c, hpsi=None):
        c = np.ascontiguousarray(c, dtype=np.complex128)
        v = np.ascontiguousarray(v, dtype=float)
        h = np.zeros_like(c) if hpsi is None else np.array(hpsi, dtype=np.complex128)
        lib().qbo_rs_mul_add(self.h, _ptr(v), _ptr(c), c.shape[0], _ptr(h))
        return h

    def kset_basis(self, ik):
        """k-point ik of a set: idx (ngw x 3), |k+G|^2, k+G (3 x ngw)."""
        n = lib().qbo_kset_basis(self.h, ik, None, None, None)
        idx = np.zeros(3 * n, dtype=np.int32)
        kpg2 = np.zeros(n)
        kpgx = np.zeros(3 * n)
        lib().qbo_kset_basis(self.h, ik, _ptr(idx, _ip), _ptr(kpg2), _ptr(kpgx))
        return dict(idx=idx.reshape(n, 3), kpg2=kpg2, kpgx=kpgx.reshape(3, n))

    def _pack(self, c):
        """a k-point set's per-k coefficient list as one buffer"""
        if self.nkp:
            return np.ascontiguousarray(np.concatenate([np.ravel(x) for x in c]),
                                        dtype=np.complex128)
        return np.ascontiguousarray(c, dtype=np.complex128)

    def update_density(self, c):
        c = self._pack(c)
        lib().qbo_update_density(self.h, _ptr(c))

    def update_vhxc(self, stress=False):
        lib().qbo_update_vhxc(self.h, int(stress))

    def get(self, what, isp=0):
        npr = lib().qbo_nlm(self.h, isp)
        sizes = {0: (self.N, float), 1: (self.N, float), 2: (self.N, float),
                 3: (self.ngv, complex), 4: (self.ngv, complex), 5: (self.ngv, complex),
                 6: (self.ngv, float), 7: (self.ngv, float), 8: (self.ngv, float),
                 9: (npr * self.ngw, float), 10: (self.ngv, complex),
                 11: (3 * self.sysdef.nat, float), 12: (npr * 6 * self.ngw, float),
                 13: (self.ngv, complex), 14: (self.N, float), 15: (self.N, float),
                 16: (self.N, float), 17: (self.N, float), 18: (self.ngv, complex),
                 19: (self.ngv, float)}
        n, t = sizes[what]
        out = np.zeros(n, dtype=np.complex128 if t is complex else np.float64)
        lib().qbo_get(self.h, what, isp, out.ctypes.data_as(_dp))
        return out

    def set_vr(self, v):
        v = np.ascontiguousarray(v, dtype=float)
        lib().qbo_set_vr(self.h, _ptr(v))

    def energy(self, c, hpsi=True, forces=True, stress=False):
        shapes = [np.shape(x) for x in c] if self.nkp else None
        c = self._pack(c)
        h = np.zeros_like(c) if hpsi else None
        e = np.zeros(14)
        f = np.zeros(3 * self.sysdef.nat)
        sig = np.zeros(42)
        rc = lib().qbo_energy(self.h, _ptr(c), int(hpsi), _ptr(h), int(forces), int(stress),
                              _ptr(e), _ptr(f), _ptr(sig))
        if rc != 0:
            raise RuntimeError(lib().qbo_last_error().decode())
        out = dict(zip(ENERGY_NAMES, e.tolist()))
        out["fion"] = f.reshape(-1, 3)
        out["sigma"] = sig.reshape(7, 6)
        if h is not None and shapes is not None:
            parts, off = [], 0
            for sh in shapes:
                n = sh[0] * sh[1]
                parts.append(h[off:off + n].reshape(sh))
                off += n
            h = parts
        out["hpsi"] = h
        return out

    def nl_energy(self, c, hpsi=None, forces=True, stress=True):
        """NonLocalPotential::energy: (enl, fion (nat x 3), sigma_enl[6],
        hpsi + H*psi_NL or None)."""
        c = np.ascontiguousarray(c, dtype=np.complex128)
        h = None if hpsi is None else np.array(hpsi, dtype=np.complex128)
        f = np.zeros(3 * self.sysdef.nat)
        sig = np.zeros(6)
        enl = lib().qbo_nl_energy(self.h, _ptr(c), int(h is not None), _ptr(h), int(forces),
                                  int(stress), _ptr(f), _ptr(sig))
        return enl, f.reshape(-1, 3), sig, h

    def evaluate(self, c, hpsi=True, forces=True, stress=False):
        """density -> update_vhxc -> energy (src/BOSampleStepper.cpp:1207-1217)."""
        self.update_density(c)
        self.update_vhxc(stress)
        return self.energy(c, hpsi, forces, stress)

    def update_rhor(self):
        """ChargeDensity::update_rhor: rhor from the current rhog."""
        lib().qbo_update_rhor(self.h)

    def mixer(self, alpha=0.5, rc_kerker=10.0, ndim=3, rc1=3.0):
        return Mixer(self, alpha, rc_kerker, ndim, rc1)

    def species_core(self, isp, g):
        """(rhocoreg, drhocoreg, zcore) of species isp at |G| = g: the core
        density's transform (Species::drhocoreg) and its integral."""
        g = np.ascontiguousarray(g, dtype=float)
        rho, drho = np.zeros(g.size), np.zeros(g.size)
        z = lib().qbo_species_core(self.h, isp, g.size, _ptr(g), _ptr(rho), _ptr(drho))
        return rho, drho, z

    def species_eval(self, isp, g, nop=None):
        """(vlocg, dvlocg, vnlg[l], dvnlg[l]) of species isp at |G| = g; one
        row per projector channel l = 0 .. lmax."""
        g = np.ascontiguousarray(g, dtype=float)
        n = g.size
        have = lib().qbo_nop(self.h, isp)
        if nop is not None and nop != have:
            raise ValueError(f"species {isp} has {have} projector channels")
        nop = have
        v, dv = np.zeros(n), np.zeros(n)
        vn, dvn = np.zeros(nop * n), np.zeros(nop * n)
        lib().qbo_species_eval(self.h, isp, n, _ptr(g), _ptr(v), _ptr(dv), _ptr(vn), _ptr(dvn))
        return v, dv, vn.reshape(nop, n), dvn.reshape(nop, n)


def ncpp_kb(sp):
    """(projectors (lkb+1, nmesh), D_l) of an NCPP species in the
    one-projector-per-l form (Species::initialize_ncpp, nquad = 0)."""
    vps = np.ascontiguousarray(sp.vps, dtype=float)
    phi = np.ascontiguousarray(sp.phi, dtype=float)
    proj = np.zeros((sp.lmax + 1) * sp.nmesh)
    dl = np.zeros(sp.lmax + 1)
    lkb = lib().qbo_ncpp_kb(sp.lmax, sp.llocal, sp.nmesh, sp.deltar, _ptr(vps), _ptr(phi),
                            _ptr(proj), _ptr(dl))
    if lkb < 0:
        raise RuntimeError(lib().qbo_last_error().decode())
    return proj[:(lkb + 1) * sp.nmesh].reshape(lkb + 1, sp.nmesh), dl[:lkb + 1]


def compute_density_occ(o, c, occ):
    """SlaterDet::compute_density for arbitrary bands/occupations of o's basis."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    occ = np.ascontiguousarray(occ, dtype=float)
    rho = np.zeros(o.N)
    lib().qbo_compute_density_occ(o.h, _ptr(c), c.shape[0], _ptr(occ), _ptr(rho))
    return rho


class Mixer:
    """Charge mixing of BOSampleStepper (Kerker weights, AndersonMixer when
    ndim > 0) on an Oracle's rhog/rhor; update() follows update_density()."""

    def __init__(self, o, alpha, rc_kerker, ndim, rc1):
        self.o = o
        self.h = lib().qbo_mixer_create(o.h, alpha, rc_kerker, ndim, rc1)

    def restart(self):
        lib().qbo_mixer_restart(self.h)

    def update(self):
        lib().qbo_mixer_update(self.h, self.o.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qbo_mixer_destroy(self.h)
            self.h = None


def _c2(c):
    return np.array(c, dtype=np.complex128, order="C")


def sd_dot(c, d):
    """SlaterDet::dot for (nst, ngw) coefficient arrays (real basis)."""
    c, d = _c2(c), _c2(d)
    return lib().qbo_sd_dot(_ptr(c), _ptr(d), c.shape[1], c.shape[0])


def prec_ekin(c, g2):
    c = _c2(c)
    g2 = np.ascontiguousarray(g2, dtype=float)
    e = np.zeros(c.shape[0])
    lib().qbo_prec_ekin(_ptr(c), _ptr(g2), c.shape[1], c.shape[0], _ptr(e))
    return e


def prec_diag(ekin, g2, ecutprec=0.0):
    return lib().qbo_prec_diag(ekin, g2, ecutprec)


def residual(c, d):
    """(d - c A, A) with A = 2 c^T d - c0 d0^T (PSD/PSDA residual)."""
    c, d = _c2(c), _c2(d)
    nst = c.shape[0]
    a = np.zeros(nst * nst)
    lib().qbo_residual(_ptr(c), _ptr(d), c.shape[1], nst, _ptr(a))
    return d, a.reshape(nst, nst).T


def polar(x, tol=1e-6, maxiter=20):
    """DoubleMatrix::polar of a square matrix; returns (factor, iterations)."""
    xc = np.array(np.asarray(x, dtype=float).T, order="C")   # column major
    it = lib().qbo_polar(_ptr(xc), xc.shape[0], tol, maxiter)
    if it < 0:
        raise RuntimeError("polar: singular matrix")
    return xc.T.copy(), it


def ortho_align(c, cl):
    c, cl = _c2(c), _c2(cl)
    rc = lib().qbo_ortho_align(_ptr(c), _ptr(cl), c.shape[1], c.shape[0])
    if rc != 0:
        raise RuntimeError(f"ortho_align failed ({rc})")
    return c


class WfStepper:
    """SD (kind 0, param dt2bye), PSD (1) and PSDA (2, param ecutprec)
    wavefunction steppers; PSDA state is kept between updates."""
    KINDS = {"sd": 0, "psd": 1, "psda": 2}

    def __init__(self, kind, param, g2):
        self.kind = self.KINDS[kind] if isinstance(kind, str) else int(kind)
        self.param = float(param)
        self.g2 = np.ascontiguousarray(g2, dtype=float)
        self.c_last = None
        self.d_last = None
        self.extrapolate = C.c_int(0)

    def preprocess(self):
        self.extrapolate.value = 0

    def update(self, c, d):
        """Returns (c_new, d_new): d holds H c on entry."""
        c, d = _c2(c), _c2(d)
        if self.c_last is None:
            self.c_last = np.zeros_like(c)
            self.d_last = np.zeros_like(c)
        rc = lib().qbo_wf_update(self.kind, self.param, _ptr(self.g2), _ptr(c), _ptr(d),
                                 c.shape[1], c.shape[0], _ptr(self.c_last), _ptr(self.d_last),
                                 C.byref(self.extrapolate))
        if rc != 0:
            raise RuntimeError(f"wavefunction update failed ({rc})")
        return c, d


class Anderson:
    def __init__(self, m, nmax):
        self.m = m
        self.h = lib().qbo_anderson_create(m, nmax)

    def update(self, x, f):
        x = np.ascontiguousarray(x, dtype=float)
        f = np.ascontiguousarray(f, dtype=float)
        xb, fb = np.zeros(self.m), np.zeros(self.m)
        lib().qbo_anderson_update(self.h, _ptr(x), _ptr(f), _ptr(xb), _ptr(fb))
        return xb, fb

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qbo_anderson_destroy(self.h)
            self.h = None


def gram(c):
    """SlaterDet::gram on (nst, ngw) complex coefficients (in place copy)."""
    c = np.array(c, dtype=np.complex128, order="C")
    nst, ngw = c.shape
    rc = lib().qbo_gram(_ptr(c), ngw, nst)
    if rc != 0:
        raise RuntimeError(f"overlap not positive definite at pivot {rc}")
    return c


def gram_complex(c):
    """SlaterDet::gram of the complex (k != 0) basis on (nst, ngw)
    coefficients: S = C^H C = L L^H, C <- C L^-H (a copy)."""
    c = np.array(c, dtype=np.complex128, order="C")
    nst, ngw = c.shape
    rc = lib().qbo_gram_complex(_ptr(c), ngw, nst)
    if rc != 0:
        raise RuntimeError(f"overlap not positive definite at pivot {rc}")
    return c


def lda(rho):
    rho = np.ascontiguousarray(rho, dtype=float)
    e, v = np.zeros_like(rho), np.zeros_like(rho)
    lib().qbo_lda(rho.size, _ptr(rho), _ptr(e), _ptr(v))
    return e, v


def lsda(rup, rdn):
    """Spin-polarized LDA per point: (exc, v_up, v_dn)."""
    rup = np.ascontiguousarray(rup, dtype=float)
    rdn = np.ascontiguousarray(rdn, dtype=float)
    e, vu, vd = np.zeros_like(rup), np.zeros_like(rup), np.zeros_like(rup)
    lib().qbo_lsda(rup.size, _ptr(rup), _ptr(rdn), _ptr(e), _ptr(vu), _ptr(vd))
    return e, vu, vd


def sinft(y):
    y = np.array(y, dtype=float)
    lib().qbo_sinft(y.size, _ptr(y))
    return y


def cosft1(y):
    """y has n+1 entries; returns the transformed first n (y[n] unchanged)."""
    y = np.array(y, dtype=float)
    lib().qbo_cosft1(y.size - 1, _ptr(y))
    return y


def spline(x, y, yp1, ypn, bl, br):
    x = np.ascontiguousarray(x, dtype=float)
    y = np.ascontiguousarray(y, dtype=float)
    y2 = np.zeros_like(x)
    lib().qbo_spline(x.size, _ptr(x), _ptr(y), C.c_double(yp1), C.c_double(ypn), bl, br, _ptr(y2))
    return y2


def splintd(xa, ya, y2a, x):
    y, dy = C.c_double(), C.c_double()
    xa = np.ascontiguousarray(xa, dtype=float)
    ya = np.ascontiguousarray(ya, dtype=float)
    y2a = np.ascontiguousarray(y2a, dtype=float)
    lib().qbo_splintd(xa.size, _ptr(xa), _ptr(ya), _ptr(y2a), C.c_double(x), C.byref(y), C.byref(dy))
    return y.value, dy.value


def fft1d(x, sign):
    x = np.array(x, dtype=np.complex128)
    lib().qbo_fft1d(_ptr(x), x.size, sign)
    return x


def cell(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    b = np.zeros(9)
    vol = C.c_double()
    lib().qbo_cell(_ptr(a), _ptr(b), C.byref(vol))
    return b.reshape(3, 3), vol.value


def fold_in_ws(a, v):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    v = np.array(v, dtype=float)
    lib().qbo_fold_in_ws(_ptr(a), _ptr(v))
    return v


# --------------------------------------------------------------- reference
_ref = None


def reflib():
    """The reference's own spline/LDA/UnitCell sources compiled in _ref/."""
    global _ref
    if _ref is None:
        if not REFLIB.exists():
            return None
        _ref = C.CDLL(str(REFLIB))
    return _ref

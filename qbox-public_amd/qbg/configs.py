"""Synthetic Qbox inputs: cells, SLPP species tables, atom positions and
Gamma-point wavefunction coefficients.

A species is given as Qbox's ``norm_conserving_semilocal_pseudopotential``
would hold it (src/SpeciesHandler.cpp:200-241): a local potential
v_loc(r) = -Z erf(r/rc)/r and one projector per l (l = 0, 1) with radial form
r^l exp(-r^2 / (2 sigma^2)), on a linear mesh, plus D_l.  Qbox reads species
through a 12-digit scientific print (src/Species.cpp:1240-1241 via
src/SpeciesCmd.cpp:42-58), so every radial value is rounded the same way here.
``rcps`` is fixed at 1.5 bohr for every species (src/AtomSet.cpp:43-44).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

RCPS = 1.5          # src/AtomSet.cpp:43-44
DELTAR = 0.01       # mesh spacing (bohr)
NMESH = 2000        # radial points


def _round12(x: np.ndarray) -> np.ndarray:
    """Round through '%.12e' like Species::print followed by a re-read."""
    return np.array([float("%.12e" % v) for v in np.asarray(x, dtype=float)])


@dataclass
class SpeciesDef:
    zval: int
    rc: float               # erf width of the local potential
    sigma: float            # Gaussian width of the projectors
    d: tuple = (2.0, -1.0)  # D_0, D_1 (Ha)
    deltar: float = DELTAR
    nmesh: int = NMESH
    rcps: float = RCPS
    # non-linear core correction: a Gaussian core density A exp(-(r/rcore)^2),
    # given as (A, rcore); () for none
    core: tuple = ()

    @property
    def lmax(self) -> int:
        return len(self.d) - 1

    def core_density(self):
        """The core density on the radial mesh (Species nlcc_), or None."""
        if not self.core:
            return None
        a, rcore = self.core
        r = np.arange(self.nmesh) * self.deltar
        return _round12(a * np.exp(-(r / rcore) ** 2))

    def radial(self):
        """Return (vloc[nmesh], proj[(lmax+1)*nmesh], dl[lmax+1])."""
        r = np.arange(self.nmesh) * self.deltar
        v = np.empty_like(r)
        v[0] = -self.zval * 2.0 / (math.sqrt(math.pi) * self.rc)
        rr = r[1:]
        v[1:] = -self.zval * np.array([math.erf(x) for x in rr / self.rc]) / rr
        g = np.exp(-r * r / (2.0 * self.sigma ** 2))
        proj = [g] + [r ** l * g for l in range(1, self.lmax + 1)]
        vloc = _round12(v)
        proj = np.concatenate([_round12(p) for p in proj])
        return vloc, proj, np.array(self.d, dtype=float)


SPECIES_A = SpeciesDef(zval=4, rc=1.0, sigma=0.8)
SPECIES_B = SpeciesDef(zval=6, rc=0.7, sigma=0.6)
# with a d projector r^2 exp(-r^2 / (2 sigma^2)), D_2 = 0.7 Ha
SPECIES_D = SpeciesDef(zval=4, rc=1.0, sigma=0.8, d=(2.0, -1.0, 0.7))
# with d and f projectors, r^3 exp(-r^2 / (2 sigma^2)) with D_3 = 0.5 Ha
SPECIES_F = SpeciesDef(zval=4, rc=1.0, sigma=0.8, d=(2.0, -1.0, 0.7, 0.5))
# with a core density 0.4 exp(-(r/0.9)^2) (about 0.75 electrons)
SPECIES_C = SpeciesDef(zval=4, rc=1.0, sigma=0.8, d=(2.0, -1.0, 0.7), core=(0.4, 0.9))


@dataclass
class SystemDef:
    name: str
    cell: np.ndarray                 # 3x3, rows are a0, a1, a2 (bohr)
    ecut_ry: float
    species: list                    # list[SpeciesDef]
    na: list                         # atoms per species
    nst: int
    seed: int
    min_sep: float
    tau: np.ndarray = field(default=None)   # (nat, 3), species-major
    # nspin 2: bands [0, nst_up) are spin up, the rest spin down; occupations
    # per band (at most 1 with two spins), Wavefunction nst_[ispin] and
    # SlaterDet occ_ (src/Wavefunction.cpp:326-336, 424-447)
    nspin: int = 1
    nst_up: int = 0
    occ_bands: np.ndarray = field(default=None)
    # Wavefunction's reference cell (rows a0 a1 a2), which fixes the basis
    # membership and the grids (src/Basis.cpp:314-325); None: the cell itself
    refcell: np.ndarray = field(default=None)
    # one k-point in units of the reciprocal vectors b0 b1 b2 (weight 1);
    # None or 0: Gamma (the real basis); k != 0 gives the complex full-sphere
    # basis (src/Basis.cpp:237, 446-481)
    kpoint: np.ndarray = field(default=None)

    @property
    def ecut(self) -> float:         # Hartree (src/Ecut.h:38-57)
        return 0.5 * self.ecut_ry

    @property
    def nat(self) -> int:
        return int(sum(self.na))

    @property
    def occ(self) -> np.ndarray:
        if self.occ_bands is not None:
            return np.asarray(self.occ_bands, dtype=float)
        return np.full(self.nst, 2.0)

    def build_positions(self) -> "SystemDef":
        rng = np.random.Generator(np.random.PCG64(self.seed))
        self.tau = random_positions(rng, self.cell, self.nat, self.min_sep)
        return self


def random_positions(rng, cell, nat, min_sep, max_tries=200000):
    """Uniform positions in the cell, rejection-sampled to a minimum-image
    separation of at least ``min_sep``."""
    A = np.asarray(cell, dtype=float)
    Ainv = np.linalg.inv(A)
    pos = []
    tries = 0
    while len(pos) < nat:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms at this separation")
        f = rng.random(3)
        x = f @ A
        ok = True
        for y in pos:
            d = x - y
            s = d @ Ainv
            s -= np.round(s)
            if np.linalg.norm(s @ A) < min_sep:
                ok = False
                break
        if ok:
            pos.append(x)
    return np.array(pos)


def cubic(L):
    return np.diag([L, L, L]).astype(float)


BASE_SEED = 20261202


def config(index: int) -> SystemDef:
    """The five workloads (index 0..4) plus small test systems (index >= 10)."""
    if index == 0:
        s = SystemDef("c1: 8 atoms L=16 ecut=30Ry 16 bands", cubic(16.0), 30.0,
                      [SPECIES_A], [8], 16, BASE_SEED + 0, 2.5)
    elif index == 1:
        s = SystemDef("c2: 64 atoms L=24 ecut=60Ry 128 bands", cubic(24.0), 60.0,
                      [SPECIES_A], [64], 128, BASE_SEED + 1, 2.5)
    elif index == 2:
        s = SystemDef("c3: 256 atoms L=36 ecut=70Ry 512 bands", cubic(36.0), 70.0,
                      [SPECIES_A, SPECIES_B], [128, 128], 512, BASE_SEED + 2, 2.2)
    elif index == 4:
        s = SystemDef("c5: 128 atoms 28x30x32 ecut=60Ry 256 bands",
                      np.diag([28.0, 30.0, 32.0]), 60.0,
                      [SPECIES_A, SPECIES_B], [64, 64], 256, BASE_SEED + 4, 2.2)
    elif index == 10:
        s = SystemDef("t0: 2 atoms L=8 ecut=12Ry 4 bands", cubic(8.0), 12.0,
                      [SPECIES_A], [2], 4, BASE_SEED + 10, 2.0)
    elif index == 11:
        s = SystemDef("t1: 2+1 atoms 7x8x9 ecut=16Ry 5 bands", np.diag([7.0, 8.0, 9.0]),
                      16.0, [SPECIES_A, SPECIES_B], [2, 1], 5, BASE_SEED + 11, 2.0)
    elif index == 13:
        # d projectors (l = 2) beside an s/p species, orthorhombic cell
        s = SystemDef("t3: 2+1 atoms with d projectors 8x8.5x9 ecut=18Ry 6 bands",
                      np.diag([8.0, 8.5, 9.0]), 18.0, [SPECIES_D, SPECIES_B], [2, 1], 6,
                      BASE_SEED + 13, 2.0)
    elif index == 14:
        # f projectors (l = 3) beside an s/p species, orthorhombic cell
        s = SystemDef("t4: 2+1 atoms with f projectors 8x8.5x9 ecut=18Ry 6 bands",
                      np.diag([8.0, 8.5, 9.0]), 18.0, [SPECIES_F, SPECIES_B], [2, 1], 6,
                      BASE_SEED + 14, 2.0)
    elif index == 15:
        # core densities (non-linear core correction) on a d species beside a
        # species without one, orthorhombic cell
        s = SystemDef("t5: 2+1 atoms with core densities 8x8.5x9 ecut=18Ry 6 bands",
                      np.diag([8.0, 8.5, 9.0]), 18.0, [SPECIES_C, SPECIES_B], [2, 1], 6,
                      BASE_SEED + 15, 2.0)
    elif index == 16:
        # one k-point off Gamma (complex full-sphere basis), d projectors
        # beside an s/p species, orthorhombic cell
        s = SystemDef("t6: 2+1 atoms at k = (0.25, 0.1, -0.2) 8x8.5x9 ecut=18Ry 6 bands",
                      np.diag([8.0, 8.5, 9.0]), 18.0, [SPECIES_D, SPECIES_B], [2, 1], 6,
                      BASE_SEED + 16, 2.0, kpoint=np.array([0.25, 0.1, -0.2]))
    elif index == 12:
        # non-orthogonal (monoclinic) cell
        a = np.array([[9.0, 0.0, 0.0], [1.5, 8.5, 0.0], [0.0, 0.7, 8.0]])
        s = SystemDef("t2: 3 atoms monoclinic ecut=14Ry 6 bands", a, 14.0,
                      [SPECIES_A], [3], 6, BASE_SEED + 12, 2.0)
    else:
        raise ValueError(index)
    return s.build_positions()


# line lengths with compiled four-step / prime-factor kernels (csrc/fft4_*.hip)
FFT_LENGTHS = (22, 24, 28, 64, 90, 96, 112, 126, 128, 144, 160, 168, 198, 210, 224)


def length_grid_system(k: int, nst: int = 5):
    """Grid k of the fifteen that put every compiled line length on every
    axis: (L[k], L[k+5], L[k+10]) with indices mod 15.  The cell has edges
    0.3 bohr per grid point and the cutoff is as large as the forced grid
    allows (density-basis half-width 0.45 of each axis), so the wavefunction
    sphere fills the share of the active rows a natural grid has (c3: 96 of
    210).  One atom of the first species at the origin.  Returns
    (SystemDef, grid)."""
    L = FFT_LENGTHS
    n = len(L)
    grid = (L[k % n], L[(k + 5) % n], L[(k + 10) % n])
    h, x = 0.3, 0.45
    ecut_ha = (2.0 * math.pi * x / h) ** 2 / 8.0
    s = SystemDef(f"lengths[{k}]: grid {grid[0]}x{grid[1]}x{grid[2]}",
                  np.diag([h * m for m in grid]), 2.0 * ecut_ha, [SPECIES_A], [1], nst,
                  BASE_SEED + 200 + k, 0.0, tau=np.zeros((1, 3)))
    return s, grid


def config4_systems():
    """Twelve uneven Gamma-point systems (first species), for batching."""
    rng = np.random.Generator(np.random.PCG64(BASE_SEED + 3))
    out = []
    for k in range(12):
        n = int(rng.choice([8, 16, 32, 64, 96, 128]))
        L = 16.0 * (n / 8.0) ** (1.0 / 3.0)
        ecut = float(rng.uniform(40.0, 80.0))
        s = SystemDef(f"c4[{k}]: {n} atoms L={L:.3f} ecut={ecut:.2f}Ry", cubic(L), ecut,
                      [SPECIES_A], [n], 2 * n, BASE_SEED + 100 + k, 2.5)
        out.append(s.build_positions())
    return out


def random_coefficients(seed: int, g2: np.ndarray, nst: int) -> np.ndarray:
    """(N(0,1) + i N(0,1)) / (1 + |G|^2) on the half sphere, Im c(G=0) = 0.

    Returns complex128 array of shape (nst, ngw) (band-major, i.e. Qbox's
    column-major c[ig + ngw*n])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ngw = g2.shape[0]
    z = rng.standard_normal((nst, ngw, 2))
    c = (z[..., 0] + 1j * z[..., 1]) / (1.0 + g2)[None, :]
    c[:, 0] = c[:, 0].real
    return np.ascontiguousarray(c)


def random_coefficients_k(seed: int, kpg2: np.ndarray, nst: int) -> np.ndarray:
    """(N(0,1) + i N(0,1)) / (1 + |k+G|^2) on the full sphere of a complex
    (k != 0) basis, shape (nst, ngw)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rng.standard_normal((nst, kpg2.shape[0], 2))
    return np.ascontiguousarray((z[..., 0] + 1j * z[..., 1]) / (1.0 + kpg2)[None, :])


def orthonormal_k(c: np.ndarray) -> np.ndarray:
    """Cholesky orthonormalization of complex-basis bands (rows of c):
    S = c c^H = L L^H, c <- L^-1 c."""
    S = c @ c.conj().T
    L = np.linalg.cholesky(S)
    return np.ascontiguousarray(np.linalg.solve(L, c))


def supercell_x2(s: SystemDef, nst_factor: int = 1) -> SystemDef:
    """The cell doubled along a0, every atom repeated at +a0 (species-major
    order kept: each species' atoms, then their copies)."""
    a = np.array(s.cell, dtype=float)
    a2 = a.copy()
    a2[0] *= 2.0
    tau, na, off = [], [], 0
    t = np.asarray(s.tau, dtype=float)
    for n in s.na:
        blk = t[off:off + n]
        tau += [blk, blk + a[0]]
        na.append(2 * n)
        off += n
    return SystemDef(s.name + " x2", a2, s.ecut_ry, list(s.species), na, s.nst * nst_factor,
                     s.seed, s.min_sep, tau=np.concatenate(tau),
                     occ_bands=None if s.occ_bands is None else np.asarray(s.occ_bands) * 2.0)


def gram_numpy(c: np.ndarray) -> np.ndarray:
    """Gamma-convention Cholesky orthonormalization in FP64 (host, numpy):
    S = 2 P^T P - p0 p0^T with P the real view, C <- C L^-T."""
    c = c.copy()
    c[:, 0] = c[:, 0].real
    P = c.view(np.float64).reshape(c.shape[0], -1)       # (nst, 2 ngw)
    S = 2.0 * (P @ P.T) - np.outer(P[:, 0], P[:, 0])
    L = np.linalg.cholesky(S)
    X = np.linalg.solve(L, P)                             # L X = P  -> rows
    return np.ascontiguousarray(X).view(np.complex128).reshape(c.shape)


def si4gs(si) -> SystemDef:
    """The system of test/si4gs/test.i: four Si atoms in a 16 bohr cube,
    ecut 6 Ry, 8 bands; ``si`` is the species read from Si_VBC_LDA-1.0.xml."""
    tau = np.array([[3.7, 0.0, 0.0], [0.0, 2.2, 0.0], [-3.7, 0.0, 0.0], [0.0, -2.2, 0.0]])
    return SystemDef("si4gs: Si4 L=16 ecut=6Ry 8 bands", cubic(16.0), 6.0, [si], [4], 8,
                     0, 0.0, tau=tau)


def h2ogs(o, h, ecut_ry=70.0) -> SystemDef:
    """The system of test/h2ogs/test.i: one water molecule in a 16 bohr cube,
    4 bands (ecut 70 Ry there); species from O_HSCV_PBE-1.0.xml and
    H_HSCV_PBE-1.0.xml (evaluated here with the LDA functional)."""
    tau = np.array([[0.0, 0.0, 0.0], [1.128, 1.444, 0.0], [1.128, -1.444, 0.0]])
    return SystemDef(f"h2ogs: H2O L=16 ecut={ecut_ry:g}Ry 4 bands", cubic(16.0), ecut_ry, [o, h],
                     [1, 2], 4, 0, 0.0, tau=tau)


def o2gs(o, ecut_ry=70.0) -> SystemDef:
    """The system of test/o2gs/test.i: the oxygen dimer (atoms at +-1.16 bohr
    along x) in an 11 bohr cube, triplet: nspin 2, delta_spin 1, nempty 4,
    12 electrons (ecut 70 Ry there); species from O_HSCV_PBE-1.0.xml
    (evaluated here with the LDA functional, as in the other tests)."""
    tau = np.array([[1.16, 0.0, 0.0], [-1.16, 0.0, 0.0]])
    s = SystemDef(f"o2gs: O2 L=11 ecut={ecut_ry:g}Ry", cubic(11.0), ecut_ry, [o], [2], 6,
                  0, 0.0, tau=tau)
    return spin_polarized(s, 12, 1, 4)


def spin_polarized(s: SystemDef, nel: int, delta_spin: int = 1, nempty: int = 0) -> SystemDef:
    """The nspin = 2 form of a system with nel electrons: nst_up = (nel+1)/2 +
    delta_spin + nempty bands up and nel/2 - delta_spin + nempty down, the
    lowest (nel+1)/2 + delta_spin up and nel/2 - delta_spin down occupied
    with 1 (Wavefunction::set_nspin / update_occ, src/Wavefunction.cpp:326-336,
    424-447)."""
    import dataclasses
    nup_occ = (nel + 1) // 2 + delta_spin
    ndn_occ = nel // 2 - delta_spin
    nst_up = nup_occ + nempty
    nst_dn = ndn_occ + nempty
    occ = np.zeros(nst_up + nst_dn)
    occ[:nup_occ] = 1.0
    occ[nst_up:nst_up + ndn_occ] = 1.0
    return dataclasses.replace(s, name=s.name + f" nspin 2 (up {nst_up}, down {nst_dn})",
                               nst=nst_up + nst_dn, nspin=2, nst_up=nst_up, occ_bands=occ)

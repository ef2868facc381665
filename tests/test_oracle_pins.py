"""External pins of the oracle's energy terms: results the mathematics fixes,
independent of how the oracle computes them.

- Local pseudopotential: the reference transforms v_loc(r) + Z erf(r/rcps)/r
  (src/Species.cpp:648-700).  For v_loc = -Z erf(r/rc)/r the transform is
  -4 pi Z (exp(-G^2 rc^2/4) - exp(-G^2 rcps^2/4)) / G^2, with the value
  pi Z (rc^2 - rcps^2) at G = 0; checked with its derivative at spline
  accuracy.
- Projectors: 4 pi int r^2 j_l(Gr) p_l(r) dr (src/Species.cpp:782-940) of the
  Gaussian projectors p_l = r^l exp(-r^2/(2 s^2)) are
  (2 pi)^{3/2} s^{3+2l} G^l exp(-G^2 s^2/2) (l = 0, 1, 2, 3).
- d and f harmonics: the five real l = 2 and seven l = 3 harmonics of twnl
  satisfy the addition theorem sum_m y_m(u) y_m(u') = ((2l+1) / 4 pi)
  P_l(u.u') (src/NonLocalPotential.cpp:627-760, 963-1245).
- Nonlocal energy: one atom and one Gaussian orbital psi (s and p type) in a
  large cell; enl = occ D_l <beta_lm|psi>^2 with beta_lm = p_l(r) Y_lm and the
  overlap integral in closed form (src/NonLocalPotential.cpp:1469-1571).
  This fixes D_l, the projector weights and the 1/omega of the reference.
- Ion-ion energy: esr - eself + ehart_p - pi Q^2 / (2 omega eta^2) (eta =
  1/(sqrt 2 rcps)) is the Ewald energy of the point ions in a neutralizing
  background, computed here by an independent Ewald sum with another
  splitting parameter (src/EnergyFunctional.cpp:1097-1192, :327-351).
- Smearing width: for a neutral system etotal does not depend on rcps
  (it moves energy between eps, ehart, esr and eself, src/AtomSet.cpp:43-44).
- Core density (non-linear core correction): the transform of a Gaussian
  core density A exp(-(r/rc)^2) is A pi^{3/2} rc^3 exp(-G^2 rc^2/4)
  (src/Species.cpp:1056-1100, 698-717), and the core density on the grid is
  the periodic sum of the atoms' Gaussians, half of it per spin with two
  spins (src/EnergyFunctional.cpp:1064-1083).
"""
import dataclasses
import math

import numpy as np
import pytest
from scipy.special import erfc

from qbg import configs as cf


def _species_oracle(oracle_mod, sp, L=16.0, ecut_ry=30.0):
    s = cf.SystemDef("one atom", cf.cubic(L), ecut_ry, [sp], [1], 1, 1, 0.0,
                     tau=np.zeros((1, 3)))
    return oracle_mod.Oracle(s)


def _closed_forms(sp, g):
    """(vloc, dvloc, p0, dp0, p1, dp1) of the closed forms at g > 0."""
    Z, rc, rp, s = sp.zval, sp.rc, sp.rcps, sp.sigma
    gg = g * g
    a, b = np.exp(-gg * rc ** 2 / 4), np.exp(-gg * rp ** 2 / 4)
    v = -4 * math.pi * Z * (a - b) / gg
    dv = -4 * math.pi * Z * ((-rc ** 2 * g / 2 * a + rp ** 2 * g / 2 * b) / gg - 2 * (a - b) / (gg * g))
    e = np.exp(-gg * s * s / 2)
    p0 = (2 * math.pi) ** 1.5 * s ** 3 * e
    p1 = (2 * math.pi) ** 1.5 * s ** 5 * g * e
    return v, dv, p0, -(s * s) * g * p0, p1, (2 * math.pi) ** 1.5 * s ** 5 * e * (1 - gg * s * s)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("sp", [cf.SPECIES_A, cf.SPECIES_B])
def test_species_transforms_closed_form(oracle_mod, sp):
    """At the spline nodes g_i = i pi / (ndft dr) the tables are the
    transforms themselves (1e-10); between nodes the cubic spline adds its
    interpolation error (a few 1e-7 at dg = 0.077 / bohr)."""
    o = _species_oracle(oracle_mod, sp)
    dg = math.pi / (o.ndft * sp.deltar)
    nodes = np.arange(1, int(12.0 / dg)) * dg
    v, dv, vn, dvn = o.species_eval(0, nodes)
    cv, cdv, p0, dp0, p1, dp1 = _closed_forms(sp, nodes)
    assert _rel(v, cv) < 1e-10 and _rel(vn[0], p0) < 1e-10 and _rel(vn[1], p1) < 1e-10
    v0 = o.species_eval(0, np.zeros(1))[0][0]
    assert abs(v0 - math.pi * sp.zval * (sp.rc ** 2 - sp.rcps ** 2)) < 1e-10 * abs(v0)
    mid = nodes + 0.5 * dg
    v, dv, vn, dvn = o.species_eval(0, mid)
    cv, cdv, p0, dp0, p1, dp1 = _closed_forms(sp, mid)
    assert _rel(v, cv) < 1e-6 and _rel(vn[0], p0) < 1e-6 and _rel(vn[1], p1) < 1e-6
    assert _rel(dv, cdv) < 1e-5 and _rel(dvn[0], dp0) < 1e-5 and _rel(dvn[1], dp1) < 1e-5


def test_d_projector_transform_closed_form(oracle_mod):
    sp = cf.SPECIES_D
    o = _species_oracle(oracle_mod, sp)
    dg = math.pi / (o.ndft * sp.deltar)
    s = sp.sigma
    for g, tol in ((np.arange(1, int(12.0 / dg)) * dg, 1e-10),
                   (np.arange(1, int(12.0 / dg)) * dg + 0.5 * dg, 1e-6)):
        vn = o.species_eval(0, g, nop=3)[2]
        e = np.exp(-g * g * s * s / 2)
        assert _rel(vn[2], (2 * math.pi) ** 1.5 * s ** 7 * g * g * e) < tol
        assert _rel(vn[0], (2 * math.pi) ** 1.5 * s ** 3 * e) < tol
    assert o.species_eval(0, np.zeros(1), nop=3)[2][2][0] == 0.0


def test_f_projector_transform_closed_form(oracle_mod):
    sp = cf.SPECIES_F
    o = _species_oracle(oracle_mod, sp)
    dg = math.pi / (o.ndft * sp.deltar)
    s = sp.sigma
    for g, tol in ((np.arange(1, int(12.0 / dg)) * dg, 1e-10),
                   (np.arange(1, int(12.0 / dg)) * dg + 0.5 * dg, 1e-6)):
        vn = o.species_eval(0, g, nop=4)[2]
        e = np.exp(-g * g * s * s / 2)
        assert _rel(vn[3], (2 * math.pi) ** 1.5 * s ** 9 * g ** 3 * e) < tol
        assert _rel(vn[2], (2 * math.pi) ** 1.5 * s ** 7 * g * g * e) < tol
    assert o.species_eval(0, np.zeros(1), nop=4)[2][3][0] == 0.0


def test_core_density_transform_closed_form(oracle_mod):
    sp = cf.SPECIES_C
    o = _species_oracle(oracle_mod, sp)
    dg = math.pi / (o.ndft * sp.deltar)
    A, rc = sp.core
    zref = A * math.pi ** 1.5 * rc ** 3
    nodes = np.arange(1, int(12.0 / dg)) * dg
    for g, tol in ((nodes, 1e-10), (nodes + 0.5 * dg, 1e-6)):
        rho, drho, z = o.species_core(0, g)
        ref = zref * np.exp(-g * g * rc * rc / 4)
        assert _rel(rho, ref) < tol
        if tol > 1e-10:
            assert _rel(drho, -0.5 * g * rc * rc * ref) < 1e-5
    rho0, drho0, z = o.species_core(0, np.zeros(1))
    assert abs(z - zref) < 1e-9 * zref and rho0[0] == z and abs(drho0[0]) < 1e-12
    # a species without a core density has none
    assert np.all(_species_oracle(oracle_mod, cf.SPECIES_A).species_core(0, nodes)[0] == 0.0)


@pytest.mark.parametrize("spin", [False, True])
def test_core_density_real_space(oracle_mod, spin):
    s = dataclasses.replace(cf.config(15), ecut_ry=40.0)
    if spin:
        s = cf.spin_polarized(s, 12, 0, 0)
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(7, o.basis(0)["g2"], s.nst))
    if spin:
        u = s.nst_up
        c = np.concatenate([cf.gram_numpy(c[:u]), cf.gram_numpy(c[u:])])
    o.evaluate(c, hpsi=False, forces=False)
    got = o.get(17)
    n0, n1, n2 = o.np0, o.np1, o.np2
    L = np.diag(s.cell)
    x = (np.arange(n0) / n0)[:, None, None] * L[0]
    y = (np.arange(n1) / n1)[None, :, None] * L[1]
    z = (np.arange(n2) / n2)[None, None, :] * L[2]
    A, rc = cf.SPECIES_C.core
    ref = np.zeros((n0, n1, n2))
    for t in s.tau[:2]:                       # the two atoms of the core species
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                for k in (-1, 0, 1):
                    d2 = ((x - t[0] - i * L[0]) ** 2 + (y - t[1] - j * L[1]) ** 2 +
                          (z - t[2] - k * L[2]) ** 2)
                    ref += A * np.exp(-d2 / rc ** 2)
    ref = ref.transpose(2, 1, 0).ravel()      # i + np0 (j + np1 k)
    if spin:
        ref = 0.5 * ref
    err = np.abs(got - ref).max() / ref.max()
    print(f"  core density on the grid: max rel err {err:.2e}")
    assert err < 1e-7


def test_f_harmonics_addition_theorem(oracle_mod):
    s = dataclasses.replace(cf.config(14), na=[1, 0], species=[cf.SPECIES_F, cf.SPECIES_B],
                            tau=np.zeros((1, 3)))
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    ngw = o.ngw
    tw = o.get(9, 0).reshape(16, ngw)
    g = np.sqrt(b["g2"])
    v = o.species_eval(0, g, nop=4)[2][3]
    rng = np.random.default_rng(6)
    i = rng.integers(1, ngw, 400)
    j = rng.integers(1, ngw, 400)
    u = (b["gx"] / np.where(g > 0, g, 1.0)).T
    x = np.einsum("ij,ij->i", u[i], u[j])
    lhs = np.einsum("mi,mi->i", tw[9:16][:, i], tw[9:16][:, j])
    rhs = v[i] * v[j] * 7 / (4 * math.pi) * 0.5 * (5 * x ** 3 - 3 * x)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(v).max() ** 2


def test_d_harmonics_addition_theorem(oracle_mod):
    s = dataclasses.replace(cf.config(13), na=[1, 0], species=[cf.SPECIES_D, cf.SPECIES_B],
                            tau=np.zeros((1, 3)))
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    ngw = o.ngw
    tw = o.get(9, 0).reshape(9, ngw)
    g = np.sqrt(b["g2"])
    v = o.species_eval(0, g, nop=3)[2][2]
    rng = np.random.default_rng(5)
    i = rng.integers(1, ngw, 400)
    j = rng.integers(1, ngw, 400)
    u = (b["gx"] / np.where(g > 0, g, 1.0)).T
    cos = np.einsum("ij,ij->i", u[i], u[j])
    lhs = np.einsum("mi,mi->i", tw[4:9][:, i], tw[4:9][:, j])
    rhs = v[i] * v[j] * 5 / (4 * math.pi) * 0.5 * (3 * cos * cos - 1)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(v).max() ** 2
    # and the p block: sum_m y_m y_m' = (3 / 4 pi) u.u'
    vp = o.species_eval(0, g, nop=3)[2][1]
    lhs = np.einsum("mi,mi->i", tw[1:4][:, i], tw[1:4][:, j])
    assert np.abs(lhs - vp[i] * vp[j] * 3 / (4 * math.pi) * cos).max() < 1e-12 * np.abs(vp).max() ** 2


def _gaussian_orbital(b, omega, a, kind):
    """Plane-wave coefficients (half sphere) of a normalized Gaussian orbital
    of width a at the origin: s = exp(-r^2/2a^2), p = x exp(-r^2/2a^2);
    c(G) = omega^{-1/2} int psi e^{-iGr} d^3r."""
    g2 = b["g2"]
    gx = b["gx"][0]
    fg = (2 * math.pi * a * a) ** 1.5 * np.exp(-g2 * a * a / 2)
    if kind == "s":
        n = (math.pi * a * a) ** -0.75
        c = n * fg + 0j
    elif kind == "p":
        n = ((a * a / 2) * (math.pi * a * a) ** 1.5) ** -0.5
        c = n * (-1j * a * a * gx) * fg
    elif kind == "d":                       # d: x y exp(-r^2/2a^2)
        gy = b["gx"][1]
        n = ((a * a / 2) ** 2 * (math.pi * a * a) ** 1.5) ** -0.5
        c = n * (-(a ** 4) * gx * gy) * fg + 0j
    else:                                   # f: x y z exp(-r^2/2a^2)
        gy, gz = b["gx"][1], b["gx"][2]
        n = ((a * a / 2) ** 3 * (math.pi * a * a) ** 1.5) ** -0.5
        c = n * (1j * a ** 6 * gx * gy * gz) * fg
    return (c / math.sqrt(omega))[None, :]


@pytest.mark.parametrize("sp,kind", [(cf.SPECIES_A, "s"), (cf.SPECIES_A, "p"),
                                     (cf.SPECIES_B, "s"), (cf.SPECIES_B, "p"),
                                     (cf.SPECIES_D, "s"), (cf.SPECIES_D, "d"),
                                     (cf.SPECIES_F, "f")])
def test_nonlocal_energy_single_atom(oracle_mod, sp, kind):
    L, a = 22.0, 1.1
    s = cf.SystemDef("one atom, one orbital", cf.cubic(L), 50.0, [sp], [1], 1, 1, 0.0,
                     tau=np.zeros((1, 3)), occ_bands=np.array([2.0]))
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = _gaussian_orbital(b, L ** 3, a, kind)
    # the orbital is normalized on the half sphere (Gamma convention) up to
    # the basis truncation
    norm = 2 * np.sum(np.abs(c[0]) ** 2) - abs(c[0, 0]) ** 2
    assert abs(norm - 1) < 1e-12
    enl = o.nl_energy(c, None, forces=False, stress=False)[0]
    sg = sp.sigma
    s2 = 1.0 / (1.0 / sg ** 2 + 1.0 / a ** 2)
    if kind == "s":
        ov = (4 * math.pi) ** -0.5 * (math.pi * a * a) ** -0.75 * (2 * math.pi * s2) ** 1.5
        ref = 2.0 * sp.d[0] * ov ** 2
    elif kind == "p":
        n = ((a * a / 2) * (math.pi * a * a) ** 1.5) ** -0.5
        ov = math.sqrt(3 / (4 * math.pi)) * n * s2 * (2 * math.pi * s2) ** 1.5
        ref = 2.0 * sp.d[1] * ov ** 2
    elif kind == "d":   # beta = r^2 exp(-r^2/2s^2) sqrt(15/4pi) x y / r^2
        n = ((a * a / 2) ** 2 * (math.pi * a * a) ** 1.5) ** -0.5
        ov = math.sqrt(15 / (4 * math.pi)) * n * s2 ** 2 * (2 * math.pi * s2) ** 1.5
        ref = 2.0 * sp.d[2] * ov ** 2
    else:   # beta = r^3 exp(-r^2/2s^2) sqrt(105/4pi) x y z / r^3
        n = ((a * a / 2) ** 3 * (math.pi * a * a) ** 1.5) ** -0.5
        ov = math.sqrt(105 / (4 * math.pi)) * n * s2 ** 3 * (2 * math.pi * s2) ** 1.5
        ref = 2.0 * sp.d[3] * ov ** 2
    assert abs(enl - ref) < 1e-6 * abs(ref), (enl, ref)   # spline accuracy of twnl


@pytest.mark.parametrize("kind,llocal", [("s", 2), ("p", 2), ("p", 0), ("d", 0)])
def test_semilocal_quadrature_single_atom(oracle_mod, kind, llocal):
    """Semilocal quadrature form (nquad > 0): projectors 4 pi j_l(G r_q) r_q
    Y_lm at the trapezoid points r_q = (q+1) rquad/(nquad+1) with weights
    (v_l(r_q) - v_loc(r_q)) h (src/NonLocalPotential.cpp:104-229, 317-960).
    For one atom and a Gaussian orbital psi = R(r) Y_lm the projection onto
    the shell r_q is r_q R(r_q), so enl = occ sum_q h dv_l(r_q) r_q^2 R(r_q)^2,
    evaluated here with scipy's clamped cubic spline of the semilocal
    potentials (v'(0) = 0, v'(r_max) = Z/r_max^2, src/Species.cpp:162-184)."""
    import dataclasses
    from pathlib import Path
    from scipy.interpolate import CubicSpline
    from qbg.species import read_species
    si = read_species(Path(__file__).resolve().parent / "golden" / "Si_VBC_LDA-1.0.xml")
    nquad, rquad = 24, 2.6
    sp = dataclasses.replace(si, nquad=nquad, rquad=rquad, llocal=llocal)
    L, a = 22.0, 1.1
    s = cf.SystemDef("one atom, one orbital", cf.cubic(L), 50.0, [sp], [1], 1, 1, 0.0,
                     tau=np.zeros((1, 3)), occ_bands=np.array([2.0]))
    o = oracle_mod.Oracle(s)
    c = _gaussian_orbital(o.basis(0), L ** 3, a, kind)
    enl = o.nl_energy(c, None, forces=False, stress=False)[0]
    l = {"s": 0, "p": 1, "d": 2}[kind]
    ndft = 1
    while ndft * sp.deltar < 40.0:
        ndft *= 2
    r = np.arange(ndft) * sp.deltar

    def vspl(ll):
        v = np.empty(ndft)
        v[:sp.nmesh] = sp.vps[ll]
        v[sp.nmesh:] = -sp.zval / r[sp.nmesh:]
        return CubicSpline(r, v, bc_type=((1, 0.0), (1, sp.zval / r[-1] ** 2)))

    h = rquad / (nquad + 1)
    rq = h * np.arange(1, nquad + 1)
    dv = vspl(l)(rq) - vspl(llocal)(rq)
    g = np.exp(-rq ** 2 / (2 * a * a))
    if kind == "s":
        R = math.sqrt(4 * math.pi) * (math.pi * a * a) ** -0.75 * g
    elif kind == "p":
        R = (math.sqrt(4 * math.pi / 3) * ((a * a / 2) * (math.pi * a * a) ** 1.5) ** -0.5
             * rq * g)
    else:
        R = (math.sqrt(4 * math.pi / 15) * ((a * a / 2) ** 2 * (math.pi * a * a) ** 1.5) ** -0.5
             * rq * rq * g)
    ref = 2.0 * np.sum(h * dv * rq * rq * R * R)
    print(f"  {kind} (llocal {llocal}): enl {enl:.14e} quadrature sum {ref:.14e}")
    assert abs(enl - ref) < 1e-9 * abs(ref)


def _ewald(cell, tau, z, eta):
    """Ewald energy of point charges z at tau in a neutralizing background."""
    A = np.asarray(cell, float)
    omega = abs(np.linalg.det(A))
    B = 2 * math.pi * np.linalg.inv(A).T
    n = len(z)
    Q = z.sum()
    # real space
    rmax = 40.0 / eta ** 0 if False else 7.0 / eta
    nr = [int(rmax / np.linalg.norm(A[i])) + 2 for i in range(3)]
    er = 0.0
    for i in range(n):
        for j in range(n):
            d = tau[i] - tau[j]
            for i0 in range(-nr[0], nr[0] + 1):
                for i1 in range(-nr[1], nr[1] + 1):
                    rr = d + i0 * A[0] + i1 * A[1] + np.outer(np.arange(-nr[2], nr[2] + 1), A[2])
                    r = np.linalg.norm(rr, axis=1)
                    if i == j and i0 == 0 and i1 == 0:
                        r = r[r > 1e-12]
                    er += 0.5 * z[i] * z[j] * np.sum(erfc(eta * r) / r)
    # reciprocal space
    gmax = 2 * eta * 6.5
    ng = [int(gmax / np.linalg.norm(B[i])) + 2 for i in range(3)]
    h = np.arange(-ng[0], ng[0] + 1)
    k = np.arange(-ng[1], ng[1] + 1)
    l_ = np.arange(-ng[2], ng[2] + 1)
    H, K, Lg = np.meshgrid(h, k, l_, indexing="ij")
    G = (H[..., None] * B[0] + K[..., None] * B[1] + Lg[..., None] * B[2]).reshape(-1, 3)
    G2 = (G * G).sum(1)
    G, G2 = G[G2 > 1e-12], G2[G2 > 1e-12]
    S = (z[None, :] * np.exp(-1j * G @ tau.T)).sum(1)
    erec = 2 * math.pi / omega * np.sum(np.abs(S) ** 2 * np.exp(-G2 / (4 * eta * eta)) / G2)
    eself = eta / math.sqrt(math.pi) * np.sum(z * z)
    ebg = math.pi * Q * Q / (2 * omega * eta * eta)
    return er + erec - eself - ebg


@pytest.mark.parametrize("idx", [0, 11, 12])
def test_ion_energy_against_ewald(oracle_mod, idx):
    s = cf.config(idx)
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(9, o.basis(0)["g2"], s.nst))
    r = o.evaluate(c, hpsi=False, forces=False)
    z = np.concatenate([[sp.zval] * n for sp, n in zip(s.species, s.na)]).astype(float)
    omega = abs(np.linalg.det(s.cell))
    eta = 1.0 / (math.sqrt(2.0) * cf.RCPS)
    qbox = r["esr"] - r["eself"] + r["ehart_p"] - math.pi * z.sum() ** 2 / (2 * omega * eta ** 2)
    for eta2 in (0.45, 0.7):
        ref = _ewald(s.cell, s.tau, z, eta2)
        assert abs(qbox - ref) < 1e-12 * abs(ref), (idx, eta2, qbox, ref)


def _with_rcps(s, rcps):
    return dataclasses.replace(s, species=[dataclasses.replace(sp, rcps=rcps)
                                           for sp in s.species])


@pytest.mark.parametrize("idx,nst", [(10, 4), (11, 7)])
def test_etotal_independent_of_rcps(oracle_mod, idx, nst):
    s = dataclasses.replace(cf.config(idx), nst=nst, ecut_ry=40.0)
    assert 2 * nst == sum(sp.zval * n for sp, n in zip(s.species, s.na))     # neutral
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(21, o.basis(0)["g2"], nst))
    out = {}
    for rcps in (1.0, 1.2, 1.5):
        r = oracle_mod.Oracle(_with_rcps(s, rcps)).evaluate(c, hpsi=False, forces=False)
        out[rcps] = r
    et = [out[k]["etotal"] for k in out]
    # the parts do move with rcps; their sum does not
    assert abs(out[1.0]["esr"] - out[1.5]["esr"]) > 1e-3
    assert max(et) - min(et) < 1e-7, et

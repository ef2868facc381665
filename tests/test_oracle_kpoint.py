"""The oracle's complex (k != 0) basis: one k-point of weight 1, the
wavefunction basis the full sphere |k+G| < sqrt(2 ecut) (src/Basis.cpp:237,
446-481), one band per transform (src/SlaterDet.cpp:225-298, 410-477),
kinetic energy with 1/2 |k+G|^2 (src/EnergyFunctional.cpp:441-571, 789-829),
complex projections anl^H c, H psi, forces and stress with k + G
(src/NonLocalPotential.cpp:1429-1445, 1572-1630, 1722-1775, 1990-2066).

What pins it, beside the reference's formulas:
- the complex branch at k = 0 on coefficients of real functions (c(-G) =
  conj c(G)) gives the real basis's energies, forces and H psi to rounding:
  a dropped 1/2 of the kinetic energy, a wrong projection factor or a missing
  -G partner fails it;
- Bloch's theorem: a band at k = b0/2 of a cell is a k = 0 band of the cell
  doubled along a0 (G' = k + G on the doubled cell's lattice), so the
  doubled cell with twice the occupation has twice every energy term and the
  same force on each atom and its copy, and the same H psi; a k + G taken as
  G anywhere (kinetic, projectors, their phases) fails it;
- time reversal: k -> -k with c(G) -> conj c(-G) leaves every energy and
  force unchanged and maps H psi the same way;
- H psi and the forces are energy gradients (central differences), and the
  stress is the strain derivative of the energy at fixed k in units of b
  (Richardson-extrapolated central differences, each term).
"""
import dataclasses

import numpy as np
import pytest

from qbg import configs as cf

TERMS = ["ekin", "eps", "enl", "ehart", "exc", "esr", "eself", "etotal"]


def _index(idx):
    return {tuple(int(v) for v in r): i for i, r in enumerate(idx)}


def _kcoeffs(o, seed, nst):
    kpg2, _, real = o.kpg()
    assert not real
    return cf.orthonormal_k(cf.random_coefficients_k(seed, kpg2, nst))


def test_complex_branch_at_gamma_equals_real_basis(oracle_mod):
    s = cf.config(13)
    o_r = oracle_mod.Oracle(s)
    o_c = oracle_mod.Oracle(s, force_complex=True)
    assert o_c.ngw == 2 * o_r.ngw - 1
    b_r, b_c = o_r.basis(0), o_c.basis(0)
    pos = _index(b_c["idx"])
    c_r = cf.gram_numpy(cf.random_coefficients(1601, b_r["g2"], s.nst))
    c_c = np.zeros((s.nst, o_c.ngw), dtype=complex)
    plus = np.array([pos[tuple(int(v) for v in g)] for g in b_r["idx"]])
    minus = np.array([pos[tuple(-int(v) for v in g)] for g in b_r["idx"]])
    c_c[:, minus] = np.conj(c_r)
    c_c[:, plus] = c_r
    r_r = o_r.evaluate(c_r, hpsi=True, forces=True, stress=True)
    r_c = o_c.evaluate(c_c, hpsi=True, forces=True, stress=True)
    assert np.abs(r_c["sigma"] - r_r["sigma"]).max() < 1e-11 * np.abs(r_r["sigma"]).max()
    for t in TERMS:
        print(f"  {t} real {r_r[t]:.12e} complex {r_c[t]:.12e}")
        assert abs(r_c[t] - r_r[t]) <= 1e-11 * max(1.0, abs(r_r[t])), t
    assert np.abs(r_c["fion"] - r_r["fion"]).max() < 1e-10
    h_c = r_c["hpsi"]
    assert np.abs(h_c[:, plus] - r_r["hpsi"]).max() < 1e-11 * np.abs(r_r["hpsi"]).max()
    assert np.abs(h_c[:, minus] - np.conj(r_r["hpsi"])).max() < 1e-11 * np.abs(r_r["hpsi"]).max()


@pytest.mark.parametrize("idx", [11, 13])
def test_half_zone_band_is_doubled_cell_gamma_band(oracle_mod, idx):
    base = cf.config(idx)
    nst = 3
    p = dataclasses.replace(base, nst=nst, occ_bands=np.ones(nst), kpoint=np.array([0.5, 0.0, 0.0]))
    sc = cf.supercell_x2(dataclasses.replace(p, kpoint=None))
    o_p = oracle_mod.Oracle(p)
    grid_s = (2 * o_p.np0, o_p.np1, o_p.np2)
    o_s = oracle_mod.Oracle(sc, grid=grid_s, force_complex=True)
    assert (o_s.np0, o_s.np1, o_s.np2) == grid_s
    c_p = _kcoeffs(o_p, 1700 + idx, nst)
    pos = _index(o_s.basis(0)["idx"])
    # k + G = (h + 1/2) b0 + k b1 + l b2 = (2h + 1) b0' + k b1 + l b2
    img = np.array([pos[(2 * int(h) + 1, int(k), int(l))] for h, k, l in o_p.basis(0)["idx"]])
    kpg2_p = o_p.kpg()[0]
    kpg2_s = o_s.kpg()[0]
    assert np.allclose(kpg2_s[img], kpg2_p, rtol=1e-13, atol=1e-13)
    # every doubled-cell vector at odd h' is an image (same cutoff sphere)
    odd = np.array([i for i, g in enumerate(o_s.basis(0)["idx"]) if g[0] % 2 != 0])
    assert sorted(odd) == sorted(img)
    c_s = np.zeros((nst, o_s.ngw), dtype=complex)
    c_s[:, img] = c_p
    r_p = o_p.evaluate(c_p, hpsi=True, forces=True, stress=True)
    r_s = o_s.evaluate(c_s, hpsi=True, forces=True, stress=True)
    # the stress is intensive: every part equal
    assert np.abs(r_s["sigma"] - r_p["sigma"]).max() < 1e-10 * np.abs(r_p["sigma"]).max()
    for t in TERMS:
        print(f"  {t} 2 x cell {2 * r_p[t]:.12e} doubled cell {r_s[t]:.12e}")
        assert abs(r_s[t] - 2.0 * r_p[t]) <= 1e-10 * max(1.0, abs(r_p[t])), t
    # the force on each atom and on its copy at +a0
    fp = r_p["fion"]
    fs = r_s["fion"]
    off_p = off_s = 0
    for n in p.na:
        for copy in range(2):
            d = np.abs(fs[off_s + copy * n: off_s + (copy + 1) * n] - fp[off_p:off_p + n]).max()
            assert d < 1e-9, (copy, d)
        off_p += n
        off_s += 2 * n
    hs = r_s["hpsi"]
    assert np.abs(hs[:, img] - r_p["hpsi"]).max() < 1e-10 * np.abs(r_p["hpsi"]).max()
    mask = np.ones(o_s.ngw, dtype=bool)
    mask[img] = False
    assert np.abs(hs[:, mask]).max() < 1e-12


def test_time_reversal(oracle_mod):
    s = cf.config(16)
    o_k = oracle_mod.Oracle(s)
    o_m = oracle_mod.Oracle(s, kpoint=-np.asarray(s.kpoint))
    c = _kcoeffs(o_k, 1801, s.nst)
    pos = _index(o_m.basis(0)["idx"])
    img = np.array([pos[tuple(-int(v) for v in g)] for g in o_k.basis(0)["idx"]])
    assert o_m.ngw == o_k.ngw
    c_m = np.zeros_like(c)
    c_m[:, img] = np.conj(c)
    r_k = o_k.evaluate(c, hpsi=True, forces=True)
    r_m = o_m.evaluate(c_m, hpsi=True, forces=True)
    for t in TERMS:
        assert abs(r_m[t] - r_k[t]) <= 1e-11 * max(1.0, abs(r_k[t])), t
    assert np.abs(r_m["fion"] - r_k["fion"]).max() < 1e-10
    assert np.abs(r_m["hpsi"][:, img] - np.conj(r_k["hpsi"])).max() < 1e-11 * np.abs(r_k["hpsi"]).max()


def test_hpsi_is_energy_gradient_at_k(oracle_mod):
    s = cf.config(16)
    o = oracle_mod.Oracle(s)
    c = _kcoeffs(o, 1901, s.nst)
    r = o.evaluate(c, hpsi=True, forces=False)
    assert abs(r["total_charge"] - s.occ.sum()) < 1e-9
    rng = np.random.default_rng(9)
    d = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    d *= 1e-4 / np.abs(d).max()
    # dE = sum_n occ_n 2 Re <d_n | H psi_n> on the full sphere
    pred = sum(s.occ[n] * 2.0 * np.real(np.vdot(d[n], r["hpsi"][n])) for n in range(s.nst))
    ep = o.evaluate(c + d, hpsi=False, forces=False)["etotal"]
    em = o.evaluate(c - d, hpsi=False, forces=False)["etotal"]
    fd = (ep - em) / 2.0
    print(f"  dE fd {fd:.10e} pred {pred:.10e}")
    assert abs(fd - pred) <= 1e-6 * abs(pred)


def test_forces_are_energy_gradient_at_k(oracle_mod):
    s = cf.config(16)
    o = oracle_mod.Oracle(s)
    c = _kcoeffs(o, 2001, s.nst)
    r = o.evaluate(c, hpsi=False, forces=True)
    h = 1e-4
    for a, j in ((0, 0), (1, 2), (s.nat - 1, 1)):
        e = []
        for sgn in (1.0, -1.0):
            t = np.array(s.tau, dtype=float)
            t[a, j] += sgn * h
            o2 = oracle_mod.Oracle(dataclasses.replace(s, tau=t))
            e.append(o2.evaluate(c, hpsi=False, forces=False)["etotal"])
        fd = -(e[0] - e[1]) / (2 * h)
        print(f"  f[{a},{j}] fd {fd:.10e} oracle {r['fion'][a, j]:.10e}")
        assert abs(fd - r["fion"][a, j]) <= 1e-6 * max(1.0, abs(fd))


PAIRS = {"ekin": 1, "eps": 2, "enl": 3, "ehart": 4, "exc": 5, "esr": 6, "etotal": 0}
VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]


@pytest.mark.parametrize("idx", [16, 12])
def test_stress_vs_strain_fd_at_k(oracle_mod, idx):
    s = cf.config(idx)
    if s.kpoint is None:
        s = dataclasses.replace(s, kpoint=np.array([-0.3, 0.2, 0.45]))
    ref = s.cell * 1.04                   # one basis for every strained cell
    o0 = oracle_mod.Oracle(s, refcell=ref)
    c = _kcoeffs(o0, 2100 + idx, s.nst)
    sigma = o0.evaluate(c, hpsi=False, forces=False, stress=True)["sigma"]
    assert np.allclose(sigma[0], sigma[1:].sum(axis=0), rtol=0, atol=1e-14)
    omega = abs(np.linalg.det(s.cell))
    idx0 = o0.basis(0)["idx"]

    def diff(u):
        F = np.eye(3) + u
        out = []
        for sg in (1.0, -1.0):
            Fs = np.eye(3) + sg * u
            o = oracle_mod.Oracle(dataclasses.replace(s, cell=s.cell @ Fs.T, tau=s.tau @ Fs.T),
                                  refcell=ref)
            assert np.array_equal(o.basis(0)["idx"], idx0)
            out.append(o.evaluate(c, hpsi=False, forces=False))
        return {t: out[0][t] - out[1][t] for t in PAIRS}

    rng = np.random.default_rng(idx)
    v = 1e-4 * rng.uniform(-1.0, 1.0, 6)
    u = np.zeros((3, 3))
    for k, (i, j) in enumerate(VOIGT):
        u[i, j] = u[j, i] = v[k]
    d1, d2 = diff(u), diff(0.5 * u)
    for term, row in PAIRS.items():
        de = (8.0 * d2[term] - d1[term]) / 3.0
        st = sigma[row]
        pred = -2.0 * omega * (sum(v[k] * st[k] for k in range(3)) +
                               2.0 * sum(v[k] * st[k] for k in range(3, 6)))
        print(f"{term:7s} fd {de: .10e} sigma {pred: .10e} diff {de - pred: .2e}")
        assert abs(de - pred) <= 2e-7 * abs(de) + 1e-12, (term, de, pred)


def test_gram_complex(oracle_mod):
    """SlaterDet::gram of the complex basis: the result is orthonormal, equals
    numpy's Cholesky orthonormalization (S = c c^H = L L^H, c <- L^-1 c for
    band rows), and at k = 0 on coefficients of real functions equals the
    real basis's gram."""
    s = cf.config(16)
    o = oracle_mod.Oracle(s)
    kpg2 = o.kpg()[0]
    c = cf.random_coefficients_k(2400, kpg2, s.nst)
    g = oracle_mod.gram_complex(c)
    assert np.abs(g @ g.conj().T - np.eye(s.nst)).max() < 1e-13
    assert np.abs(g - cf.orthonormal_k(c)).max() < 1e-12 * np.abs(g).max()
    # the Gamma limit
    s0 = cf.config(13)
    o_r = oracle_mod.Oracle(s0)
    o_c = oracle_mod.Oracle(s0, force_complex=True)
    b_r = o_r.basis(0)
    pos = _index(o_c.basis(0)["idx"])
    plus = np.array([pos[tuple(int(v) for v in x)] for x in b_r["idx"]])
    minus = np.array([pos[tuple(-int(v) for v in x)] for x in b_r["idx"]])
    c_r = cf.random_coefficients(2401, b_r["g2"], s0.nst)
    c_r[:, 0] = c_r[:, 0].real
    c_c = np.zeros((s0.nst, o_c.ngw), dtype=complex)
    c_c[:, minus] = np.conj(c_r)
    c_c[:, plus] = c_r
    g_r = oracle_mod.gram(c_r)
    g_c = oracle_mod.gram_complex(c_c)
    assert np.abs(g_c[:, plus] - g_r).max() < 1e-12
    assert np.abs(g_c[:, minus] - np.conj(g_r)).max() < 1e-12


# ------------------------------------------------------------ k-point sets
def _kset_coeffs(o, nkp, nst, seed):
    out = []
    for ik in range(nkp):
        kpg2 = o.kset_basis(ik)["kpg2"]
        out.append(cf.orthonormal_k(cf.random_coefficients_k(seed + ik, kpg2, nst)))
    return out


def test_kset_repeated_point_is_single_point(oracle_mod):
    """A set holding one k-point twice with weights 0.3 and 0.7 and the same
    bands at both is that k-point alone (the weights sum, every part)."""
    s = cf.config(16)
    k = np.asarray(s.kpoint)
    o1 = oracle_mod.Oracle(s)
    o2 = oracle_mod.Oracle(s, kpoints=[k, k], weights=[0.3, 0.7])
    c = _kcoeffs(o1, 2700, s.nst)
    r1 = o1.evaluate(c, hpsi=True, forces=True, stress=True)
    r2 = o2.evaluate([c, c], hpsi=True, forces=True, stress=True)
    for t in TERMS:
        assert abs(r2[t] - r1[t]) <= 1e-11 * max(1.0, abs(r1[t])), t
    assert np.abs(r2["fion"] - r1["fion"]).max() < 1e-11
    assert np.abs(r2["sigma"] - r1["sigma"]).max() < 1e-11 * np.abs(r1["sigma"]).max()
    for h in r2["hpsi"]:
        assert np.abs(h - r1["hpsi"]).max() < 1e-12 * np.abs(r1["hpsi"]).max()


@pytest.mark.parametrize("idx", [11, 13])
def test_kset_is_doubled_cell(oracle_mod, idx):
    """The set {0, b0/2} with weights 1/2 samples the zone of the cell doubled
    along a0: the doubled cell's Gamma bands (the images of both k-points'
    bands) have twice every energy term, the same force on each atom and its
    copy, the same stress, and the images of both k-points' H psi."""
    base = cf.config(idx)
    nst = 3
    p = dataclasses.replace(base, nst=nst, occ_bands=np.ones(nst), kpoint=None)
    kps = [np.zeros(3), np.array([0.5, 0.0, 0.0])]
    o_p = oracle_mod.Oracle(p, kpoints=kps, weights=[0.5, 0.5])
    sc = dataclasses.replace(cf.supercell_x2(p), nst=2 * nst, occ_bands=np.ones(2 * nst))
    grid_s = (2 * o_p.np0, o_p.np1, o_p.np2)
    o_s = oracle_mod.Oracle(sc, grid=grid_s, force_complex=True)
    cp = _kset_coeffs(o_p, 2, nst, 2800 + idx)
    pos = _index(o_s.basis(0)["idx"])
    c_s = np.zeros((2 * nst, o_s.ngw), dtype=complex)
    imgs = []
    for ik in range(2):
        idx_k = o_p.kset_basis(ik)["idx"]
        img = np.array([pos[(2 * int(h) + ik, int(k), int(l))] for h, k, l in idx_k])
        c_s[ik * nst:(ik + 1) * nst][:, img] = cp[ik]
        imgs.append(img)
    r_p = o_p.evaluate(cp, hpsi=True, forces=True, stress=True)
    r_s = o_s.evaluate(c_s, hpsi=True, forces=True, stress=True)
    for t in TERMS:
        print(f"  {t} 2 x cell {2 * r_p[t]:.12e} doubled cell {r_s[t]:.12e}")
        assert abs(r_s[t] - 2.0 * r_p[t]) <= 1e-10 * max(1.0, abs(r_p[t])), t
    assert np.abs(r_s["sigma"] - r_p["sigma"]).max() < 1e-10 * np.abs(r_p["sigma"]).max()
    fp, fs = r_p["fion"], r_s["fion"]
    off_p = off_s = 0
    for n in p.na:
        for copy in range(2):
            assert np.abs(fs[off_s + copy * n: off_s + (copy + 1) * n] -
                          fp[off_p:off_p + n]).max() < 1e-9
        off_p += n
        off_s += 2 * n
    hs = r_s["hpsi"]
    for ik in range(2):
        hp = r_p["hpsi"][ik]
        assert np.abs(hs[ik * nst:(ik + 1) * nst][:, imgs[ik]] - hp).max() < 1e-10 * np.abs(hp).max()


def test_kset_spin_identical_sets_equal_unpolarized(oracle_mod):
    """A k-point set with two spins whose up and down bands are the same, each
    with occupation 1, equals the unpolarized set with occupation 2."""
    s = dataclasses.replace(cf.config(16), kpoint=None)
    kps, wts = [[0.25, 0.1, -0.2], [-0.25, 0.35, 0.1]], [0.25, 0.75]
    o1 = oracle_mod.Oracle(s, kpoints=kps, weights=wts)
    cs = _kset_coeffs(o1, 2, s.nst, 3000)
    s2 = dataclasses.replace(s, nst=2 * s.nst, nspin=2, nst_up=s.nst,
                             occ_bands=np.ones(2 * s.nst))
    o2 = oracle_mod.Oracle(s2, kpoints=kps, weights=wts)
    r1 = o1.evaluate(cs, hpsi=True, forces=True, stress=True)
    r2 = o2.evaluate([np.concatenate([c, c]) for c in cs], hpsi=True, forces=True, stress=True)
    for t in TERMS:
        assert abs(r2[t] - r1[t]) <= 1e-11 * max(1.0, abs(r1[t])), t
    assert np.abs(r2["fion"] - r1["fion"]).max() < 1e-11
    assert np.abs(r2["sigma"] - r1["sigma"]).max() < 1e-11 * np.abs(r1["sigma"]).max()
    for h2, h1 in zip(r2["hpsi"], r1["hpsi"]):
        assert np.abs(h2[:s.nst] - h1).max() < 1e-11 * np.abs(h1).max()
        assert np.abs(h2[s.nst:] - h1).max() < 1e-11 * np.abs(h1).max()

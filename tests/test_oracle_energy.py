"""Pinning the oracle's energy functional (a6-a16) independently of its own
formulas:

  - LDA, spline/splintd and UnitCell against the reference's own sources
    compiled into oracle/_ref (src/LDAFunctional.cpp, src/spline.cpp,
    src/UnitCell.cpp), and sinft/cosft1 against direct trigonometric sums;
  - the spin-polarized LDA against the reference's LDAFunctional with two
    densities, against its own derivative (v_s = d(rho e)/d rho_s by central
    differences) and its zeta = 0 reduction;
  - H*psi is the gradient of the total energy: dE/de along a direction d
    equals sum_n occ_n 2 <d_n|H psi_n> (full-sphere inner product), by
    central differences of complete evaluations;
  - forces are -dE/dtau (central differences at fixed coefficients), the
    short-range ion forces sum to zero (Newton's third law), the total
    charge is the sum of the occupations;
  - nspin 2: identical up and down band sets with occupation 1 reproduce
    nspin 1 with occupation 2; swapping the spin sets leaves the energy
    unchanged; the gradient and force identities hold per spin."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from qbg import configs as cf


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _need_ref(oracle_mod):
    r = oracle_mod.reflib()
    if r is None:
        pytest.skip("reference sources not compiled (oracle/_ref)")
    return r


def full_dot(a, b):
    """Re sum over the full sphere of conj(a) b for half-sphere (Gamma)
    coefficients of real functions: 2 Re sum_half - Re(a0 b0)."""
    return 2.0 * np.real(np.vdot(a, b)) - np.real(np.conj(a[0]) * b[0])


# ------------------------------------------------------------ reference
def test_lda_matches_reference(oracle_mod):
    ref = _need_ref(oracle_mod)
    rng = np.random.default_rng(1)
    rho = np.concatenate([10 ** rng.uniform(-8, 1.5, 4000), [0.0, 1e-30, 0.2387]])
    e, v = oracle_mod.lda(rho)
    er, vr = np.zeros_like(rho), np.zeros_like(rho)
    ref.ref_lda(C.c_int(rho.size), _p(rho), _p(er), _p(vr))
    assert np.allclose(e, er, rtol=1e-13, atol=0)
    assert np.allclose(v, vr, rtol=1e-13, atol=0)


def test_lsda_matches_reference(oracle_mod):
    ref = _need_ref(oracle_mod)
    rng = np.random.default_rng(2)
    n = 4000
    r = 10 ** rng.uniform(-8, 1.5, n)
    z = rng.uniform(-1, 1, n)
    z[:20] = 1.0
    z[20:40] = -1.0
    rup, rdn = 0.5 * r * (1 + z), 0.5 * r * (1 - z)
    rup[40], rdn[40] = 0.0, 0.0
    e, vu, vd = oracle_mod.lsda(rup, rdn)
    er, vur, vdr = np.zeros(n), np.zeros(n), np.zeros(n)
    ref.ref_lsda(C.c_int(n), _p(rup), _p(rdn), _p(er), _p(vur), _p(vdr))
    for a, b in ((e, er), (vu, vur), (vd, vdr)):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-15)


def test_lsda_derivatives_and_limits(oracle_mod):
    rng = np.random.default_rng(3)
    n = 200
    r = 10 ** rng.uniform(-4, 1, n)
    z = rng.uniform(-0.95, 0.95, n)
    rup, rdn = 0.5 * r * (1 + z), 0.5 * r * (1 - z)
    e, vu, vd = oracle_mod.lsda(rup, rdn)
    for which, v in ((0, vu), (1, vd)):
        h = 1e-6 * r
        d = [h if which == 0 else 0 * h, h if which == 1 else 0 * h]
        ep, _, _ = oracle_mod.lsda(rup + d[0], rdn + d[1])
        em, _, _ = oracle_mod.lsda(rup - d[0], rdn - d[1])
        fd = ((r + h) * ep - (r - h) * em) / (2 * h)
        assert np.allclose(v, fd, rtol=2e-6, atol=1e-9), which
    # unpolarized limit is the LDA exactly
    e0, vu0, vd0 = oracle_mod.lsda(r / 2, r / 2)
    el, vl = oracle_mod.lda(r)
    assert np.array_equal(e0, el) and np.array_equal(vu0, vl) and np.array_equal(vd0, vl)
    # fully polarized exchange is 2^(1/3) times the paramagnetic exchange:
    # at high density (rs -> 0) e_x dominates, e(zeta=1)/e(zeta=0) -> 2^(1/3)
    rr = np.array([1e6])
    e1, _, _ = oracle_mod.lsda(rr, 0 * rr)
    ex, _ = oracle_mod.lda(rr)
    assert abs(e1[0] / ex[0] - 2 ** (1 / 3)) < 2e-2


def test_spline_matches_reference(oracle_mod):
    ref = _need_ref(oracle_mod)
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.uniform(0.05, 0.2, 60))
    y = np.sin(x) * np.exp(-0.1 * x)
    for bl, br, yp1, ypn in ((0, 0, 0.3, -0.2), (1, 1, 0.0, 0.0), (0, 1, 1.0, 0.0)):
        y2 = oracle_mod.spline(x, y, yp1, ypn, bl, br)
        y2r = np.zeros_like(x)
        ref.ref_spline(C.c_int(x.size), _p(x.copy()), _p(y.copy()), C.c_double(yp1),
                       C.c_double(ypn), C.c_int(bl), C.c_int(br), _p(y2r))
        assert np.allclose(y2, y2r, rtol=1e-12, atol=1e-14)
        for t in rng.uniform(x[0], x[-1], 20):
            a, da = oracle_mod.splintd(x, y, y2, t)
            yy, dy = C.c_double(), C.c_double()
            ref.ref_splintd(C.c_int(x.size), _p(x.copy()), _p(y.copy()), _p(y2r.copy()),
                            C.c_double(t), C.byref(yy), C.byref(dy))
            assert abs(a - yy.value) < 1e-13 and abs(da - dy.value) < 1e-12


def test_unit_cell_matches_reference(oracle_mod):
    ref = _need_ref(oracle_mod)
    rng = np.random.default_rng(5)
    for _ in range(10):
        a = np.diag(rng.uniform(6, 12, 3)) + rng.uniform(-1.5, 1.5, (3, 3))
        b, vol = oracle_mod.cell(a)
        br, volr = np.zeros(9), C.c_double()
        ref.ref_cell(_p(np.ascontiguousarray(a.ravel())), _p(br), C.byref(volr))
        assert np.allclose(b.ravel(), br, rtol=1e-13, atol=1e-14)
        assert abs(vol - volr.value) < 1e-10 * abs(vol)
        # b = 2 pi A^-T
        assert np.allclose(b, 2 * np.pi * np.linalg.inv(a).T, rtol=1e-12)
        v = rng.uniform(-20, 20, 3)
        w = oracle_mod.fold_in_ws(a, v)
        wr = v.copy()
        ref.ref_fold_in_ws(_p(np.ascontiguousarray(a.ravel())), _p(wr))
        assert np.allclose(w, wr, atol=1e-12)


def test_sine_cosine_transforms_direct(oracle_mod):
    n = 64
    rng = np.random.default_rng(6)
    y = rng.standard_normal(n)
    y[0] = 0.0
    s = oracle_mod.sinft(y)
    k = np.arange(n)
    direct = np.array([np.sum(y * np.sin(np.pi * j * k / n)) for j in range(n)])
    assert np.allclose(s, direct, atol=1e-12)
    yc = rng.standard_normal(n + 1)
    c = oracle_mod.cosft1(yc)
    kk = np.arange(1, n)
    dc = np.array([0.5 * (yc[0] + (-1) ** j * yc[n]) + np.sum(yc[1:n] * np.cos(np.pi * j * kk / n))
                   for j in range(n)])
    assert np.allclose(c[:n], dc, atol=1e-12)


# ------------------------------------------------------- energy identities
def _system(idx, spin=None):
    s = cf.config(idx)
    if spin is not None:
        nel, dspin, nempty = spin
        s = cf.spin_polarized(s, nel, dspin, nempty)
    return s


def _coeffs(oracle_mod, s, seed):
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.random_coefficients(seed, b["g2"], s.nst)
    if getattr(s, "nspin", 1) == 2:
        u = s.nst_up
        c = np.concatenate([cf.gram_numpy(c[:u]), cf.gram_numpy(c[u:])])
    else:
        c = cf.gram_numpy(c)
    return o, c


CASES = [(11, None), (12, None), (11, (9, 1, 1)), (10, (5, 0, 1)), (13, None), (14, None),
         (15, None), (15, (12, 1, 1))]


@pytest.mark.parametrize("idx,spin", CASES)
def test_hpsi_is_energy_gradient(oracle_mod, idx, spin):
    s = _system(idx, spin)
    o, c = _coeffs(oracle_mod, s, 300 + idx)
    r = o.evaluate(c, hpsi=True, forces=False)
    assert abs(r["total_charge"] - s.occ.sum()) < 1e-9
    rng = np.random.default_rng(7)
    d = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    d[:, 0] = d[:, 0].real
    d *= 1e-4 / np.abs(d).max()
    pred = sum(s.occ[n] * 2.0 * full_dot(d[n], r["hpsi"][n]) for n in range(s.nst))
    h = 1.0
    ep = o.evaluate(c + h * d, hpsi=False, forces=False)["etotal"]
    em = o.evaluate(c - h * d, hpsi=False, forces=False)["etotal"]
    fd = (ep - em) / (2 * h)
    print(f"  dE/de fd {fd:.10e} pred {pred:.10e}")
    assert abs(fd - pred) <= 1e-6 * abs(pred)


@pytest.mark.parametrize("idx,spin", CASES)
def test_forces_are_energy_gradient(oracle_mod, idx, spin):
    s = _system(idx, spin)
    o, c = _coeffs(oracle_mod, s, 400 + idx)
    r = o.evaluate(c, hpsi=False, forces=True)
    assert np.abs(o.get(11).reshape(-1, 3).sum(axis=0)).max() < 1e-10   # esr pairs
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


@pytest.mark.parametrize("idx,spin", [(11, None), (15, None), (15, (12, 1, 1))])
def test_dxc_is_exc_minus_valence_potential_energy(oracle_mod, idx, spin):
    """dxc = exc - omega/N sum_s sum_r rho_s(r) vxc_s(r) over the valence
    densities: with a core density the reference adds the core part of
    sum rho vxc back in G space (src/EnergyFunctional.cpp:275-306), so the
    identity checks that term, its factor 2, its G = 0 term and its spin
    factor against a real-space sum."""
    s = _system(idx, spin)
    o, c = _coeffs(oracle_mod, s, 600 + idx)
    r = o.evaluate(c, hpsi=False, forces=False)
    omega = abs(np.linalg.det(s.cell))
    pv = np.sum(o.get(0) * o.get(2))
    if getattr(s, "nspin", 1) == 2:
        pv += np.sum(o.get(14) * o.get(16))
    ref = r["exc"] - omega / o.N * pv
    print(f"  dxc {r['dxc']:.14e} exc - <rho vxc> {ref:.14e}")
    assert abs(r["dxc"] - ref) <= 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("idx", [11, 15])
def test_spin_identical_sets_equal_unpolarized(oracle_mod, idx):
    s1 = cf.config(idx)
    o1, c = _coeffs(oracle_mod, s1, 500)
    r1 = o1.evaluate(c, hpsi=True, forces=True, stress=True)
    s2 = dataclasses.replace(s1, nst=2 * s1.nst, nspin=2, nst_up=s1.nst,
                             occ_bands=np.ones(2 * s1.nst))
    o2 = oracle_mod.Oracle(s2)
    r2 = o2.evaluate(np.concatenate([c, c]), hpsi=True, forces=True, stress=True)
    for t in ["ekin", "eps", "enl", "ehart", "exc", "dxc", "etotal", "total_charge"]:
        assert abs(r2[t] - r1[t]) <= 1e-11 * max(1.0, abs(r1[t])), t
    assert np.allclose(r2["fion"], r1["fion"], rtol=0, atol=1e-11)
    assert np.allclose(r2["sigma"], r1["sigma"], rtol=0, atol=1e-11)
    n = s1.nst
    assert np.allclose(r2["hpsi"][:n], r1["hpsi"], rtol=0, atol=1e-11)
    assert np.allclose(r2["hpsi"][n:], r1["hpsi"], rtol=0, atol=1e-11)
    assert np.allclose(o2.get(0), o2.get(14), rtol=0, atol=1e-14)        # rho up = down


def test_spin_swap_symmetry(oracle_mod):
    s = _system(11, (9, 1, 1))
    o, c = _coeffs(oracle_mod, s, 600)
    r = o.evaluate(c, hpsi=True, forces=True)
    u = s.nst_up
    occ = s.occ
    sw = dataclasses.replace(s, nst_up=s.nst - u, occ_bands=np.concatenate([occ[u:], occ[:u]]))
    osw = oracle_mod.Oracle(sw)
    rsw = osw.evaluate(np.concatenate([c[u:], c[:u]]), hpsi=True, forces=True)
    assert abs(rsw["etotal"] - r["etotal"]) <= 1e-11 * abs(r["etotal"])
    assert np.allclose(rsw["fion"], r["fion"], atol=1e-11)
    assert np.allclose(rsw["hpsi"][s.nst - u:], r["hpsi"][:u], atol=1e-11)
    # the polarized system differs from its unpolarized density's LDA
    assert np.abs(o.get(0) - o.get(14)).max() > 1e-3


# ---------------------------------------------------------- nspin 2 SCF
def scf_spin(oracle_mod, s, o, c, kind, param, nitscf, nite=0, ndim=3, delta_ratio=0.01):
    """BOSampleStepper's loop (src/BOSampleStepper.cpp:782-977) with one
    wavefunction stepper per spin (the spin loop of wf_stepper->update):
    charge mixing of both spins' densities in one vector when nite > 0, then
    up to max(nite, 1) steps at fixed potential, stopped early when the
    eigenvalue sum changes by less than delta_ratio times the change of the
    Hartree energy; returns bands and energies."""
    g2 = o.basis(0)["g2"]
    sets = [(0, s.nst_up), (s.nst_up, s.nst - s.nst_up)] if s.nspin == 2 else [(0, s.nst)]
    sts = [oracle_mod.WfStepper(kind, param, g2) for _ in sets]
    mx = o.mixer(0.5, 10.0, ndim, 3.0) if nite > 0 else None
    for st in sts:
        st.preprocess()
    es = []
    c = c.copy()
    ehart = delta_ehart = 0.0
    for itscf in range(nitscf):
        o.update_density(c)
        if mx is not None:
            mx.update()
        o.update_vhxc(False)
        if nite > 0:
            for st in sts:
                st.preprocess()
        eig_m = None
        for ite in range(max(nite, 1)):
            r = o.energy(c, hpsi=True, forces=False)
            if ite == 0:
                if itscf == 0:
                    ehart, delta_ehart = r["ehart"], 0.0
                elif itscf == 1:
                    ehart, delta_ehart = r["ehart"], abs(r["ehart"] - ehart)
                else:
                    ehart, delta_ehart = r["ehart"], min(delta_ehart, abs(r["ehart"] - ehart))
            eig = oracle_mod.sd_dot(c, r["hpsi"])
            for st, (b0, n) in zip(sts, sets):
                cn, _ = st.update(c[b0:b0 + n], r["hpsi"][b0:b0 + n])
                c[b0:b0 + n] = cn
            if ite > 0 and abs(eig - eig_m) < delta_ratio * delta_ehart:
                break
            eig_m = eig
        es.append(r["etotal"])
    return c, es


def test_spin_scf_identical_sets_follow_nspin1(oracle_mod):
    s1 = cf.config(10)
    o1, c = _coeffs(oracle_mod, s1, 900)
    _, e1 = scf_spin(oracle_mod, s1, o1, c, "psda", 0.0, 8)
    s2 = dataclasses.replace(s1, nst=2 * s1.nst, nspin=2, nst_up=s1.nst,
                             occ_bands=np.ones(2 * s1.nst))
    o2 = oracle_mod.Oracle(s2)
    _, e2 = scf_spin(oracle_mod, s2, o2, np.concatenate([c, c]), "psda", 0.0, 8)
    assert np.allclose(e1, e2, rtol=1e-11, atol=0), (e1, e2)


def test_spin_mixing_identical_sets_follow_nspin1(oracle_mod):
    """Charge mixing of two identical spin densities (both in one Anderson
    vector) follows the nspin 1 SCF with mixing: the spin halves carry half
    the total, the Anderson sums scale by one half and its solution does not
    change."""
    from test_oracle_scf import scf
    s1 = cf.config(10)
    o1, c = _coeffs(oracle_mod, s1, 903)
    _, e1 = scf(oracle_mod, s1, o1, c, "psda", 0.0, 6, 3)
    s2 = dataclasses.replace(s1, nst=2 * s1.nst, nspin=2, nst_up=s1.nst,
                             occ_bands=np.ones(2 * s1.nst))
    o2 = oracle_mod.Oracle(s2)
    _, e2 = scf_spin(oracle_mod, s2, o2, np.concatenate([c, c]), "psda", 0.0, 6, 3)
    assert np.allclose(e1, e2, rtol=1e-10, atol=0), (e1, e2)


def test_spin_mixing_first_step_is_kerker(oracle_mod):
    """The second mixer call with ndim 3 takes the Kerker step on each spin,
    rhog_s = rin_s + alpha wk (rout_s - rin_s): the total rhog follows the
    same step; a spin whose bands did not change keeps its density; the spin
    densities in r add up to the inverse transform of the mixed total."""
    s = _system(11, (9, 1, 1))
    o, c = _coeffs(oracle_mod, s, 904)
    _, c2 = _coeffs(oracle_mod, s, 905)
    u = s.nst_up
    c2[:u] = c[:u]                        # only the down bands change
    alpha, rck = 0.4, 6.0
    g2 = o.basis(1)["g2"]
    q0 = 2 * np.pi / rck
    wk = np.where(g2 != 0.0, g2 / (g2 + q0 * q0), 1.0)
    o.update_density(c2)
    g_out, dn_out = o.get(3).copy(), o.get(14).copy()
    mx = o.mixer(alpha, rck, 3, 3.0)
    o.update_density(c)
    g_in, up_in, dn_in = o.get(3).copy(), o.get(0).copy(), o.get(14).copy()
    mx.update()                           # memorizes
    o.update_density(c2)
    mx.update()                           # mixes
    up, dn, g = o.get(0).copy(), o.get(14).copy(), o.get(3).copy()
    assert np.allclose(g, g_in + alpha * wk * (g_out - g_in), rtol=1e-11, atol=1e-14)
    assert np.abs(up - up_in).max() < 1e-12 * np.abs(up_in).max()
    assert np.abs(dn - dn_in).max() > 1e-3 * np.abs(dn_out - dn_in).max()
    o.update_rhor()                       # rhor (up slot) = inverse transform of the total
    assert np.abs(o.get(0) - (up + dn)).max() < 1e-12 * np.abs(up + dn).max()


def test_spin_scf_reaches_stationary_point(oracle_mod):
    """A spin-polarized SCF (PSDA, 40 iterations) converges, and random
    perturbations of the final bands (kept orthonormal per spin) raise the
    energy."""
    s = _system(11, (9, 1, 1))
    o, c = _coeffs(oracle_mod, s, 901)
    c, es = scf_spin(oracle_mod, s, o, c, "psda", 0.0, 60)
    assert abs(es[-1] - es[-2]) < 1e-7 * abs(es[-1])
    o.update_density(c)
    o.update_vhxc(False)
    e0 = o.energy(c, hpsi=False, forces=False)["etotal"]
    u = s.nst_up
    rng = np.random.default_rng(5)
    for _ in range(2):
        d = 1e-3 * (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
        d[:, 0] = d[:, 0].real
        cp = c + d
        cp = np.concatenate([cf.gram_numpy(cp[:u]), cf.gram_numpy(cp[u:])])
        o.update_density(cp)
        o.update_vhxc(False)
        assert o.energy(cp, hpsi=False, forces=False)["etotal"] > e0

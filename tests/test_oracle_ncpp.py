"""Norm-conserving species read from the reference's own test potentials
(test/potentials/*.xml, copied under tests/golden/) and their
Kleinman-Bylander form in the oracle (Species::initialize_ncpp,
src/Species.cpp:71-507), pinned by independent quadratures:

  - the weight of each nonlocal channel is 1 / <phi_l| v_l - v_loc |phi_l>;
  - the projector's transform equals 4 pi int r^2 j_l(Gr) (v_l - v_loc) phi_l dr
    (direct quadrature with scipy's spherical Bessel functions), which checks
    the sine/cosine-transform route of the reference;
  - the KB operator reproduces the semilocal channel on its own reference
    state: w_l <phi_l|dv phi_l>^2 = <phi_l|dv|phi_l>;
and whole systems of the reference's tests (test/si4gs, test/h2ogs) hold
their electron count and a variational SCF minimum."""
from pathlib import Path

import numpy as np
import pytest
from scipy.integrate import simpson
from scipy.special import spherical_jn

from qbg import configs as cf
from qbg.species import read_species

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def si():
    return read_species(GOLD / "Si_VBC_LDA-1.0.xml")


@pytest.fixture(scope="module")
def oxy():
    return read_species(GOLD / "O_HSCV_PBE-1.0.xml")


@pytest.fixture(scope="module")
def hyd():
    return read_species(GOLD / "H_HSCV_PBE-1.0.xml")


def test_read_species(si, oxy, hyd):
    assert (si.symbol, si.zval, si.lmax, si.llocal, si.nmesh, si.deltar) == \
        ("Si", 4, 2, 2, 1000, 0.01)
    assert (oxy.zval, oxy.lmax, oxy.llocal, oxy.nmesh) == (6, 1, 1, 2208)
    assert (hyd.zval, hyd.lmax, hyd.llocal, hyd.nmesh) == (1, 0, 0, 2537)
    # the first and last mesh values as printed in the file
    assert si.vps[0][0] == 6.01180353958
    r = np.arange(si.nmesh) * si.deltar
    # pseudo-wavefunctions are normalized: int phi^2 r^2 dr = 1
    for l in (0, 1):
        assert abs(simpson(si.phi[l] ** 2 * r * r, x=r) - 1.0) < 1e-3
    # tails approach -Z/r
    assert abs(si.vps[2][-1] * r[-1] + si.zval) < 1e-2


def test_core_density_xml_round_trip(si):
    """A <core_density> element (src/SpeciesHandler.cpp:214-220) is read,
    written where Species::print puts it (src/Species.cpp:1256, 1310-1316)
    and read back unchanged; species without one have none."""
    import dataclasses
    from qbg.species import parse_species, species_xml
    assert si.core_density() is None
    r = np.arange(si.nmesh) * si.deltar
    core = 0.3 * np.exp(-(r / 1.1) ** 2)
    text = species_xml(dataclasses.replace(si, nlcc=core), "si_core")
    assert text.index("<core_density") < text.index("<projector")
    back = parse_species(text)
    assert np.array_equal(back.core_density(), core)
    assert np.array_equal(back.vps, si.vps) and np.array_equal(back.phi, si.phi)


@pytest.mark.parametrize("name", ["si", "oxy"])
def test_kb_weights_and_identity(oracle_mod, name, request):
    sp = request.getfixturevalue(name)
    proj, dl = oracle_mod.ncpp_kb(sp)
    r = np.arange(sp.nmesh) * sp.deltar
    vl = sp.vps[sp.llocal]
    for l in range(proj.shape[0]):
        if l == sp.llocal:
            assert dl[l] == 0.0 and not proj[l].any()
            continue
        dv = sp.vps[l] - vl
        e = simpson(dv * sp.phi[l] ** 2 * r * r, x=r)
        assert abs(dl[l] * e - 1.0) < 1e-6, (l, dl[l], e)
        assert np.allclose(proj[l], dv * sp.phi[l], rtol=0, atol=1e-15)
        # KB acting on its reference state reproduces the semilocal value
        beta_phi = simpson(proj[l] * sp.phi[l] * r * r, x=r)
        assert abs(dl[l] * beta_phi ** 2 - e) < 1e-6 * abs(e)


def si_with_d(si):
    """Si_VBC_LDA with llocal = 0, so the p and d channels are nonlocal; the
    file holds no d pseudo-wavefunction, so a normalized r^2 exp(-r^2/2.88)
    stands in for it (a synthetic d channel for the l = 2 code paths)."""
    import dataclasses
    r = np.arange(si.nmesh) * si.deltar
    phi2 = r * r * np.exp(-r * r / 2.88)
    phi2 /= np.sqrt(simpson(phi2 ** 2 * r * r, x=r))
    phi = np.array(si.phi, dtype=float).copy()
    phi[2] = phi2
    return dataclasses.replace(si, llocal=0, phi=phi)


@pytest.mark.parametrize("with_d", [False, True])
def test_projector_transform_by_quadrature(oracle_mod, si, with_d):
    """The oracle's projector tables w_l(G) (sinft/cosft route) against direct
    quadrature of 4 pi int r^2 j_l(Gr) f_l(r) dr, for Si as given (s and p
    nonlocal) and with nonlocal p and d channels."""
    sp = si_with_d(si) if with_d else si
    llocal = sp.llocal
    s = cf.si4gs(sp)
    o = oracle_mod.Oracle(s)
    proj, dl = oracle_mod.ncpp_kb(sp)
    r = np.arange(sp.nmesh) * sp.deltar
    g = np.array([0.3, 1.1, 2.7, 4.0, 6.5])
    v, dv, vn, dvn = o.species_eval(0, g)
    for l in range(proj.shape[0]):
        if l == llocal:
            assert not vn[l].any()
            continue
        direct = np.array([4 * np.pi * simpson(r * r * spherical_jn(l, gg * r) * proj[l], x=r)
                           for gg in g])
        assert np.allclose(vn[l], direct, rtol=2e-5, atol=1e-6), (l, vn[l], direct)


def test_si_d_channel_system(oracle_mod, si):
    """Si with nonlocal p and d channels in the si4gs cell: charge, and
    forces as the gradient of the energy."""
    import dataclasses
    s = cf.si4gs(si_with_d(si))
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(6, o.basis(0)["g2"], s.nst))
    r = o.evaluate(c, hpsi=False, forces=True)
    assert abs(r["total_charge"] - 2 * s.nst) < 1e-8
    assert abs(r["enl"]) > 1e-3
    h = 1e-4
    t = np.array(s.tau, dtype=float)
    for a, j in ((0, 0), (1, 1)):
        e = []
        for sgn in (1.0, -1.0):
            tt = t.copy()
            tt[a, j] += sgn * h
            e.append(oracle_mod.Oracle(dataclasses.replace(s, tau=tt)).evaluate(
                c, hpsi=False, forces=False)["etotal"])
        fd = -(e[0] - e[1]) / (2 * h)
        assert abs(fd - r["fion"][a, j]) <= 1e-6 * max(1.0, abs(fd)), (a, j, fd)


def test_si4_and_h2o_charge(oracle_mod, si, oxy, hyd):
    for s in (cf.si4gs(si), cf.h2ogs(oxy, hyd, ecut_ry=30.0)):
        o = oracle_mod.Oracle(s)
        b = o.basis(0)
        c = cf.gram_numpy(cf.random_coefficients(3, b["g2"], s.nst))
        r = o.evaluate(c, hpsi=True, forces=True)
        assert abs(r["total_charge"] - 2 * s.nst) < 1e-8
        assert np.isfinite(r["etotal"])


def test_o2gs_triplet_charge_and_moment(oracle_mod, oxy):
    """test/o2gs/test.i: 12 electrons, delta_spin 1, nempty 4: 11 bands up
    (7 occupied), 9 down (5 occupied); the densities integrate to 7 and 5."""
    s = cf.o2gs(oxy, ecut_ry=30.0)
    assert (s.nst, s.nst_up) == (20, 11)
    assert s.occ.sum() == 12 and s.occ[:11].sum() == 7
    o = oracle_mod.Oracle(s)
    c = cf.random_coefficients(5, o.basis(0)["g2"], s.nst)
    c = np.concatenate([cf.gram_numpy(c[:11]), cf.gram_numpy(c[11:])])
    r = o.evaluate(c, hpsi=True, forces=True)
    assert abs(r["total_charge"] - 12) < 1e-8
    dv = abs(np.linalg.det(np.asarray(s.cell))) / o.get(0).size
    assert abs(o.get(0).sum() * dv - 7) < 1e-8 and abs(o.get(14).sum() * dv - 5) < 1e-8


def test_si4_scf_is_variational(oracle_mod, si):
    """The settings of test/si4gs/test.i: PSDA, ecutprec 2 Ry (1 Ha),
    run 0 200 (200 SCF iterations of one step, no charge mixing); the
    minimum reached is -15.3729448 Ha."""
    from test_oracle_scf import perturbed_energy, scf
    s = cf.si4gs(si)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(4, b["g2"], s.nst))
    c, es = scf(oracle_mod, s, o, c, "psda", 1.0, 200, 0)
    assert abs(es[-1] - es[-2]) < 1e-7, es[-4:]
    assert abs(es[-1] - (-15.3729448)) < 1e-6
    o.update_density(c)
    o.update_vhxc(False)
    e0 = o.energy(c, hpsi=False, forces=False)["etotal"]
    for t in range(2):
        assert perturbed_energy(oracle_mod, o, c, 1e-2, t) > e0


# ----------------------------------------- semilocal quadrature (nquad > 0)
def _quad_si4gs(si, llocal):
    import dataclasses
    return cf.si4gs(dataclasses.replace(si, nquad=16, rquad=2.6, llocal=llocal))


@pytest.mark.parametrize("llocal", [2, 0])
def test_quadrature_forces_and_hpsi_are_gradients(oracle_mod, si, llocal):
    """The quadrature projectors (src/NonLocalPotential.cpp:317-960) in a
    whole evaluation: forces are -dE/dtau and H*psi is dE/dc* by central
    differences (p and s, or p and d channels nonlocal)."""
    import dataclasses
    s = _quad_si4gs(si, llocal)
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(17, o.basis(0)["g2"], s.nst))
    r = o.evaluate(c, hpsi=True, forces=True)
    h = 1e-4
    for a, j in ((0, 0), (1, 1), (3, 2)):
        e = []
        for sgn in (1.0, -1.0):
            t = np.array(s.tau, dtype=float)
            t[a, j] += sgn * h
            e.append(oracle_mod.Oracle(dataclasses.replace(s, tau=t)).evaluate(
                c, hpsi=False, forces=False)["etotal"])
        fd = -(e[0] - e[1]) / (2 * h)
        print(f"  f[{a},{j}] fd {fd:.10e} oracle {r['fion'][a, j]:.10e}")
        assert abs(fd - r["fion"][a, j]) <= 1e-6 * max(1.0, abs(fd))
    rng = np.random.default_rng(3)
    d = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    d[:, 0] = d[:, 0].real
    d *= 1e-4 / np.abs(d).max()
    full = lambda x, y: 2.0 * np.real(np.vdot(x, y)) - np.real(np.conj(x[0]) * y[0])
    pred = sum(s.occ[n] * 2.0 * full(d[n], r["hpsi"][n]) for n in range(s.nst))
    ep = o.evaluate(c + d, hpsi=False, forces=False)["etotal"]
    em = o.evaluate(c - d, hpsi=False, forces=False)["etotal"]
    assert abs((ep - em) / 2 - pred) <= 1e-6 * abs(pred)


def test_quadrature_stress_vs_strain_fd(oracle_mod, si):
    """sigma_enl and the total stress of a quadrature species against
    Richardson-extrapolated central strain differences (the method of
    test/si2stress/check_stress.sh:20-90)."""
    from test_oracle_stress import strained, sym
    s = _quad_si4gs(si, 0)
    ref = s.cell * 1.04
    o0 = oracle_mod.Oracle(s, refcell=ref)
    c = cf.gram_numpy(cf.random_coefficients(19, o0.basis(0)["g2"], s.nst))
    sig = o0.evaluate(c, hpsi=False, forces=False, stress=True)["sigma"]
    omega = abs(np.linalg.det(s.cell))
    v = 1e-4 * np.random.default_rng(5).uniform(-1.0, 1.0, 6)

    def diff(u):
        ep = oracle_mod.Oracle(strained(s, u), refcell=ref).evaluate(c, hpsi=False, forces=False)
        em = oracle_mod.Oracle(strained(s, -u), refcell=ref).evaluate(c, hpsi=False, forces=False)
        return {t: ep[t] - em[t] for t in ("enl", "etotal")}

    d1, d2 = diff(sym(v)), diff(sym(0.5 * v))
    for term, row in (("enl", 3), ("etotal", 0)):
        de = (8.0 * d2[term] - d1[term]) / 3.0
        st = sig[row]
        pred = -2.0 * omega * (sum(v[k] * st[k] for k in range(3)) +
                               2.0 * sum(v[k] * st[k] for k in range(3, 6)))
        print(f"  {term}: fd {de:.10e} sigma {pred:.10e}")
        assert abs(de - pred) <= 2e-7 * abs(de) + 1e-12

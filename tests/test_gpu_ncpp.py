"""Norm-conserving species from the reference's test potentials on the
device: the systems of test/si4gs, test/h2ogs and test/o2gs (nspin 2)
evaluated against the FP64
oracle (energies, H*psi, forces to 1e-10; the mixed modes to the 1e-6
Ha/atom bar), the si4gs SCF run (PSDA, ecutprec 2 Ry, 200 iterations)
reaching the oracle's minimum, Si with nonlocal p and d channels, and
species the library does not take (nonlocal f)."""
import dataclasses
from pathlib import Path

import numpy as np
import pytest

from qbg import configs as cf
from qbg.species import read_species

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
TERMS = ["ekin", "eps", "enl", "ehart", "exc", "esr", "eself", "etotal", "total_charge"]


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def systems():
    si = read_species(GOLD / "Si_VBC_LDA-1.0.xml")
    o = read_species(GOLD / "O_HSCV_PBE-1.0.xml")
    h = read_species(GOLD / "H_HSCV_PBE-1.0.xml")
    return {"si4gs": cf.si4gs(si), "h2ogs": cf.h2ogs(o, h, ecut_ry=70.0),
            "o2gs": cf.o2gs(o, ecut_ry=70.0)}


@pytest.mark.parametrize("name", ["si4gs", "h2ogs", "o2gs"])
@pytest.mark.parametrize("prec", ["fp64", "bf16x2"])
def test_evaluate(ctx, oracle_mod, name, prec):
    import qbg
    s = systems()[name]
    print(f"[case] {name} {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.random_coefficients(11, b["g2"], s.nst)
    if s.nspin == 2:
        u = s.nst_up
        c = np.concatenate([cf.gram_numpy(c[:u]), cf.gram_numpy(c[u:])])
    else:
        c = cf.gram_numpy(c)
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, stress=True, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True, stress=True)
    for t in TERMS:
        if prec == "fp64":
            assert abs(got[t] - ref[t]) <= 1e-10 * max(abs(ref[t]), 1e-12), (t, got[t], ref[t])
        else:
            assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
    ftol, htol = (1e-10, 1e-10) if prec == "fp64" else (1e-5, 1e-5)
    assert np.abs(got["fion"] - ref["fion"]).max() <= ftol * max(1.0, np.abs(ref["fion"]).max())
    assert relerr(hp.download(), ref["hpsi"]) < htol
    stol = 1e-10 if prec == "fp64" else 1e-5
    assert np.abs(got["sigma"] - ref["sigma"]).max() < stol * np.abs(ref["sigma"][0]).max()
    ef.release()


def test_si4gs_scf(ctx, oracle_mod):
    """test/si4gs/test.i: PSDA, ecutprec 2 Ry, run 0 200, from random
    coefficients; the oracle's loop ends at -15.3729448 Ha."""
    import qbg
    from qbg.scf import ScfLoop
    s = systems()["si4gs"]
    print("[case] si4gs scf", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    b = ef.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(4, b["g2"], s.nst))
    dc = ctx.to_device(c)
    dh = ctx.array(c.shape, np.complex128)
    loop = ScfLoop(ef, dc, dh, wf_dyn="psda", param=1.0, graphs=True)
    hist = loop.run(200, 0)
    e = [h["etotal"] for h in hist]
    assert abs(e[-1] - e[-2]) < 1e-7
    assert abs(e[-1] - (-15.3729448)) < 1e-6, e[-1]
    loop.release()
    ef.release()


@pytest.mark.parametrize("prec", ["fp64", "bf16x2"])
def test_d_channel_evaluate(ctx, oracle_mod, prec):
    """Si with nonlocal p and d channels (a synthetic d pseudo-wavefunction,
    tests/test_oracle_ncpp.py) in the si4gs cell against the oracle: the
    l = 2 projector tables, the (-i)^2 phase and the d strain derivatives."""
    import qbg
    from test_oracle_ncpp import si_with_d
    s = cf.si4gs(si_with_d(read_species(GOLD / "Si_VBC_LDA-1.0.xml")))
    print(f"[case] si4gs with d projectors {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    assert ef.nproj == 4 * 9
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(13, o.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, stress=True, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True, stress=True)
    for t in TERMS:
        if prec == "fp64":
            assert abs(got[t] - ref[t]) <= 1e-10 * max(abs(ref[t]), 1e-12), (t, got[t], ref[t])
        else:
            assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
    ftol = 1e-10 if prec == "fp64" else 1e-5
    assert np.abs(got["fion"] - ref["fion"]).max() <= ftol * max(1.0, np.abs(ref["fion"]).max())
    assert relerr(hp.download(), ref["hpsi"]) < ftol
    assert np.abs(got["sigma"] - ref["sigma"]).max() < ftol * np.abs(ref["sigma"][0]).max()
    ef.release()


@pytest.mark.parametrize("prec", ["fp64", "bf16x2"])
def test_f_projectors_evaluate(ctx, oracle_mod, prec):
    """A semilocal species with s, p, d and f projectors (Gaussians r^l
    exp(-r^2/2s^2), src/Species.cpp:948-1046) beside an s/p species in an
    orthorhombic cell against the oracle: the j_3 transform, the seven f
    harmonics of twnl, the (-i)^3 phase, and the f strain derivatives (in the
    sign the strain finite differences confirm, tests/test_oracle_stress.py)."""
    import qbg
    s = cf.config(14)
    print(f"[case] {s.name} {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    assert ef.nproj == 2 * 16 + 1 * 4
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(14, o.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, stress=True, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True, stress=True)
    for t in TERMS:
        if prec == "fp64":
            assert abs(got[t] - ref[t]) <= 1e-10 * max(abs(ref[t]), 1e-12), (t, got[t], ref[t])
        else:
            assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
    ftol = 1e-10 if prec == "fp64" else 1e-5
    assert np.abs(got["fion"] - ref["fion"]).max() <= ftol * max(1.0, np.abs(ref["fion"]).max())
    assert relerr(hp.download(), ref["hpsi"]) < ftol
    assert np.abs(got["sigma"] - ref["sigma"]).max() < ftol * np.abs(ref["sigma"][0]).max()
    ef.release()


@pytest.mark.parametrize("prec", ["fp64", "bf16x2"])
def test_core_density_ncpp(ctx, oracle_mod, prec):
    """Si (norm-conserving, Kleinman-Bylander) with a core density
    0.3 exp(-(r/1.1)^2) on its mesh in the si4gs cell against the oracle:
    the non-linear core correction of a norm-conserving species
    (src/Species.cpp:194, 1056-1100; src/EnergyFunctional.cpp:275-306,
    650-659, 855-856, 1064-1083), energies, forces, H*psi and stress."""
    import qbg
    si = read_species(GOLD / "Si_VBC_LDA-1.0.xml")
    r = np.arange(si.nmesh) * si.deltar
    s = cf.si4gs(dataclasses.replace(si, nlcc=0.3 * np.exp(-(r / 1.1) ** 2)))
    print(f"[case] si4gs with a core density {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(15, o.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, stress=True, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True, stress=True)
    for t in TERMS:
        if prec == "fp64":
            assert abs(got[t] - ref[t]) <= 1e-10 * max(abs(ref[t]), 1e-12), (t, got[t], ref[t])
        else:
            assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
    ftol = 1e-10 if prec == "fp64" else 1e-5
    assert np.abs(got["fion"] - ref["fion"]).max() <= ftol * max(1.0, np.abs(ref["fion"]).max())
    assert relerr(hp.download(), ref["hpsi"]) < ftol
    assert np.abs(got["sigma"] - ref["sigma"]).max() < ftol * np.abs(ref["sigma"][0]).max()
    if prec == "fp64":
        assert relerr(ef.get(11), o.get(17)) < 1e-10
    ef.release()


@pytest.mark.parametrize("llocal,prec", [(2, "fp64"), (0, "fp64"), (0, "bf16x2")])
def test_quadrature_evaluate(ctx, oracle_mod, llocal, prec):
    """Si in the semilocal quadrature form (nquad = 16, rquad = 2.6;
    src/NonLocalPotential.cpp:104-229, 317-960) in the si4gs cell against
    the oracle: 64 (s, p) or 128 (p, d) projectors per atom, their weights
    from the splined semilocal potentials, energies, forces, H*psi and all
    seven stress rows."""
    import qbg
    si = read_species(GOLD / "Si_VBC_LDA-1.0.xml")
    s = cf.si4gs(dataclasses.replace(si, nquad=16, rquad=2.6, llocal=llocal))
    print(f"[case] si4gs quadrature form llocal {llocal} {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    assert ef.nproj == 4 * 16 * (4 if llocal == 2 else 8)
    o = oracle_mod.Oracle(s)
    c = cf.gram_numpy(cf.random_coefficients(16, o.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, stress=True, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True, stress=True)
    for t in TERMS:
        if prec == "fp64":
            assert abs(got[t] - ref[t]) <= 1e-10 * max(abs(ref[t]), 1e-12), (t, got[t], ref[t])
        else:
            assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
    ftol = 1e-10 if prec == "fp64" else 1e-5
    assert np.abs(got["fion"] - ref["fion"]).max() <= ftol * max(1.0, np.abs(ref["fion"]).max())
    assert relerr(hp.download(), ref["hpsi"]) < ftol
    assert np.abs(got["sigma"] - ref["sigma"]).max() < ftol * np.abs(ref["sigma"][0]).max()
    ef.release()


def test_f_channel_refused(ctx):
    """A nonlocal f channel (l = 3) of a norm-conserving potential is
    refused: the reference's Kleinman-Bylander construction
    (src/Species.cpp:330-505) has no l = 3 transform."""
    import qbg
    si = read_species(GOLD / "Si_VBC_LDA-1.0.xml")
    vps = np.vstack([si.vps, si.vps[2:3]])
    phi = np.vstack([si.phi, si.phi[1:2]])
    bad = dataclasses.replace(si, lmax=3, llocal=0, vps=vps, phi=phi)
    with pytest.raises(qbg.QbgError) as e:
        qbg.EnergyFunctional(ctx, cf.si4gs(bad))
    assert e.value.code == -1


def test_o2gs_scf_with_mixing(ctx, oracle_mod):
    """test/o2gs/test.i's run (nspin 2, charge mixing, five steps per SCF
    iteration at fixed potential) at ecut 30 Ry for six SCF iterations
    against the oracle's loop, per iteration."""
    import qbg
    from qbg.scf import ScfLoop
    from test_oracle_energy import scf_spin
    o_sp = read_species(GOLD / "O_HSCV_PBE-1.0.xml")
    s = cf.o2gs(o_sp, ecut_ry=30.0)
    print("[case] o2gs scf", flush=True)
    o = oracle_mod.Oracle(s)
    c = cf.random_coefficients(12, o.basis(0)["g2"], s.nst)
    u = s.nst_up
    c = np.concatenate([cf.gram_numpy(c[:u]), cf.gram_numpy(c[u:])])
    co, es = scf_spin(oracle_mod, s, o, c, "psda", 0.0, 6, 5)
    ef = qbg.EnergyFunctional(ctx, s)
    dc = ctx.to_device(c)
    dh = ctx.array(c.shape, np.complex128)
    loop = ScfLoop(ef, dc, dh, wf_dyn="psda", param=0.0, graphs=True)
    hist = loop.run(6, 5)
    ed = [h["etotal"] for h in hist]
    assert np.allclose(ed, es, rtol=1e-9, atol=0), (ed, es)
    assert ed[-1] < ed[0]
    loop.release()
    ef.release()

"""nspin 2 on the device (qbg_ef_set_nspin) against the oracle: per-spin
densities and potentials, the spin-polarized LDA, energies, forces, stress
and H*psi to 1e-10 in FP64 and to the mixed-precision bar in the split modes;
the H*psi pass from the kept transforms of both spins equal to the
recomputing pass; identical up and down sets equal to nspin 1; the SCF loop
with one stepper per spin, without and with charge mixing of both spins,
against the oracle."""
import dataclasses

import numpy as np
import pytest

from qbg import configs as cf

pytestmark = pytest.mark.gpu

TERMS = ["ekin", "eps", "enl", "ehart", "exc", "dxc", "esr", "etotal", "total_charge"]


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def spin_coeffs(s, g2, seed):
    c = cf.random_coefficients(seed, g2, s.nst)
    u = s.nst_up
    return np.concatenate([cf.gram_numpy(c[:u]), cf.gram_numpy(c[u:])])


# t5: core densities (each spin carries half the core density)
SYSTEMS = {"t1": (11, 9, 1, 1), "t2": (12, 11, 1, 0), "c2": (1, 255, 1, 2),
           "t5": (15, 12, 1, 1)}


@pytest.mark.parametrize("name,prec", [("t1", "fp64"), ("t2", "fp64"), ("t1", "bf16x2"),
                                       ("c2", "bf16x2"), ("t5", "fp64"), ("t5", "bf16x2")])
def test_spin_parity(ctx, oracle_mod, name, prec):
    import qbg
    idx, nel, ds, ne = SYSTEMS[name]
    s = cf.spin_polarized(cf.config(idx), nel, ds, ne)
    print(f"[case] {s.name} {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    o = oracle_mod.Oracle(s)
    c = spin_coeffs(s, o.basis(0)["g2"], 700 + idx)
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    stress = prec == "fp64"
    got = ef.evaluate(dc, hp, forces=True, stress=stress, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True, stress=stress)
    if prec == "fp64":
        for t in TERMS:
            assert abs(got[t] - ref[t]) <= 1e-10 * max(abs(ref[t]), 1e-12), (t, got[t], ref[t])
        assert relerr(got["fion"], ref["fion"]) < 1e-10
        assert relerr(hp.download(), ref["hpsi"]) < 1e-10
        scale = np.abs(ref["sigma"][0]).max()
        assert np.abs(got["sigma"] - ref["sigma"]).max() / scale < 1e-10
        for dev, orc in ((0, 0), (8, 14), (1, 1), (9, 15), (2, 2), (10, 16)):
            assert relerr(ef.get(dev), o.get(orc)) < 1e-10, (dev, orc)
        if name == "t5":
            assert relerr(ef.get(11), o.get(17)) < 1e-10
            assert relerr(ef.get(12), o.get(18)) < 1e-10
    else:
        for t in TERMS:
            assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
        assert np.abs(got["fion"] - ref["fion"]).max() < 1e-5
        assert relerr(hp.download(), ref["hpsi"]) < 1e-5
    ef.release()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spin_kept_transforms(ctx, prec):
    import qbg
    s = cf.spin_polarized(cf.config(11), 9, 1, 1)
    print(f"[case] spin kept transforms {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    c = spin_coeffs(s, ef.basis(0)["g2"], 801)
    dc = ctx.to_device(c)
    h1 = ctx.array(c.shape, np.complex128)
    h2 = ctx.array(c.shape, np.complex128)
    r1 = ef.evaluate(dc, h1, forces=True, prec=prec)
    ef.update_density(dc, prec)
    ef.update_vhxc(False)
    r2 = ef.energy(dc, h2, forces=True, prec=prec)
    assert np.array_equal(h1.download(), h2.download())
    assert r1["etotal"] == r2["etotal"]
    ef.release()


def test_spin_identical_sets_equal_nspin1(ctx):
    import qbg
    s1 = cf.config(12)
    print("[case] spin identical sets", flush=True)
    ef1 = qbg.EnergyFunctional(ctx, s1)
    c = cf.gram_numpy(cf.random_coefficients(802, ef1.basis(0)["g2"], s1.nst))
    dc1 = ctx.to_device(c)
    h1 = ctx.array(c.shape, np.complex128)
    r1 = ef1.evaluate(dc1, h1, forces=True)
    s2 = dataclasses.replace(s1, nst=2 * s1.nst, nspin=2, nst_up=s1.nst,
                             occ_bands=np.ones(2 * s1.nst))
    ef2 = qbg.EnergyFunctional(ctx, s2)
    cc = np.concatenate([c, c])
    dc2 = ctx.to_device(cc)
    h2 = ctx.array(cc.shape, np.complex128)
    r2 = ef2.evaluate(dc2, h2, forces=True)
    for t in TERMS:
        assert abs(r2[t] - r1[t]) <= 1e-11 * max(1.0, abs(r1[t])), t
    a1, a2 = h1.download(), h2.download()
    n = s1.nst
    assert np.abs(a2[:n] - a1).max() < 1e-11 and np.abs(a2[n:] - a1).max() < 1e-11
    # back to one spin: the same functional as ef1 on the doubled set with occ 1
    ef2.set_nspin(1)
    r3 = ef2.evaluate(dc2, h2, forces=True)
    assert abs(r3["etotal"] - r1["etotal"]) <= 1e-11 * abs(r1["etotal"])
    ef1.release()
    ef2.release()


def test_spin_refusals(ctx):
    """A stepper over all bands refuses nspin 2 (one per spin is needed), a
    mixer made for nspin 2 refuses a functional switched back to nspin 1,
    and nst_up beyond nst is refused."""
    import qbg
    s = cf.spin_polarized(cf.config(10), 5, 0, 1)
    ef = qbg.EnergyFunctional(ctx, s)
    with pytest.raises(qbg.QbgError) as e:
        qbg.WfStepper(ef, "psd", 0.0)
    assert e.value.code == -1
    mx = qbg.Mixer(ef, 0.5, 10.0, 3, 3.0)
    c = spin_coeffs(s, ef.basis(0)["g2"], 806)
    ef.set_nspin(1)
    ef.update_density(ctx.to_device(c), "fp64")
    with pytest.raises(qbg.QbgError):
        mx.update()
    mx.release()
    with pytest.raises(qbg.QbgError):
        ef.set_nspin(2, s.nst + 1)
    ef.release()


@pytest.mark.parametrize("graphs", [False, True])
def test_spin_scf_matches_oracle(ctx, oracle_mod, graphs):
    """nspin 2 SCF on the device (one stepper per spin, no charge mixing)
    against the oracle's loop, per iteration; graph replays included."""
    import qbg
    from qbg.scf import ScfLoop
    from test_oracle_energy import scf_spin
    s = cf.spin_polarized(cf.config(11), 9, 1, 1)
    print(f"[case] spin scf graphs={graphs}", flush=True)
    o = oracle_mod.Oracle(s)
    c = spin_coeffs(s, o.basis(0)["g2"], 950)
    co, es = scf_spin(oracle_mod, s, o, c, "psda", 0.0, 6)
    ef = qbg.EnergyFunctional(ctx, s)
    dc = ctx.to_device(c)
    dh = ctx.array(c.shape, np.complex128)
    loop = ScfLoop(ef, dc, dh, wf_dyn="psda", param=0.0, graphs=graphs)
    hist = loop.run(6, 0)
    ed = [h["etotal"] for h in hist]
    assert np.allclose(ed, es, rtol=1e-10, atol=0), (ed, es)
    assert relerr(dc.download(), co) < 1e-8
    loop.release()
    ef.release()


@pytest.mark.parametrize("graphs", [False, True])
def test_spin_scf_mixing_matches_oracle(ctx, oracle_mod, graphs):
    """nspin 2 SCF with charge mixing (both spins' densities in one Anderson
    vector, three steps per iteration at fixed potential) against the
    oracle's loop, per iteration; graph replays included."""
    import qbg
    from qbg.scf import ScfLoop
    from test_oracle_energy import scf_spin
    s = cf.spin_polarized(cf.config(11), 9, 1, 1)
    print(f"[case] spin scf mixing graphs={graphs}", flush=True)
    o = oracle_mod.Oracle(s)
    c = spin_coeffs(s, o.basis(0)["g2"], 951)
    co, es = scf_spin(oracle_mod, s, o, c, "psda", 0.0, 5, 3)
    ef = qbg.EnergyFunctional(ctx, s)
    dc = ctx.to_device(c)
    dh = ctx.array(c.shape, np.complex128)
    loop = ScfLoop(ef, dc, dh, wf_dyn="psda", param=0.0, graphs=graphs)
    hist = loop.run(5, 3)
    ed = [h["etotal"] for h in hist]
    assert np.allclose(ed, es, rtol=1e-9, atol=0), (ed, es)
    assert relerr(dc.download(), co) < 1e-7
    for dev, orc in ((0, 0), (8, 14)):
        assert relerr(ef.get(dev), o.get(orc)) < 1e-9, (dev, orc)
    loop.release()
    ef.release()

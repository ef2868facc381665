"""The class-level entry points of the boundary beyond the energy functional.

NonLocalPotential (qbg_nl_*, src/NonLocalPotential.h:61-68): enl, the
nonlocal forces and stress and H*psi_NL added into dsd, against the oracle's
NonLocalPotential::energy (src/NonLocalPotential.cpp:1258-2069), nspin 1 and
per spin of an nspin 2 system; update_twnl after a cell change.

EnergyFunctional::cell_moved (qbg_ef_cell_moved, src/EnergyFunctional.cpp:
1200-1245): a functional moved to a strained cell gives bit for bit what a
functional created on that cell (same reference cell) gives.
"""
import dataclasses

import numpy as np
import pytest

from qbg import configs as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


# H*psi_NL alone, relative to its own largest entry: the 2-part BF16 split of
# the projector products keeps about 16 bits (1.5e-5), so its bar is that of
# two roundings; on the whole H*psi (kinetic and local parts included) the
# mixed modes meet 1e-5 (tests/test_gpu_energy.py)
NL_HPSI_TOL = {"fp32": 1e-5, "bf16x2": 3e-5}


def check_mixed(s, enl, renl, f, rf, sig, rsig):
    """The metric's bar for the mixed modes: enl within 1e-6 Ha per atom,
    forces within 1e-5 Ha/bohr, and the stress within what an energy error
    of 1e-6 Ha per atom per unit strain gives, 1e-6 nat / omega (the 2-part
    BF16 split keeps about 16 bits, so a relative bar of 1e-5 on the small
    nonlocal stress alone would sit at its rounding level)."""
    omega = abs(np.linalg.det(s.cell))
    assert abs(enl - renl) / s.nat < 1e-6
    assert np.abs(f - rf).max() < 1e-5
    assert np.abs(sig - rsig).max() < 1e-6 * s.nat / omega


@pytest.mark.parametrize("idx", [11, 12, 1])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x2"])
def test_nl_energy_vs_oracle(ctx, oracle_mod, idx, prec):
    import qbg
    s = cf.config(idx)
    print(f"[case] {s.name} {prec}", flush=True)
    o = oracle_mod.Oracle(s)
    ef = qbg.EnergyFunctional(ctx, s)
    nl = qbg.NonLocalPotential(ef, 0)
    c = cf.gram_numpy(cf.random_coefficients(3000 + idx, o.basis(0)["g2"], s.nst))
    rng = np.random.default_rng(idx)
    h0 = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    dsd = ctx.to_device(h0)
    enl, f, sig = nl.energy(ctx.to_device(c), dsd, forces=True, stress=True, prec=prec)
    renl, rf, rsig, rh = o.nl_energy(c, np.zeros_like(c), forces=True, stress=True)
    hd = dsd.download() - h0
    if prec == "fp64":
        assert abs(enl - renl) <= 1e-10 * abs(renl)
        assert relerr(f, rf) < 1e-10
        assert relerr(sig, rsig) < 1e-10
        assert relerr(hd, rh) < 1e-10
    else:
        check_mixed(s, enl, renl, f, rf, sig, rsig)
        assert relerr(hd, rh) < NL_HPSI_TOL[prec]
    # without forces and stress: zeroed outputs, same energy
    enl2, f2, sig2 = nl.energy(ctx.to_device(c), None, forces=False, stress=False, prec=prec)
    assert enl2 == enl and not f2.any() and not sig2.any()


@pytest.mark.parametrize("prec", ["fp64", "bf16x2"])
def test_nl_energy_per_spin(ctx, oracle_mod, prec):
    import qbg
    s = cf.spin_polarized(cf.config(11), 16, 1, 1)
    ef = qbg.EnergyFunctional(ctx, s)
    c = cf.gram_numpy(cf.random_coefficients(77, ef.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    occ = s.occ
    for ispin, (b0, nb) in enumerate([(0, s.nst_up), (s.nst_up, s.nst - s.nst_up)]):
        print(f"[case] {s.name} spin {ispin} {prec}", flush=True)
        sub = dataclasses.replace(s, nst=nb, nspin=1, nst_up=0, occ_bands=occ[b0:b0 + nb])
        o = oracle_mod.Oracle(sub)
        nl = qbg.NonLocalPotential(ef, ispin)
        dsd = ctx.array((nb, ef.ngw), np.complex128)
        dsd.zero()
        enl, f, sig = nl.energy(dc.offset(16 * ef.ngw * b0), dsd, forces=True, stress=True,
                                prec=prec)
        renl, rf, rsig, rh = o.nl_energy(c[b0:b0 + nb], np.zeros((nb, ef.ngw), complex),
                                         forces=True, stress=True)
        if prec == "fp64":
            assert abs(enl - renl) <= 1e-10 * abs(renl)
            assert relerr(f, rf) < 1e-10 and relerr(sig, rsig) < 1e-10
            assert relerr(dsd.download(), rh) < 1e-10
        else:
            check_mixed(s, enl, renl, f, rf, sig, rsig)
            assert relerr(dsd.download(), rh) < NL_HPSI_TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "bf16x2"])
def test_cell_moved_equals_fresh_create(ctx, prec):
    import qbg
    s = cf.config(12)
    ref = s.cell * 1.05
    strain = np.array([[1.01, 0.002, 0.0], [0.002, 0.995, 0.001], [0.0, 0.001, 1.003]])
    cell2 = s.cell @ strain.T
    ef = qbg.EnergyFunctional(ctx, s, refcell=ref)
    c = cf.gram_numpy(cf.random_coefficients(55, ef.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    hp1 = ctx.array(c.shape, np.complex128)
    hp2 = ctx.array(c.shape, np.complex128)
    r0 = ef.evaluate(dc, hp1, forces=True, stress=True, prec=prec)
    ef.cell_moved(cell2)
    r1 = ef.evaluate(dc, hp1, forces=True, stress=True, prec=prec)
    assert r1["etotal"] != r0["etotal"]
    fresh = qbg.EnergyFunctional(ctx, dataclasses.replace(s, cell=cell2), refcell=ref)
    r2 = fresh.evaluate(dc, hp2, forces=True, stress=True, prec=prec)
    for k in r2:
        assert np.array_equal(np.asarray(r1[k]), np.asarray(r2[k])), k
    assert np.array_equal(hp1.download(), hp2.download())
    # the basis vectors follow the cell, the membership does not
    b1, b2 = ef.basis(0), fresh.basis(0)
    assert np.array_equal(b1["idx"], b2["idx"]) and np.array_equal(b1["gx"], b2["gx"])
    # back to the first cell: the first results again
    ef.cell_moved(s.cell)
    r3 = ef.evaluate(dc, hp1, forces=True, stress=True, prec=prec)
    assert np.array_equal(np.asarray(r3["sigma"]), np.asarray(r0["sigma"]))
    assert r3["etotal"] == r0["etotal"]
    # update_twnl on an unchanged cell leaves the nonlocal energy unchanged
    nl = qbg.NonLocalPotential(ef, 0)
    e1 = nl.energy(dc, None, prec=prec)[0]
    nl.update_twnl()
    assert nl.energy(dc, None, prec=prec)[0] == e1


def test_cell_moved_rejects_membership_change(ctx):
    import qbg
    s = cf.config(10)
    ef = qbg.EnergyFunctional(ctx, s)            # reference cell = the cell itself
    with pytest.raises(qbg.QbgError):
        ef.cell_moved(s.cell * 1.2)              # would need a larger basis

"""GPU parity of the whole evaluation (density -> update_vhxc -> energy with
H*psi and forces) and of SlaterDet::gram against the FP64 oracle.

Index maps built by the library's host setup are compared bit-exactly with
the oracle's; energies, H*psi, forces and intermediate fields agree to
1e-10 relative (arrays: max abs difference / max abs value)."""
import numpy as np
import pytest

from qbg import configs as cf

pytestmark = pytest.mark.gpu

TERMS = ["ekin", "eps", "enl", "ehart", "ehart_e", "ehart_ep", "ehart_p", "exc", "dxc", "esr",
         "eself", "ecoul", "etotal", "total_charge"]


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module", params=[10, 11, 12, 13, 15, 0, 1])
def system(request, ctx, oracle_mod):
    import qbg
    idx = request.param
    s = cf.config(idx)
    print(f"[case] config {idx}: {s.name}", flush=True)
    o = oracle_mod.Oracle(s)
    ef = qbg.EnergyFunctional(ctx, s)
    return idx, s, o, ef


def test_maps_bit_exact(system):
    idx, s, o, ef = system
    assert (ef.ngw, ef.ngv, ef.np0, ef.np1, ef.np2) == (o.ngw, o.ngv, o.np0, o.np1, o.np2)
    for which in (0, 1):
        a, b = ef.basis(which), o.basis(which)
        for k in ("idx", "ip", "im", "zvec_to_val"):
            assert np.array_equal(a[k], b[k]), (which, k)
        assert np.array_equal(a["g2"], b["g2"])
        assert np.array_equal(a["gx"], b["gx"])


def test_setup_tables(system):
    idx, s, o, ef = system
    # structure factor and local-potential tables (atoms_moved)
    assert relerr(ef.get(4), o.get(4)) < 1e-12     # vion_local_g
    assert relerr(ef.get(5), o.get(5)) < 1e-12     # rhopst
    assert relerr(ef.get(7), o.get(11)) < 1e-12    # fion_esr


def test_evaluate_fp64(system, ctx):
    idx, s, o, ef = system
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(1000 + idx, b["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True)
    ref = o.evaluate(c, hpsi=True, forces=True)
    for t in TERMS:
        err = abs(got[t] - ref[t]) / max(abs(ref[t]), 1e-12)
        assert err < 1e-10, (t, got[t], ref[t], err)
    assert relerr(ef.get(0), o.get(0)) < 1e-10          # rhor
    assert relerr(ef.get(3), o.get(3)) < 1e-10          # rhog
    assert relerr(ef.get(1), o.get(1)) < 1e-10          # v_r
    assert relerr(hp.download(), ref["hpsi"]) < 1e-10
    assert relerr(got["fion"], ref["fion"]) < 1e-10
    if any(sp.core_density() is not None for sp in s.species):
        assert relerr(ef.get(11), o.get(17)) < 1e-10     # core density on the grid
        assert relerr(ef.get(12), o.get(18)) < 1e-10     # vxc_g


# mixed precision (FP32 band FFTs; projector GEMMs on FP32 MFMA or as
# BF16-split products; FP64 reductions) against the FP64 oracle: every energy
# term within 1e-6 Ha/atom, forces within 1e-5 Ha/bohr, H*psi within 1e-5
# relative.  The 2-part split keeps about 16 bits per product; its rounding
# errors average out over the projector sums, and it meets the same bar.
MIXED_TOL = {"fp32": (1e-6, 1e-5, 1e-5), "bf16x3": (1e-6, 1e-5, 1e-5),
             "bf16x2": (1e-6, 1e-5, 1e-5)}


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x2"])
def test_evaluate_mixed(system, ctx, prec):
    idx, s, o, ef = system
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(1000 + idx, b["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, prec=prec)
    ref = o.evaluate(c, hpsi=True, forces=True)
    te, tf, th = MIXED_TOL[prec]
    for t in TERMS:
        err = abs(got[t] - ref[t]) / s.nat
        print(f"  {prec} {t:13s} {got[t]: .12e} dE/atom {err:.2e}")
        assert err < te, (t, got[t], ref[t], err)
    ferr = np.abs(got["fion"] - ref["fion"]).max()
    herr = relerr(hp.download(), ref["hpsi"])
    print(f"  {prec} fion max err {ferr:.2e}  hpsi rel err {herr:.2e}")
    assert ferr < tf
    assert herr < th


# stress: every row (total, ekin, eps, enl, ehart, exc, esr) to 1e-10 of the
# largest component of the oracle's total
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_stress(system, ctx, prec):
    idx, s, o, ef = system
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(2000 + idx, b["g2"], s.nst))
    dc = ctx.to_device(c)
    got = ef.evaluate(dc, None, forces=True, stress=True, prec=prec)
    ref = o.evaluate(c, hpsi=False, forces=True, stress=True)
    scale = np.abs(ref["sigma"][0]).max()
    err = np.abs(got["sigma"] - ref["sigma"]).max(axis=1) / scale
    print(f"  sigma rel err per row ({prec}):", " ".join(f"{x:.1e}" for x in err))
    tol = 1e-10 if prec == "fp64" else 1e-5
    assert err.max() < tol
    # stress without forces takes the same path
    got2 = ef.energy(dc, None, forces=False, stress=True, prec=prec)
    assert np.array_equal(got2["sigma"], got["sigma"])


def test_stress_strain_fd_device(ctx):
    """E(+u) - E(-u) = -2 omega tr(sigma u) (test/si2stress/check_stress.sh)
    on the device path, with a reference cell fixing the basis."""
    import dataclasses
    import qbg
    s = cf.config(11)
    ref = s.cell * 1.04
    ef0 = qbg.EnergyFunctional(ctx, s, refcell=ref)
    b = ef0.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(99, b["g2"], s.nst))
    dc = ctx.to_device(c)
    sig = ef0.evaluate(dc, None, forces=False, stress=True)["sigma"][0]
    v = 1e-4 * np.array([0.7, -0.4, 0.9, 0.3, -0.6, 0.2])
    u = np.zeros((3, 3))
    for k, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        u[i, j] = u[j, i] = v[k]
    e = []
    for sgn in (1.0, -1.0):
        F = np.eye(3) + sgn * u
        st = dataclasses.replace(s, cell=s.cell @ F.T, tau=s.tau @ F.T)
        ef = qbg.EnergyFunctional(ctx, st, refcell=ref)
        assert np.array_equal(ef.basis(0)["idx"], b["idx"])
        e.append(ef.evaluate(dc, None, forces=False, stress=False)["etotal"])
        ef.release()
    omega = abs(np.linalg.det(s.cell))
    pred = -2.0 * omega * (v[:3] @ sig[:3] + 2.0 * v[3:] @ sig[3:])
    de = e[0] - e[1]
    print(f"  fd {de:.10e} sigma {pred:.10e}")
    assert abs(de - pred) <= 2e-6 * abs(de) + 1e-11


def test_energy_state_errors(ctx):
    import qbg
    s = cf.config(10)
    ef = qbg.EnergyFunctional(ctx, s)
    c = ctx.array((s.nst, ef.ngw), np.complex128).zero()
    with pytest.raises(qbg.QbgError) as e:
        ef.energy(c)
    assert e.value.code == -5          # QBG_E_STATE: energy before update_vhxc
    ef.update_density(c)
    ef.update_vhxc(False)
    with pytest.raises(qbg.QbgError) as e:
        ef.energy(c, None, forces=False, stress=True)
    assert e.value.code == -5          # stress needs update_vhxc(compute_stress=1)


# 4, 5, 16 and 128 bands: one or many Cholesky panels, ragged inverse blocks
@pytest.mark.parametrize("idx", [10, 11, 0, 1])
def test_gram_fp64(ctx, oracle_mod, idx):
    import qbg
    s = cf.config(idx)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.random_coefficients(77 + idx, b["g2"], s.nst)
    ref = oracle_mod.gram(c)
    dc = ctx.to_device(c)
    qbg.gram(ctx, dc, o.ngw, o.ngw, s.nst)
    got = dc.download()
    assert relerr(got, ref) < 1e-10
    P = got.view(np.float64).reshape(s.nst, -1)
    S = 2 * P @ P.T - np.outer(P[:, 0], P[:, 0])
    assert np.abs(S - np.eye(s.nst)).max() < 1e-12


@pytest.mark.parametrize("nst", [129, 300])
def test_gram_blocked_cholesky(ctx, oracle_mod, nst):
    """More than two 64-column panels: the Cholesky runs panel by panel over
    the device (diagonal block, panel solve, trailing GEMM)."""
    import qbg
    import dataclasses
    s = dataclasses.replace(cf.config(1), nst=nst)
    o = oracle_mod.Oracle(cf.config(1))
    b = o.basis(0)
    c = cf.random_coefficients(88 + nst, b["g2"], nst)
    ref = oracle_mod.gram(c)
    dc = ctx.to_device(c)
    qbg.gram(ctx, dc, o.ngw, o.ngw, nst)
    got = dc.download()
    assert relerr(got, ref) < 1e-10
    P = got.view(np.float64).reshape(nst, -1)
    S = 2 * P @ P.T - np.outer(P[:, 0], P[:, 0])
    assert np.abs(S - np.eye(nst)).max() < 1e-12


@pytest.mark.parametrize("nst", [4, 200])
def test_gram_not_spd(ctx, nst):
    import qbg
    rng = np.random.default_rng(0)
    ngw = 183 if nst == 4 else 1500
    c = rng.standard_normal((nst, ngw)) + 1j * rng.standard_normal((nst, ngw))
    c[nst - 1] = 0.0                  # zero band: singular overlap, pivot exactly 0
    dc = ctx.to_device(c)
    with pytest.raises(qbg.QbgError) as e:
        qbg.gram(ctx, dc, ngw, ngw, nst)
    assert e.value.code == -4         # QBG_E_NOT_SPD
    got = dc.download()
    c[:, 0] = c[:, 0].real
    assert np.array_equal(got, c)     # coefficients untouched apart from cleanup


def test_mixed_256_tiles(ctx, oracle_mod):
    """256 bands and 256 projectors: the BF16-split products run on the
    256 x 256 tiles (projections, H*psi and the force product); the same
    bars as the 128-tile runs."""
    import dataclasses
    import qbg
    s = dataclasses.replace(cf.config(1), nst=256)
    print(f"[case] 256 tiles: {s.name} nst 256", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(4242, b["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    got = ef.evaluate(dc, hp, forces=True, prec="bf16x2")
    ref = o.evaluate(c, hpsi=True, forces=True)
    for t in TERMS:
        assert abs(got[t] - ref[t]) / s.nat < 1e-6, (t, got[t], ref[t])
    assert np.abs(got["fion"] - ref["fion"]).max() < 1e-5
    assert relerr(hp.download(), ref["hpsi"]) < 1e-5
    ef.release()

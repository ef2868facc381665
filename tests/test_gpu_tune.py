"""The launch tuner of the batched FFT passes (qbg_grid_tuning, qbg_grid_plan,
qbg_fft_splits), the counterpart of FourierTransform's FFTW_MEASURE planning
(src/FourierTransform.cpp:45-54, 830-1062).

The first batched call of a grid in a precision times the compiled splits of
np0 and np1 and the y pass's pairs per block and keeps a choice that the
library reports; grids of the same shape share it; a fixed plan overrides it
and is refused for splits that are not compiled.  The numbers the tuned
kernels produce are checked against the oracle by every other GPU test (the
tuner runs wherever a batched transform runs), and every candidate split by
tests/test_gpu_fft_lengths.py.
"""
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


# grid (90, 144, 224): both tuned axes have two compiled splits
K = 4


def grid_and_inputs(ctx, oracle_mod, nst=9):
    import qbg
    s, grid = cf.length_grid_system(K, nst)
    o = oracle_mod.Oracle(s, grid=grid)
    b = o.basis(0)
    g = qbg.Grid(ctx, o.np0, o.np1, o.np2, b["ip"], b["im"], b["zvec_to_val"])
    c = cf.random_coefficients(500 + K, b["g2"], s.nst)
    return s, o, g, ctx.to_device(c), c


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tuned_on_first_batched_call(ctx, oracle_mod, prec):
    import qbg
    from oracle import compute_density_occ
    s, o, g, dc, c = grid_and_inputs(ctx, oracle_mod)
    print(f"[case] tune {s.name} {prec}", flush=True)
    assert not g.tuning(prec)["tuned"]
    rho = ctx.array(o.N).zero()
    omega = abs(np.linalg.det(s.cell))
    g.compute_density(dc, o.ngw, s.nst, s.occ, 1.0, omega, rho, prec)
    t = g.tuning(prec)
    print(f"  choice {t}", flush=True)
    assert t["tuned"]
    assert t["x"][0] in qbg.fft_splits(o.np0) and t["x"][0] * t["x"][1] == o.np0
    assert t["y"][0] in qbg.fft_splits(o.np1) and t["y"][0] * t["y"][1] == o.np1
    assert t["yppb"] in (8, 16, 32)
    # the density with the tuned kernels is the oracle's
    ref = compute_density_occ(o, c, s.occ)
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert np.abs(rho.download() - ref).max() / np.abs(ref).max() < tol
    # a second grid of the same shape takes the cached choice
    b = o.basis(0)
    g2 = qbg.Grid(ctx, o.np0, o.np1, o.np2, b["ip"], b["im"], b["zvec_to_val"])
    rho2 = ctx.array(o.N).zero()
    g2.compute_density(dc, o.ngw, s.nst, s.occ, 1.0, omega, rho2, prec)
    assert g2.tuning(prec) == t
    assert np.array_equal(rho2.download(), rho.download())
    g2.release()
    g.release()


def test_plan_fixes_and_validates(ctx, oracle_mod):
    import qbg
    s, o, g, dc, c = grid_and_inputs(ctx, oracle_mod)
    splits = qbg.fft_splits(o.np0)
    g.plan("fp64", splits[-1], 0, 8)
    t = g.tuning("fp64")
    assert t["tuned"] and t["x"][0] == splits[-1] and t["yppb"] == 8
    with pytest.raises(qbg.QbgError) as e:
        g.plan("fp64", 5, 0, 0)              # no 5 x n2 split of this length is compiled
    assert e.value.code == -1
    # a batched call keeps the fixed plan
    rho = ctx.array(o.N).zero()
    g.compute_density(dc, o.ngw, s.nst, s.occ, 1.0, abs(np.linalg.det(s.cell)), rho, "fp64")
    assert g.tuning("fp64") == t
    g.release()


def test_gemm_plan_tuned_at_first_bf16_use(ctx):
    """The projection product's plan (tile edge, reduction splits) is timed at
    the first 2-part BF16 evaluation and kept; the tuned evaluation meets the
    metric's bar against the FP64 evaluation of the same inputs."""
    import dataclasses
    import qbg
    s = dataclasses.replace(cf.config(1), nst=256)
    print(f"[case] gemm plan {s.name} nst 256", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    assert not ef.gemm_plan()["tuned"]
    c = cf.gram_numpy(cf.random_coefficients(610, ef.basis(0)["g2"], s.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    r64 = ef.evaluate(dc, hp, forces=True, prec="fp64")
    h64 = hp.download()
    rbf = ef.evaluate(dc, hp, forces=True, prec="bf16x2")
    p = ef.gemm_plan()
    print(f"  plan {p}", flush=True)
    assert p["tuned"] and p["edge"] in (128, 256) and p["split"] >= 0
    for t in ("ekin", "eps", "enl", "ehart", "exc", "etotal"):
        assert abs(rbf[t] - r64[t]) / s.nat < 1e-6, t
    assert np.abs(rbf["fion"] - r64["fion"]).max() < 1e-5
    assert np.abs(hp.download() - h64).max() / np.abs(h64).max() < 1e-5
    ef.release()

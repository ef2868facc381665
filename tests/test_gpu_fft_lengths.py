"""GPU parity of every compiled line length on every axis against the oracle.

The four-step / prime-factor kernels are instantiated per length
(csrc/fft4_*.hip): 22, 24, 28, 64, 90, 96, 112, 126, 128, 144, 160, 168, 198,
210 and 224.  Fifteen grids (L[k], L[k+5], L[k+10]) put each length on each of
the three axes once.  The oracle runs the same basis on the same forced grid
(FourierTransform(basis, np0, np1, np2), src/FourierTransform.h:105).

Per grid and precision: single backward/forward on the wavefunction and the
density basis, pair backward/forward (src/FourierTransform.cpp:147-827),
SlaterDet::compute_density and rs_mul_add (src/SlaterDet.cpp:225-298,
410-477) with an odd band count.  FP64 to 1e-12 relative, FP32 to 2e-5.

Lengths with more than one compiled split (the candidates of the launch
tuner) run the pair transforms, the density and rs_mul_add again with every
split fixed on the x axis (np0) and the y axis (np1) through qbg_grid_plan,
and with 8 and 32 band pairs per y block.
"""
import numpy as np
import pytest

from qbg import configs as cf

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5}


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rand_c(rng, n):
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c[0] = c[0].real
    return c


@pytest.mark.parametrize("k", range(len(cf.FFT_LENGTHS)))
def test_every_length_on_every_axis(ctx, oracle_mod, k):
    import qbg
    from oracle import compute_density_occ
    s, grid = cf.length_grid_system(k)
    o = oracle_mod.Oracle(s, grid=grid)
    assert (o.np0, o.np1, o.np2) == grid
    print(f"[case] {s.name} ngw {o.ngw} ngv {o.ngv} ntrans0 {o.ntrans0}", flush=True)
    b, vb = o.basis(0), o.basis(1)
    g = qbg.Grid(ctx, o.np0, o.np1, o.np2, b["ip"], b["im"], b["zvec_to_val"])
    gv = qbg.Grid(ctx, o.np0, o.np1, o.np2, vb["ip"], vb["im"], vb["zvec_to_val"])
    assert g.four_step, "the compiled kernels must serve this grid"
    rng = np.random.default_rng(300 + k)
    # inputs and oracle results, shared by both precisions
    cases = []
    for grd, which, n in ((g, 0, o.ngw), (gv, 1, o.ngv)):
        c = rand_c(rng, n)
        x = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
        cases.append((grd, n, c, o.fft_backward(c, which), x, o.fft_forward(x, which)))
    c1, c2 = rand_c(rng, o.ngw), rand_c(rng, o.ngw)
    fpair = o.fft_backward_pair(c1, c2)
    xp = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    r1, r2 = o.fft_forward_pair(xp)
    nst = s.nst
    cb = cf.random_coefficients(400 + k, b["g2"], nst)
    occ = np.linspace(2.0, 0.5, nst)
    omega = abs(np.linalg.det(s.cell))
    rho_ref = compute_density_occ(o, cb, occ)
    v = rng.standard_normal(o.N)
    h0 = rng.standard_normal(cb.shape) + 1j * rng.standard_normal(cb.shape)
    href = o.rs_mul_add(v, cb, h0) - h0
    dcb, dv = ctx.to_device(cb), ctx.to_device(v)
    for prec, tol in TOL.items():
        print(f"[case] {s.name} {prec}", flush=True)
        for grd, n, c, fref, x, cref in cases:
            f = ctx.array(o.N, np.complex128)
            grd.fft_backward(ctx.to_device(c), f, prec)
            assert relerr(f.download(), fref) < tol, (prec, "backward", n)
            dc = ctx.array(n, np.complex128)
            grd.fft_forward(ctx.to_device(x), dc, prec)
            assert relerr(dc.download(), cref) < tol, (prec, "forward", n)
        both = ctx.to_device(np.stack([c1, c2]))
        f = ctx.array(o.N, np.complex128)
        g.fft_backward_pair(both.p, both.offset(16 * o.ngw), f, prec)
        assert relerr(f.download(), fpair) < tol, (prec, "backward pair")
        out = ctx.array((2, o.ngw), np.complex128)
        g.fft_forward_pair(ctx.to_device(xp), out.p, out.offset(16 * o.ngw), prec)
        got = out.download()
        assert relerr(got[0], r1) < tol and relerr(got[1], r2) < tol, (prec, "forward pair")
        rho = ctx.array(o.N).zero()
        g.compute_density(dcb, o.ngw, nst, occ, 1.0, omega, rho, prec)
        assert relerr(rho.download(), rho_ref) < tol, (prec, "density")
        hp = ctx.to_device(h0)
        g.rs_mul_add(dv, dcb, o.ngw, nst, hp, o.ngw, prec)
        assert relerr(hp.download() - h0, href) < tol, (prec, "rs_mul_add")
        # every other compiled split of np0 (x pass) and np1 (y pass), fixed
        xs, ys = qbg.fft_splits(o.np0), qbg.fft_splits(o.np1)
        plans = [(xn1, 0, 0) for xn1 in xs[1:]] + [(0, yn1, 0) for yn1 in ys[1:]]
        plans += [(0, 0, 8), (xs[-1], ys[-1], 32)]
        for xn1, yn1, ppb in plans:
            print(f"[case] {s.name} {prec} plan x {xn1} y {yn1} ppb {ppb}", flush=True)
            g.plan(prec, xn1, yn1, ppb)
            t = g.tuning(prec)
            assert t["tuned"] and (xn1 == 0 or t["x"][0] == xn1) and (yn1 == 0 or t["y"][0] == yn1)
            f = ctx.array(o.N, np.complex128)
            g.fft_backward_pair(both.p, both.offset(16 * o.ngw), f, prec)
            assert relerr(f.download(), fpair) < tol, (prec, "backward pair", xn1, yn1)
            g.fft_forward_pair(ctx.to_device(xp), out.p, out.offset(16 * o.ngw), prec)
            got = out.download()
            assert relerr(got[0], r1) < tol and relerr(got[1], r2) < tol, (prec, "fwd pair", xn1, yn1)
            rho = ctx.array(o.N).zero()
            g.compute_density(dcb, o.ngw, nst, occ, 1.0, omega, rho, prec)
            assert relerr(rho.download(), rho_ref) < tol, (prec, "density", xn1, yn1, ppb)
            hp = ctx.to_device(h0)
            g.rs_mul_add(dv, dcb, o.ngw, nst, hp, o.ngw, prec)
            assert relerr(hp.download() - h0, href) < tol, (prec, "rs_mul_add", xn1, yn1, ppb)

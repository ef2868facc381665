"""GPU parity of the FourierTransform / SlaterDet transforms against the oracle.

Bit-exactness of the index maps holds by construction (the device uses the
maps it is given); these tests check the transformed values through the C ABI
on the same seeded inputs: FP64 to 1e-12 relative, FP32 to 2e-5.
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


def _setup(ctx, oracle_mod, idx):
    import qbg
    s = cf.config(idx)
    o = oracle_mod.Oracle(s)
    print(f"[case] config {idx} {s.name} grid {o.np0}x{o.np1}x{o.np2}", flush=True)
    b = o.basis(0)
    g = qbg.Grid(ctx, o.np0, o.np1, o.np2, b["ip"], b["im"], b["zvec_to_val"])
    vb = o.basis(1)
    gv = qbg.Grid(ctx, o.np0, o.np1, o.np2, vb["ip"], vb["im"], vb["zvec_to_val"])
    return s, o, b, g, gv


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rand_c(rng, n):
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c[0] = c[0].real
    return c


@pytest.mark.parametrize("idx", [10, 11, 12, 0])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_single_and_pair_transforms(ctx, oracle_mod, idx, prec):
    s, o, b, g, gv = _setup(ctx, oracle_mod, idx)
    rng = np.random.default_rng(idx)
    for grid, which, n in ((g, 0, o.ngw), (gv, 1, o.ngv)):
        c = rand_c(rng, n)
        f = ctx.array(o.N, np.complex128)
        grid.fft_backward(ctx.to_device(c), f, prec)
        ref = o.fft_backward(c, which)
        assert relerr(f.download(), ref) < TOL[prec]
        # forward of an arbitrary complex grid function
        x = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
        dx = ctx.to_device(x)
        dc = ctx.array(n, np.complex128)
        grid.fft_forward(dx, dc, prec)
        assert relerr(dc.download(), o.fft_forward(x, which)) < TOL[prec]
    c1, c2 = rand_c(rng, o.ngw), rand_c(rng, o.ngw)
    both = ctx.to_device(np.stack([c1, c2]))
    f = ctx.array(o.N, np.complex128)
    g.fft_backward_pair(both.p, both.offset(16 * o.ngw), f, prec)
    assert relerr(f.download(), o.fft_backward_pair(c1, c2)) < TOL[prec]
    x = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    out = ctx.array((2, o.ngw), np.complex128)
    g.fft_forward_pair(ctx.to_device(x), out.p, out.offset(16 * o.ngw), prec)
    r1, r2 = o.fft_forward_pair(x)
    got = out.download()
    assert relerr(got[0], r1) < TOL[prec] and relerr(got[1], r2) < TOL[prec]


@pytest.mark.parametrize("idx,nst", [(10, 4), (11, 5), (12, 6), (0, 16), (0, 15)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_density_and_rs_mul_add(ctx, oracle_mod, idx, nst, prec):
    s, o, b, g, gv = _setup(ctx, oracle_mod, idx)
    rng = np.random.default_rng(100 + idx)
    c = cf.random_coefficients(5 + idx, b["g2"], nst)
    occ = np.linspace(2.0, 0.5, nst)
    omega = abs(np.linalg.det(s.cell))
    dc = ctx.to_device(c)
    rho0 = rng.random(o.N)
    rho = ctx.to_device(rho0)                       # density is accumulated (+=)
    g.compute_density(dc, o.ngw, nst, occ, 1.0, omega, rho, prec)
    # oracle with the same occupations
    from oracle import compute_density_occ
    ref = rho0 + compute_density_occ(o, c, occ)
    assert relerr(rho.download() - rho0, ref - rho0) < TOL[prec]
    v = rng.standard_normal(o.N)
    h0 = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    hp = ctx.to_device(h0)
    g.rs_mul_add(ctx.to_device(v), dc, o.ngw, nst, hp, o.ngw, prec)
    href = o.rs_mul_add(v, c, h0)
    assert relerr(hp.download() - h0, href - h0) < TOL[prec]


def test_bad_maps_rejected(ctx, oracle_mod):
    import qbg
    s = cf.config(10)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    ip = b["ip"].copy()
    ip[5] = ip[6]                                   # duplicate slot
    with pytest.raises(qbg.QbgError) as e:
        qbg.Grid(ctx, o.np0, o.np1, o.np2, ip, b["im"], b["zvec_to_val"])
    assert e.value.code == -1
    with pytest.raises(qbg.QbgError):
        qbg.Grid(ctx, o.np0, o.np1, 13, b["ip"], b["im"], b["zvec_to_val"])   # 13 not a radix

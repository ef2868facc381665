"""qbg_gram_noalloc (SlaterDet::gram with its temporaries in the library's
fixed per-device array) against qbg_gram, through the C ABI: two calls in a
row, the second on a smaller input, both equal to qbg_gram's result to FP64
rounding and orthonormal in the Gamma convention; a size past the fixed
maximum is refused; a singular overlap is reported as by qbg_gram and leaves
the array usable."""
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


def gamma_overlap(c):
    P = c.view(np.float64).reshape(c.shape[0], -1)
    return 2 * P @ P.T - np.outer(P[:, 0], P[:, 0])


def coefficients(seed, ngw, nst):
    rng = np.random.default_rng(seed)
    g2 = np.sort(rng.uniform(0.0, 40.0, ngw))
    g2[0] = 0.0
    return cf.random_coefficients(seed, g2, nst)


def check_pair(ctx, c):
    import qbg
    nst, ngw = c.shape
    print(f"[case] gram_noalloc nst {nst} ngw {ngw}", flush=True)
    d_ref, d_na = ctx.to_device(c), ctx.to_device(c)
    qbg.gram(ctx, d_ref, ngw, ngw, nst)
    qbg.gram_noalloc(ctx, d_na, ngw, ngw, nst)
    ref, got = d_ref.download(), d_na.download()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err
    assert np.abs(gamma_overlap(got) - np.eye(nst)).max() < 1e-12


def test_two_calls_second_smaller(ctx):
    import qbg
    lib = ctx.lib
    assert lib.qbg_gram_noalloc_bytes() >= 8 * (2 + 6) * 1024 * 1024
    # three column blocks of the in-place update (300 = 128 + 128 + 44), then
    # a smaller system in the same array
    check_pair(ctx, coefficients(11, 5682, 300))
    check_pair(ctx, coefficients(12, 1500, 40))
    # past the maximum: refused before any work
    c = coefficients(13, 64, 8)
    d = ctx.to_device(c)
    rc = lib.qbg_gram_noalloc(ctx.h, d.p, 64, 64, 1025, 0)
    assert rc == qbg._lib.QBG_E_INVAL


def test_not_spd_then_usable(ctx):
    import qbg
    rng = np.random.default_rng(0)
    nst, ngw = 200, 1500
    c = rng.standard_normal((nst, ngw)) + 1j * rng.standard_normal((nst, ngw))
    c[nst - 1] = 0.0
    d = ctx.to_device(c)
    with pytest.raises(qbg.QbgError) as e:
        qbg.gram_noalloc(ctx, d, ngw, ngw, nst)
    assert e.value.code == qbg._lib.QBG_E_NOT_SPD
    c[:, 0] = c[:, 0].real
    assert np.array_equal(d.download(), c)
    check_pair(ctx, coefficients(14, 900, 130))

"""Pins the oracle's Basis / BasisMapping / FourierTransform restatement.

Checks come from outside the restatement: brute-force G-sphere enumeration,
the rod-ordering rule of src/Basis.cpp:355,593-597 re-derived here, numpy's
FFT on the full grid, Parseval, the closed-form Gaussian of
src/testFourierTransform.cpp:157-167, and the sizes SURVEY-derived for the
five workloads.
"""
import math

import numpy as np
import pytest

from qbg import configs as cf


@pytest.fixture(scope="module", params=[10, 11, 12, 0])
def sysinfo(request, oracle_mod):
    s = cf.config(request.param)
    return s, oracle_mod.Oracle(s)


def brute_half_sphere(b, ecut_ha):
    """All integer (h,k,l) with |G|^2 < 2 ecut in Qbox's Gamma half space."""
    two_ecut = 2.0 * ecut_ha
    bmax = math.sqrt(two_ecut) / min(np.linalg.norm(b, axis=1)) + 3
    n = int(bmax) + 2
    out = set()
    for h in range(0, n + 1):
        for k in range(-n, n + 1):
            for l in range(-n, n + 1):
                if h == 0 and (k < 0 or (k == 0 and l < 0)):
                    continue
                g = h * b[0] + k * b[1] + l * b[2]
                if g @ g < two_ecut:
                    out.add((h, k, l))
    return out


def test_sizes_match_workload_table(oracle_mod):
    # sizes of the c1/c2 workloads (rules of src/Basis.cpp:307-520 and
    # src/ChargeDensity.cpp:44-49 evaluated independently)
    exp = {0: (5682, 609, 14, (64, 64, 64), 45524),
           1: (54336, 2749, 30, (126, 126, 126), 434108)}
    for idx, (ngw, nvec, nt0, grid, ngv) in exp.items():
        o = oracle_mod.Oracle(cf.config(idx))
        assert (o.ngw, o.nvec, o.ntrans0, (o.np0, o.np1, o.np2), o.ngv) == (ngw, nvec, nt0, grid, ngv)


@pytest.mark.slow
def test_sizes_large_workloads(oracle_mod):
    exp = {2: (230668, 7209, 48, (210, 210, 210), 1845428),
           4: (105502, 4009, 37, (144, 160, 168), 843843)}
    for idx, (ngw, nvec, nt0, grid, ngv) in exp.items():
        o = oracle_mod.Oracle(cf.config(idx))
        assert (o.ngw, o.nvec, o.ntrans0, (o.np0, o.np1, o.np2), o.ngv) == (ngw, nvec, nt0, grid, ngv)


def test_basis_is_brute_force_half_sphere(sysinfo, oracle_mod):
    s, o = sysinfo
    b, _ = oracle_mod.cell(s.cell)
    bas = o.basis(0)
    got = set(map(tuple, bas["idx"].tolist()))
    assert len(got) == o.ngw
    assert got == brute_half_sphere(b, s.ecut)
    vb = o.basis(1)
    assert set(map(tuple, vb["idx"].tolist())) == brute_half_sphere(b, 4 * s.ecut)
    # g2 consistent with idx
    G = bas["idx"] @ b
    np.testing.assert_allclose((G * G).sum(1), bas["g2"], rtol=1e-14, atol=1e-14)


def test_rod_order_rule(sysinfo):
    """Rods sorted by size with ties in insertion order ((0,0), (0,k>0), then
    (h>0, k ascending)), then rod (0,0) swapped into slot 0."""
    _, o = sysinfo
    rh, rk, rl, rs = o.rods(0)
    rods = list(zip(rh.tolist(), rk.tolist(), rl.tolist(), rs.tolist()))
    ins = sorted(rods, key=lambda r: (0 if r[0] == 0 and r[1] == 0 else 1 if r[0] == 0 else 2,
                                      r[0], r[1]))
    order = sorted(range(len(ins)), key=lambda i: (ins[i][3], i))   # stable by size
    exp = [ins[i] for i in order]
    j = [i for i, r in enumerate(exp) if r[0] == 0 and r[1] == 0][0]
    exp[0], exp[j] = exp[j], exp[0]
    assert rods == exp


def test_mapping_invariants(sysinfo):
    _, o = sysinfo
    for which in (0, 1):
        b = o.basis(which)
        nvec = o.nvec if which == 0 else o.nvec_v
        ip, im, z2v = b["ip"], b["im"], b["zvec_to_val"]
        assert ip[0] == 0 and im[0] == 0
        slots = np.concatenate([ip, im[1:]])
        assert len(np.unique(slots)) == slots.size
        assert slots.min() >= 0 and slots.max() < nvec * o.np2
        assert len(np.unique(z2v)) == nvec
        assert z2v.min() >= 0 and z2v.max() < o.np0 * o.np1
        # column of +G and -G hold (h,k) and (-h,-k) modulo the grid
        idx = b["idx"]
        col_p, col_m = ip // o.np2, im // o.np2
        xy = lambda h, k: (h % o.np0) + o.np0 * (k % o.np1)
        assert np.all(z2v[col_p] == xy(idx[:, 0], idx[:, 1]))
        assert np.all(z2v[col_m] == xy(-idx[:, 0], -idx[:, 1]))
        assert np.all(ip % o.np2 == idx[:, 2] % o.np2)
        assert np.all(im % o.np2 == (-idx[:, 2]) % o.np2)


def grid_from_coeffs(o, idx, c, pair_c2=None):
    F = np.zeros((o.np2, o.np1, o.np0), dtype=complex)
    h, k, l = idx[:, 0], idx[:, 1], idx[:, 2]
    if pair_c2 is None:
        F[(-l) % o.np2, (-k) % o.np1, (-h) % o.np0] = np.conj(c)
        F[l % o.np2, k % o.np1, h % o.np0] = c
    else:
        z = c + 1j * pair_c2
        zm = np.conj(c) + 1j * np.conj(pair_c2)
        F[(-l) % o.np2, (-k) % o.np1, (-h) % o.np0] = zm
        F[l % o.np2, k % o.np1, h % o.np0] = z
    return F


def rand_c(rng, n):
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c[0] = c[0].real
    return c


def test_backward_equals_numpy_ifftn(sysinfo):
    _, o = sysinfo
    rng = np.random.default_rng(1)
    b = o.basis(0)
    c = rand_c(rng, o.ngw)
    f = o.fft_backward(c)
    F = grid_from_coeffs(o, b["idx"], c)
    ref = np.fft.ifftn(F) * F.size          # synthesis e^{+iGr}, unscaled
    np.testing.assert_allclose(f, ref.ravel(), rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.abs(f.imag).max() < 1e-12 * np.abs(f).max()   # real function


def test_backward_pair_packs_two_real_functions(sysinfo):
    _, o = sysinfo
    rng = np.random.default_rng(2)
    b = o.basis(0)
    c1, c2 = rand_c(rng, o.ngw), rand_c(rng, o.ngw)
    f = o.fft_backward_pair(c1, c2)
    np.testing.assert_allclose(f.real, o.fft_backward(c1).real, atol=1e-11)
    np.testing.assert_allclose(f.imag, o.fft_backward(c2).real, atol=1e-11)


def test_forward_equals_numpy_fftn_and_roundtrips(sysinfo):
    _, o = sysinfo
    rng = np.random.default_rng(3)
    b = o.basis(0)
    c1, c2 = rand_c(rng, o.ngw), rand_c(rng, o.ngw)
    f = o.fft_backward_pair(c1, c2)
    d1, d2 = o.fft_forward_pair(f.copy())
    np.testing.assert_allclose(d1, c1, atol=1e-13)
    np.testing.assert_allclose(d2, c2, atol=1e-13)
    # arbitrary real-space function: forward = fftn/N sampled at the basis
    g = rng.standard_normal(o.N) + 0j
    c = o.fft_forward(g.copy())
    Fg = np.fft.fftn(g.reshape(o.np2, o.np1, o.np0)) / o.N
    h, k, l = b["idx"].T
    np.testing.assert_allclose(c, Fg[l % o.np2, k % o.np1, h % o.np0], atol=1e-14)


def test_gaussian_closed_form(oracle_mod):
    """exp(-G^2 rc^2/4)/Omega synthesizes (pi^-3/2 / rc^3) sum_R exp(-|r-R|^2/rc^2)."""
    s = cf.config(0)
    o = oracle_mod.Oracle(s)
    vb = o.basis(1)
    rc = 1.0
    omega = 16.0 ** 3
    c = np.exp(-vb["g2"] * rc * rc / 4) / omega
    f = o.fft_backward(c, which=1)
    n = o.np0
    x = np.arange(n) * 16.0 / n
    d = np.minimum(x, 16.0 - x)
    r2 = d[None, None, :] ** 2 + d[None, :, None] ** 2 + d[:, None, None] ** 2
    exact = np.exp(-r2 / rc ** 2) / (math.pi ** 1.5 * rc ** 3)
    np.testing.assert_allclose(f.real.reshape(n, n, n), exact, atol=1e-10)


def test_parseval(sysinfo):
    _, o = sysinfo
    rng = np.random.default_rng(4)
    c = rand_c(rng, o.ngw)
    f = o.fft_backward(c)
    lhs = (np.abs(f) ** 2).sum() / o.N
    rhs = 2 * (np.abs(c) ** 2).sum() - abs(c[0]) ** 2
    assert abs(lhs - rhs) < 1e-12 * rhs


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 18, 20, 21,
                               22, 24, 28, 30, 33, 35, 36, 40, 42, 44, 45, 48, 56, 60, 63,
                               64, 70, 72, 77, 80, 84, 90, 96, 99, 126, 144, 160, 168, 210, 252])
def test_fft1d_vs_numpy(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(oracle_mod.fft1d(x, -1), np.fft.fft(x), atol=1e-12 * n)
    np.testing.assert_allclose(oracle_mod.fft1d(x, +1), np.fft.ifft(x) * n, atol=1e-12 * n)

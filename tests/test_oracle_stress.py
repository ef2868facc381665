"""The oracle's stress tensor against central finite differences of the
energy under strain, term by term (ekin, eps, enl, ehart, exc, esr), with the
method of test/si2stress/check_stress.sh:20-90:

    E(+u) - E(-u) = -2 omega (sum_i u_ii s_ii + 2 sum_{i<j} u_ij s_ij)

at fixed plane-wave coefficients and fixed fractional atomic positions.  With
core densities the reference puts the strain derivative of the core density's
shape into sigma_ehart (src/EnergyFunctional.cpp:650-659) while exc carries
it, so there exc and ehart are checked as one term.  A
reference cell fixes the basis membership and the grids
(src/Basis.cpp:314-325), so the strained systems share one basis and the
coefficients carry over index by index."""
import dataclasses

import numpy as np
import pytest

from qbg import configs as cf

# energy term -> row of the 7 x 6 sigma block (total, ekin, eps, enl, ehart, exc, esr)
PAIRS = {"ekin": 1, "eps": 2, "enl": 3, "ehart": 4, "exc": 5, "esr": 6, "etotal": 0}
ENERGY_TERMS = set(PAIRS)
VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]


def strained(sysdef, u):
    """Cell rows a_i -> (1+u) a_i and positions tau -> (1+u) tau."""
    F = np.eye(3) + u
    return dataclasses.replace(sysdef, cell=sysdef.cell @ F.T, tau=sysdef.tau @ F.T)


def sym(v):
    u = np.zeros((3, 3))
    for k, (i, j) in enumerate(VOIGT):
        u[i, j] = u[j, i] = v[k]
    return u


@pytest.mark.parametrize("idx", [11, 12, 13, 14, 15])
def test_stress_vs_strain_fd(oracle_mod, idx):
    s = cf.config(idx)
    ref = s.cell * 1.04                   # holds every strained sphere
    o0 = oracle_mod.Oracle(s, refcell=ref)
    b0 = o0.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(500 + idx, b0["g2"], s.nst))
    r0 = o0.evaluate(c, hpsi=False, forces=False, stress=True)
    sigma = r0["sigma"]
    # the total is the sum of the parts
    assert np.allclose(sigma[0], sigma[1:].sum(axis=0), rtol=0, atol=1e-14)
    rng = np.random.default_rng(idx)

    def diff(u):
        """E(+u) - E(-u) per energy term."""
        rp = oracle_mod.Oracle(strained(s, u), refcell=ref)
        rm = oracle_mod.Oracle(strained(s, -u), refcell=ref)
        assert np.array_equal(rp.basis(0)["idx"], b0["idx"])
        assert np.array_equal(rm.basis(1)["idx"], o0.basis(1)["idx"])
        ep = rp.evaluate(c, hpsi=False, forces=False, stress=False)
        em = rm.evaluate(c, hpsi=False, forces=False, stress=False)
        return {t: ep[t] - em[t] for t in ep if t in ENERGY_TERMS}

    for trial in range(2):
        v = 1e-4 * rng.uniform(-1.0, 1.0, 6)
        u = sym(v)
        # Richardson: the cubic term of the central difference cancels in
        # (8 D(u/2) - D(u)) / 3, which leaves a relative error near 1e-9
        d1, d2 = diff(u), diff(0.5 * u)
        ep = {t: (8.0 * d2[t] - d1[t]) / 3.0 for t in d1}
        em = {t: 0.0 for t in d1}
        omega = abs(np.linalg.det(s.cell))
        core = any(sp.core_density() is not None for sp in s.species)
        pairs = dict(PAIRS)
        if core:
            del pairs["exc"], pairs["ehart"]
            pairs["exc+ehart"] = (4, 5)
        for term, row in pairs.items():
            if isinstance(row, tuple):
                de = sum(ep[t] - em[t] for t in term.split("+"))
                st = sigma[row[0]] + sigma[row[1]]
            else:
                de = ep[term] - em[term]
                st = sigma[row]
            pred = -2.0 * omega * (sum(v[k] * st[k] for k in range(3)) +
                                   2.0 * sum(v[k] * st[k] for k in range(3, 6)))
            print(f"{term:7s} fd {de: .10e} sigma {pred: .10e} diff {de - pred: .2e}")
            assert abs(de - pred) <= 2e-7 * abs(de) + 1e-12, (term, de, pred)


def test_stress_axis_permutation(oracle_mod):
    """Relabelling the axes of a cubic cell (x, y, z) -> (y, z, x), with the
    positions and the coefficients relabelled the same way, permutes the
    stress tensor the same way."""
    s = cf.config(10)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(7, b["g2"], s.nst))
    sig = o.evaluate(c, hpsi=False, forces=False, stress=True)["sigma"][0]
    # permute axes (x, y, z) -> (y, z, x) of positions; cubic cell unchanged
    s2 = dataclasses.replace(s, tau=s.tau[:, [1, 2, 0]])
    o2 = oracle_mod.Oracle(s2)
    b2 = o2.basis(0)
    # coefficients follow the permutation of Miller indices
    key = {tuple(v): i for i, v in enumerate(b["idx"])}
    perm = []
    for v in b2["idx"]:
        w = (v[2], v[0], v[1])            # G' = P G
        if w in key:
            perm.append((key[w], False))
        else:
            perm.append((key[(-w[0], -w[1], -w[2])], True))
    c2 = np.empty_like(c)
    for i, (j, conj) in enumerate(perm):
        c2[:, i] = np.conj(c[:, j]) if conj else c[:, j]
    sig2 = o2.evaluate(c2, hpsi=False, forces=False, stress=True)["sigma"][0]
    # sigma'_{ij} = sigma_{p(i) p(j)} with x'=y, y'=z, z'=x
    full = lambda v: np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]])
    S, S2 = full(sig), full(sig2)
    P = [1, 2, 0]
    assert np.allclose(S2, S[np.ix_(P, P)], rtol=1e-9, atol=1e-12)

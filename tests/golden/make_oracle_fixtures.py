"""Write the CPU oracle's results for the full-size workloads as fixtures.

The 256-atom / 512-band mixed-precision workload (c3) and the 128-atom
orthorhombic workload with stress (c5) take the oracle minutes, too long for
a GPU test to recompute, so its results on the workloads' own seeded inputs
are committed here and the GPU tests (tests/test_gpu_fullsize.py) and the
bench compare the device against them.

Inputs: the system of qbg.configs.config(idx) and the coefficients
gram_numpy(random_coefficients(BASE_SEED + 50 + idx, g2, nst)) on its basis.
Stored: the 14 energy terms, fion, sigma (7 x 6, when computed), and H*psi
as 8192 sampled entries plus the per-band squared norms and max |H*psi|.

The twelve-system uneven batch has one system per distinct grid length;
its large members (c4[0] on 210^3, c4[6] on 198^3, c4[7] on 168^3, c4[10] on
224^3) are written the same way from config4_systems()[k] with the
coefficients gram_numpy(random_coefficients(BASE_SEED + 60 + k, g2, nst)),
energies, fion and sampled H*psi (no stress), as oracle_c4_<k>.npz.

Usage: python tests/golden/make_oracle_fixtures.py 2 4   (writes oracle_c3.npz,
oracle_c5.npz next to this script); python tests/golden/make_oracle_fixtures.py
c4-0 c4-6 c4-7 c4-10
"""
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "qbox-public_amd")]

import oracle  # noqa: E402
from qbg import configs as cf  # noqa: E402

NAMES = {2: "c3", 4: "c5"}
STRESS = {2: False, 4: True}
NSAMPLE = 8192


def coefficients(idx, g2, nst):
    return cf.gram_numpy(cf.random_coefficients(cf.BASE_SEED + 50 + idx, g2, nst))


def sample_indices(nst, ngw):
    rng = np.random.Generator(np.random.PCG64(77))
    return rng.integers(0, nst, NSAMPLE), rng.integers(0, ngw, NSAMPLE)


def make(idx):
    if isinstance(idx, str):                      # c4-<k>: a member of the uneven batch
        k = int(idx.split("-")[1])
        s = cf.config4_systems()[k]
        o = oracle.Oracle(s)
        c = cf.gram_numpy(cf.random_coefficients(cf.BASE_SEED + 60 + k, o.basis(0)["g2"],
                                                 s.nst))
        name, stress = f"c4_{k}", False
    else:
        s = cf.config(idx)
        o = oracle.Oracle(s)
        c = coefficients(idx, o.basis(0)["g2"], s.nst)
        name, stress = NAMES[idx], STRESS[idx]
    t0 = time.time()
    r = o.evaluate(c, hpsi=True, forces=True, stress=stress)
    dt = time.time() - t0
    h = r["hpsi"]
    bi, gi = sample_indices(s.nst, o.ngw)
    out = HERE / f"oracle_{name}.npz"
    np.savez(out, names=np.array(oracle.ENERGY_NAMES),
             energies=np.array([r[k] for k in oracle.ENERGY_NAMES]),
             fion=r["fion"], sigma=r["sigma"] if stress else np.zeros((0, 6)),
             hpsi_band=bi, hpsi_g=gi, hpsi_val=h[bi, gi],
             hpsi_norm2=(np.abs(h) ** 2).sum(axis=1), hpsi_absmax=np.abs(h).max(),
             sizes=np.array([o.ngw, o.ngv, o.np0, o.np1, o.np2, s.nst]))
    print(f"{out.name}: {s.name} etotal {r['etotal']:.12f} ({dt:.1f} s)", flush=True)


if __name__ == "__main__":
    for a in sys.argv[1:]:
        make(a if a.startswith("c4-") else int(a))

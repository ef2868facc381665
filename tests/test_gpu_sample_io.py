"""Wavefunctions through sample files on the device: the grid functions
written for coefficients (real backward transforms on the wavefunction grid,
qbg_ef_wf_grid) against numpy's inverse FFT of the coefficients scattered
onto the full sphere; coefficients read back from them equal the originals;
a whole sample (atomset + wavefunction, nspin 1 and 2) written, read and
evaluated gives the energy of the original system."""
import numpy as np
import pytest

from qbg import configs as cf
from qbg import sample_io as sio
from qbg.species import read_species

pytestmark = pytest.mark.gpu

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def _system(nspin):
    s = cf.si4gs(read_species(GOLD / "Si_VBC_LDA-1.0.xml"))
    if nspin == 2:
        s = cf.spin_polarized(s, 16, 1, 1)
    return s


def _coeffs(s, g2, seed):
    c = cf.random_coefficients(seed, g2, s.nst)
    if getattr(s, "nspin", 1) == 2:
        u = s.nst_up
        return np.concatenate([cf.gram_numpy(c[:u]), cf.gram_numpy(c[u:])])
    return cf.gram_numpy(c)


def test_grid_functions_vs_numpy(ctx):
    import qbg
    s = _system(1)
    print("[case] sample grid functions", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    b = ef.basis(0)
    c = _coeffs(s, b["g2"], 31)
    gf = sio.grid_functions(ef, c)
    nx, ny, nz = ef.wf_grid().np
    idx = b["idx"]
    for n in (0, s.nst - 1):
        F = np.zeros((nz, ny, nx), dtype=np.complex128)          # [k][j][i]
        h, k, l = idx[:, 0] % nx, idx[:, 1] % ny, idx[:, 2] % nz
        F[l, k, h] += c[n]
        mask = np.any(idx != 0, axis=1)
        F[(-idx[mask, 2]) % nz, (-idx[mask, 1]) % ny, (-idx[mask, 0]) % nx] += np.conj(c[n][mask])
        f = np.fft.ifftn(F) * F.size
        assert np.abs(f.imag).max() < 1e-12 * np.abs(f.real).max()
        ref = f.real.ravel()                                      # i + nx (j + ny k)
        assert np.abs(gf[n] - ref).max() < 1e-12 * np.abs(ref).max()
    back = sio.coefficients(ef, gf)
    assert np.abs(back - c).max() < 1e-13 * np.abs(c).max()
    ef.release()


@pytest.mark.parametrize("nspin", [1, 2])
def test_sample_file_round_trip_energy(ctx, tmp_path, nspin):
    import qbg
    s = _system(nspin)
    print(f"[case] sample file nspin {nspin}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    c = _coeffs(s, ef.basis(0)["g2"], 32 + nspin)
    dc = ctx.to_device(c)
    e0 = ef.evaluate(dc, None, forces=True)
    atoms = [sio.Atom(f"Si{i + 1}", "silicon", np.asarray(s.tau[i], dtype=float))
             for i in range(s.nat)]
    wf = sio.wavefunction_data(ef, c, s, nel=16, nempty=1 if nspin == 2 else 0)
    sample = sio.Sample(cell=np.asarray(s.cell, dtype=float),
                        species={"silicon": s.species[0]}, atoms=atoms, wf=wf)
    path = tmp_path / "si4.xml"
    sio.write_sample(path, sample)
    r = sio.read_sample(path)
    s2 = sio.system_from_sample(r)
    assert (s2.nst, s2.nspin, s2.nst_up) == (s.nst, getattr(s, "nspin", 1),
                                             getattr(s, "nst_up", 0) if nspin == 2 else 0)
    ef2 = qbg.EnergyFunctional(ctx, s2)
    assert ef2.ngw == ef.ngw
    c2 = sio.sample_coefficients(ef2, r)
    assert np.abs(c2 - c).max() < 1e-13 * np.abs(c).max()
    e2 = ef2.evaluate(ctx.to_device(c2), None, forces=True)
    assert abs(e2["etotal"] - e0["etotal"]) <= 1e-11 * abs(e0["etotal"])
    assert np.abs(e2["fion"] - e0["fion"]).max() < 1e-10
    ef.release()
    ef2.release()


def test_sample_round_trip_with_reference_cell(ctx, tmp_path):
    """A wavefunction on a strained cell whose basis is fixed by a reference
    cell (Qbox writes <reference_domain>) reads back onto the same basis."""
    import dataclasses
    import qbg
    s0 = _system(1)
    strain = np.array([[1.02, 0.01, 0.0], [0.01, 0.99, 0.0], [0.0, 0.0, 1.01]])
    s = dataclasses.replace(s0, cell=s0.cell @ strain.T, refcell=s0.cell * 1.05)
    print("[case] sample file with a reference cell", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    c = _coeffs(s, ef.basis(0)["g2"], 41)
    e0 = ef.evaluate(ctx.to_device(c), None, forces=True)
    atoms = [sio.Atom(f"Si{i + 1}", "silicon", np.asarray(s.tau[i], dtype=float))
             for i in range(s.nat)]
    wf = sio.wavefunction_data(ef, c, s, nel=16)
    sample = sio.Sample(cell=np.asarray(s.cell, dtype=float),
                        species={"silicon": s.species[0]}, atoms=atoms, wf=wf)
    path = tmp_path / "si4_strained.xml"
    sio.write_sample(path, sample)
    r = sio.read_sample(path)
    assert r.wf.refcell is not None
    s2 = sio.system_from_sample(r)
    ef2 = qbg.EnergyFunctional(ctx, s2)
    assert ef2.ngw == ef.ngw and np.array_equal(ef2.basis(0)["idx"], ef.basis(0)["idx"])
    c2 = sio.sample_coefficients(ef2, r)
    assert np.abs(c2 - c).max() < 1e-13 * np.abs(c).max()
    e2 = ef2.evaluate(ctx.to_device(c2), None, forces=True)
    assert abs(e2["etotal"] - e0["etotal"]) <= 1e-11 * abs(e0["etotal"])
    # a functional on another basis is refused
    ef3 = qbg.EnergyFunctional(ctx, dataclasses.replace(s2, refcell=None))
    with pytest.raises(ValueError):
        sio.sample_coefficients(ef3, r)
    for x in (ef, ef2, ef3):
        x.release()

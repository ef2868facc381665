"""Sample files on the host: base64 encoding and line layout against the
reference's own Base64Transcoder (compiled into oracle/_ref), decoding of
the reference's encoding, and document round trips (atomset with embedded
species, nspin 1 and 2 wavefunctions, base64 and text grid functions)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from qbg import configs as cf
from qbg import sample_io as sio
from qbg.species import read_species

GOLD = Path(__file__).resolve().parent / "golden"


def _ref(oracle_mod):
    r = oracle_mod.reflib()
    if r is None:
        pytest.skip("reference sources not compiled (oracle/_ref)")
    return r


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5, 53, 54, 55, 1000, 8 * 1331])
def test_base64_matches_reference(oracle_mod, n):
    ref = _ref(oracle_mod)
    data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    cap = 2 * n + 64
    buf = C.create_string_buffer(cap)
    m = ref.ref_b64_print(C.c_long(n), C.c_char_p(data), buf, C.c_long(cap))
    assert m >= 0
    assert sio.b64_lines(data) == buf.raw[:m].decode("ascii")
    # the reference decodes what this encoder writes
    text = "".join(sio.b64_lines(data).split()).encode("ascii")
    out = C.create_string_buffer(max(n, 1) + 3)
    k = ref.ref_b64_decode(C.c_long(len(text)), C.c_char_p(text), out, C.c_long(n))
    assert out.raw[:n] == data and k >= n


def _sample(nspin):
    si = read_species(GOLD / "Si_VBC_LDA-1.0.xml")
    rng = np.random.default_rng(nspin)
    atoms = [sio.Atom(f"Si{i}", "silicon", rng.uniform(-3, 3, 3), rng.uniform(-1e-3, 1e-3, 3))
             for i in range(4)]
    grid = (12, 10, 14)
    n = grid[0] * grid[1] * grid[2]
    if nspin == 2:
        sds = [sio.SlaterDetData("up", np.zeros(3), 1.0, np.array([1.0, 1.0, 0.0]),
                                 rng.standard_normal((3, n))),
               sio.SlaterDetData("down", np.zeros(3), 1.0, np.array([1.0, 0.5]),
                                 rng.standard_normal((2, n)))]
    else:
        sds = [sio.SlaterDetData(None, np.zeros(3), 1.0, np.array([2.0, 2.0, 1.25]),
                                 rng.standard_normal((3, n)))]
    cell = np.diag([16.0, 15.0, 17.0])
    wf = sio.WavefunctionData(ecut=3.0, nspin=nspin, nel=5, nempty=1, cell=cell, grid=grid,
                              sds=sds)
    return sio.Sample(cell=cell, species={"silicon": si}, atoms=atoms, wf=wf,
                      description="round trip")


@pytest.mark.parametrize("nspin", [1, 2])
@pytest.mark.parametrize("encoding", ["base64", "text"])
def test_sample_round_trip(tmp_path, nspin, encoding):
    s = _sample(nspin)
    path = tmp_path / "s.xml"
    sio.write_sample(path, s, encoding)
    r = sio.read_sample(path)
    assert r.description == s.description
    assert np.array_equal(r.cell, s.cell)
    assert list(r.species) == ["silicon"]
    sp, sp0 = r.species["silicon"], s.species["silicon"]
    assert np.array_equal(sp.vps, sp0.vps) and np.array_equal(sp.phi, sp0.phi)
    for a, b in zip(r.atoms, s.atoms):
        assert (a.name, a.species) == (b.name, b.species)
        assert np.array_equal(a.position, b.position) and np.array_equal(a.velocity, b.velocity)
    w, w0 = r.wf, s.wf
    assert (w.ecut, w.nspin, w.nel, w.nempty, w.grid) == (w0.ecut, w0.nspin, w0.nel, w0.nempty,
                                                          w0.grid)
    assert len(w.sds) == len(w0.sds)
    for a, b in zip(w.sds, w0.sds):
        assert a.spin == b.spin and a.weight == b.weight
        assert np.allclose(a.occ, b.occ, atol=1e-8)
        if encoding == "base64":
            assert np.array_equal(a.gf, b.gf)
        else:
            assert np.allclose(a.gf, b.gf, rtol=1e-15, atol=0)


def test_system_from_sample():
    s = _sample(2)
    s.atoms.append(sio.Atom("H1", "hydrogen", np.ones(3)))
    s.species["hydrogen"] = read_species(GOLD / "H_HSCV_PBE-1.0.xml")
    s.atoms.insert(1, s.atoms.pop())            # interleaved species in the file
    d = sio.system_from_sample(s)
    assert d.na == [4, 1] and d.nspin == 2 and d.nst_up == 3 and d.nst == 5
    assert np.array_equal(d.tau[4], np.ones(3))  # species-major order
    assert d.ecut_ry == 6.0
    assert np.array_equal(d.occ, [1.0, 1.0, 0.0, 1.0, 0.5])


def test_reader_rejects_bad_grid():
    s = _sample(1)
    text = sio.sample_xml(s).replace('grid_function type="double" nx="12"',
                                     'grid_function type="double" nx="13"', 1)
    with pytest.raises(ValueError):
        sio.read_sample(text)

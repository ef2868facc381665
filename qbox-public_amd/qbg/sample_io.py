"""Qbox sample files: the <fpmd:sample> documents Qbox's `save` writes and its
`load` reads (xml/sample.xsd; writers src/AtomSet.cpp:748-766,
src/Atom.cpp:31-48, src/Wavefunction.cpp:919-961, src/SlaterDet.cpp:1382-1553;
reader src/WavefunctionHandler.cpp), as inputs and outputs of the energy
functional.

A wavefunction is stored band by band as grid functions: the real function
psi_n(r) = sum_G c_n(G) exp(iGr) on the wavefunction basis' own grid
(Basis::np), base64 of little-endian doubles in 72-character lines.  Reading
turns them back into coefficients with the forward transform (scaled by
1/np012) on the same grid.  Both transforms run on the device, two bands per
complex transform (psi_n + i psi_{n+1}), through the grid of
qbg_ef_wf_grid.
"""
from __future__ import annotations

import base64
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from .species import NcppSpecies, parse_species_element, read_species, species_xml

FPMD_NS = "http://www.quantum-simulation.org/ns/fpmd/fpmd-1.0"


# ------------------------------------------------------------------ base64
def b64_lines(data: bytes) -> str:
    """Base64 of data in lines of 72 characters (Base64Transcoder::print,
    src/Base64Transcoder.cpp:277-310)."""
    s = base64.b64encode(data).decode("ascii")
    return "".join(s[i:i + 72] + "\n" for i in range(0, len(s), 72))


def b64_doubles(text: str) -> np.ndarray:
    """Little-endian doubles from base64 text (whitespace ignored)."""
    raw = base64.b64decode("".join(text.split()))
    return np.frombuffer(raw, dtype="<f8").astype(np.float64)


# ------------------------------------------------------------------- model
@dataclass
class Atom:
    name: str
    species: str
    position: np.ndarray
    velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))


@dataclass
class SlaterDetData:
    spin: str | None              # "up", "down" or None (nspin 1)
    kpoint: np.ndarray
    weight: float
    occ: np.ndarray               # density_matrix, form "diagonal"
    gf: np.ndarray                # (nbands, nx*ny*nz) real grid functions


@dataclass
class WavefunctionData:
    ecut: float                   # Hartree
    nspin: int
    nel: int
    nempty: int
    cell: np.ndarray              # 3x3, rows a, b, c
    grid: tuple
    sds: list
    refcell: np.ndarray | None = None


@dataclass
class Sample:
    cell: np.ndarray
    species: dict                 # name -> NcppSpecies
    atoms: list
    wf: WavefunctionData | None = None
    description: str = ""


def _local(tag):
    return tag.rsplit("}", 1)[-1]


def _vec(text):
    return np.array([float(x) for x in text.split()])


def _domain(el):
    return np.array([_vec(el.attrib[k]) for k in ("a", "b", "c")])


# ------------------------------------------------------------------ reading
def read_sample(path_or_text) -> Sample:
    """Parse a sample file (path) or its XML text.  Species given by href are
    read relative to the sample's directory."""
    if isinstance(path_or_text, (str, os.PathLike)) and os.path.exists(path_or_text):
        base = os.path.dirname(os.path.abspath(path_or_text))
        with open(path_or_text, "r") as f:
            text = f.read()
    else:
        base, text = os.getcwd(), str(path_or_text)
    root = ET.fromstring(text)
    if _local(root.tag) != "sample":
        raise ValueError("not a sample document")
    out = Sample(cell=np.zeros((3, 3)), species={}, atoms=[])
    for el in root:
        name = _local(el.tag)
        if name == "description":
            out.description = (el.text or "").strip()
        elif name == "atomset":
            _read_atomset(el, out, base)
        elif name == "wavefunction":
            out.wf = _read_wavefunction(el)
    return out


def _read_atomset(el, out, base):
    for sub in el:
        name = _local(sub.tag)
        if name == "unit_cell":
            out.cell = _domain(sub)
        elif name == "species":
            sname = sub.attrib["name"]
            if "href" in sub.attrib and len(sub) == 0:
                out.species[sname] = read_species(os.path.join(base, sub.attrib["href"]))
            else:
                out.species[sname] = parse_species_element(sub)
        elif name == "atom":
            pos = vel = None
            for a in sub:
                if _local(a.tag) == "position":
                    pos = _vec(a.text)
                elif _local(a.tag) == "velocity":
                    vel = _vec(a.text)
            out.atoms.append(Atom(sub.attrib["name"], sub.attrib["species"], pos,
                                  vel if vel is not None else np.zeros(3)))


def _read_wavefunction(el) -> WavefunctionData:
    wf = WavefunctionData(ecut=float(el.attrib.get("ecut", "0")), nspin=int(el.attrib["nspin"]),
                          nel=int(el.attrib["nel"]), nempty=int(el.attrib.get("nempty", "0")),
                          cell=np.zeros((3, 3)), grid=(0, 0, 0), sds=[])
    for sub in el:
        name = _local(sub.tag)
        if name == "domain":
            wf.cell = _domain(sub)
        elif name == "reference_domain":
            wf.refcell = _domain(sub)
        elif name == "grid":
            wf.grid = tuple(int(sub.attrib[k]) for k in ("nx", "ny", "nz"))
        elif name == "slater_determinant":
            wf.sds.append(_read_sd(sub, wf.grid))
    return wf


def _read_sd(el, grid) -> SlaterDetData:
    size = int(el.attrib["size"])
    occ, gfs = None, []
    n = grid[0] * grid[1] * grid[2]
    for sub in el:
        name = _local(sub.tag)
        if name == "density_matrix":
            if sub.attrib.get("form") != "diagonal":
                raise ValueError("density_matrix must be diagonal")
            occ = _vec(sub.text or "")
        elif name == "grid_function":
            dims = tuple(int(sub.attrib[k]) for k in ("nx", "ny", "nz"))
            if dims != tuple(grid):
                raise ValueError("grid_function dimensions differ from the wavefunction grid")
            if sub.attrib.get("type", "double") != "double":
                raise ValueError("only real (Gamma) grid functions are read")
            enc = sub.attrib.get("encoding", "text")
            v = b64_doubles(sub.text or "") if enc == "base64" else _vec(sub.text or "")
            if v.size != n:
                raise ValueError("grid_function size")
            gfs.append(v)
    if len(gfs) != size or occ is None or occ.size != size:
        raise ValueError("slater_determinant size")
    return SlaterDetData(spin=el.attrib.get("spin"), kpoint=_vec(el.attrib["kpoint"]),
                         weight=float(el.attrib["weight"]), occ=occ, gf=np.array(gfs))


# ------------------------------------------------------------------ writing
def _f(x):
    return repr(float(x))


def sample_xml(s: Sample, encoding="base64") -> str:
    """The sample document (Qbox's layout: atomset, then wavefunction)."""
    out = ['<?xml version="1.0" encoding="UTF-8"?>',
           f'<fpmd:sample xmlns:fpmd="{FPMD_NS}">']
    if s.description:
        out.append(f"<description>{s.description}</description>")
    out.append("<atomset>")
    c = s.cell
    out.append(f'<unit_cell a="{_f(c[0,0])} {_f(c[0,1])} {_f(c[0,2])}" '
               f'b="{_f(c[1,0])} {_f(c[1,1])} {_f(c[1,2])}" '
               f'c="{_f(c[2,0])} {_f(c[2,1])} {_f(c[2,2])}"/>')
    for name, sp in s.species.items():
        out.append(species_xml(sp, name).rstrip("\n"))
    for a in s.atoms:
        p, v = a.position, a.velocity
        out.append(f'  <atom name="{a.name}" species="{a.species}">\n'
                   f"    <position> {_f(p[0])} {_f(p[1])} {_f(p[2])} </position>\n"
                   f"    <velocity> {_f(v[0])} {_f(v[1])} {_f(v[2])} </velocity>\n  </atom>")
    out.append("</atomset>")
    if s.wf is not None:
        out.append(_wavefunction_xml(s.wf, encoding))
    out.append("</fpmd:sample>")
    return "\n".join(out) + "\n"


def _wavefunction_xml(wf: WavefunctionData, encoding) -> str:
    c = wf.cell
    out = [f'<wavefunction ecut="{_f(wf.ecut)}" nspin="{wf.nspin}" nel="{wf.nel}" '
           f'nempty="{wf.nempty}">',
           f'<domain a="{_f(c[0,0])} {_f(c[0,1])} {_f(c[0,2])}"\n'
           f'        b="{_f(c[1,0])} {_f(c[1,1])} {_f(c[1,2])}"\n'
           f'        c="{_f(c[2,0])} {_f(c[2,1])} {_f(c[2,2])}"/>']
    if wf.refcell is not None:
        r = wf.refcell
        out.append(f'<reference_domain a="{_f(r[0,0])} {_f(r[0,1])} {_f(r[0,2])}"\n'
                   f'        b="{_f(r[1,0])} {_f(r[1,1])} {_f(r[1,2])}"\n'
                   f'        c="{_f(r[2,0])} {_f(r[2,1])} {_f(r[2,2])}"/>')
    nx, ny, nz = wf.grid
    out.append(f'<grid nx="{nx}" ny="{ny}" nz="{nz}"/>')
    for sd in wf.sds:
        spin = f' spin="{sd.spin}"' if sd.spin else ""
        k = sd.kpoint
        out.append(f'<slater_determinant{spin} kpoint="{_f(k[0])} {_f(k[1])} {_f(k[2])}"\n'
                   f'  weight="{_f(sd.weight)}" size="{sd.occ.size}">')
        out.append(f'<density_matrix form="diagonal" size="{sd.occ.size}">')
        out.append(" ".join(f"{x:.8f}" for x in sd.occ))
        out.append("</density_matrix>")
        for g in sd.gf:
            if encoding == "base64":
                body = b64_lines(np.ascontiguousarray(g, dtype="<f8").tobytes())
            else:
                body = "".join(f" {x:.16e}" + ("\n" if i % 4 == 3 else "") for i, x in enumerate(g))
                body += "" if body.endswith("\n") else "\n"
            out.append(f'<grid_function type="double" nx="{nx}" ny="{ny}" nz="{nz}" '
                       f'encoding="{encoding}">\n{body}</grid_function>')
        out.append("</slater_determinant>")
    out.append("</wavefunction>")
    return "\n".join(out)


def write_sample(path, s: Sample, encoding="base64"):
    with open(path, "w") as f:
        f.write(sample_xml(s, encoding))


# ------------------------------------------- coefficients <-> grid functions
def _wf_grid(ef):
    return ef.wf_grid()


class _Band:
    """A band of a device coefficient array (what the grid calls take)."""

    def __init__(self, arr, n, ngw):
        self.p = arr.offset(16 * ngw * n)


def grid_functions(ef, c: np.ndarray) -> np.ndarray:
    """psi_n(r) on the wavefunction grid for (nst, ngw) coefficients: the
    real backward transforms (SlaterDet::print, src/SlaterDet.cpp:1423-1440),
    two bands per complex transform on the device."""
    ctx = ef.ctx
    g = _wf_grid(ef)
    c = np.ascontiguousarray(c, dtype=np.complex128)
    nst = c.shape[0]
    dc = ctx.to_device(c)
    f = ctx.array(g.N, np.complex128)
    out = np.zeros((nst, g.N))
    for n in range(0, nst, 2):
        if n + 1 < nst:
            g.fft_backward_pair(dc.offset(16 * ef.ngw * n), dc.offset(16 * ef.ngw * (n + 1)), f)
            z = f.download()
            out[n], out[n + 1] = z.real, z.imag
        else:
            g.fft_backward(_Band(dc, n, ef.ngw), f)
            out[n] = f.download().real
    dc.release()
    f.release()
    return out


def coefficients(ef, gf: np.ndarray) -> np.ndarray:
    """(nst, ngw) coefficients from real grid functions on the wavefunction
    grid: forward transforms scaled by 1/np012 (the reader,
    src/WavefunctionHandler.cpp:604), two bands per complex transform."""
    ctx = ef.ctx
    g = _wf_grid(ef)
    gf = np.asarray(gf, dtype=np.float64)
    nst = gf.shape[0]
    if gf.shape[1] != g.N:
        raise ValueError("grid functions are not on this functional's wavefunction grid")
    dc = ctx.array((nst, ef.ngw), np.complex128)
    f = ctx.array(g.N, np.complex128)
    for n in range(0, nst, 2):
        if n + 1 < nst:
            f.upload(gf[n] + 1j * gf[n + 1])
            g.fft_forward_pair(f, dc.offset(16 * ef.ngw * n), dc.offset(16 * ef.ngw * (n + 1)))
        else:
            f.upload(gf[n] + 0j)
            g.fft_forward(f, _Band(dc, n, ef.ngw))
    out = dc.download()
    dc.release()
    f.release()
    return out


def wavefunction_data(ef, c: np.ndarray, sysdef, nel: int, nempty: int = 0) -> WavefunctionData:
    """The <wavefunction> of coefficients c (nst, ngw) of the system sysdef
    (nspin and occupations from it)."""
    gf = grid_functions(ef, c)
    occ = sysdef.occ
    nspin = getattr(sysdef, "nspin", 1)
    if nspin == 2:
        u = sysdef.nst_up
        sds = [SlaterDetData("up", np.zeros(3), 1.0, occ[:u], gf[:u]),
               SlaterDetData("down", np.zeros(3), 1.0, occ[u:], gf[u:])]
    else:
        sds = [SlaterDetData(None, np.zeros(3), 1.0, occ, gf)]
    g = ef.wf_grid()
    ref = getattr(sysdef, "refcell", None)
    return WavefunctionData(ecut=sysdef.ecut, nspin=nspin, nel=nel, nempty=nempty,
                            cell=np.asarray(sysdef.cell, dtype=float), grid=g.np, sds=sds,
                            refcell=None if ref is None else np.asarray(ref, dtype=float))


def system_from_sample(s: Sample, name="sample"):
    """A system definition (species-major atoms, ecut, bands, occupations,
    nspin, and the wavefunction's reference cell when the sample has one, so
    the functional is built on the basis the grid functions were written
    for) from a sample with an atomset and a wavefunction."""
    from .configs import SystemDef
    if s.wf is None:
        raise ValueError("the sample has no wavefunction")
    order = list(s.species)
    na = [sum(1 for a in s.atoms if a.species == sp) for sp in order]
    tau = np.array([a.position for sp in order for a in s.atoms if a.species == sp])
    occ = np.concatenate([sd.occ for sd in s.wf.sds])
    nspin = s.wf.nspin
    nst_up = s.wf.sds[0].occ.size if nspin == 2 else 0
    return SystemDef(name, np.asarray(s.cell, dtype=float), 2.0 * s.wf.ecut,
                     [s.species[sp] for sp in order], na, int(occ.size), 0, 0.0, tau=tau,
                     nspin=nspin, nst_up=nst_up, occ_bands=occ,
                     refcell=None if s.wf.refcell is None else np.asarray(s.wf.refcell,
                                                                          dtype=float))


def sample_coefficients(ef, s: Sample) -> np.ndarray:
    """All bands of the sample's wavefunction (spin up first) as coefficients
    of the functional ef, which must have the basis the grid functions were
    written for: the sample's reference cell (or its cell when it has none)
    must be the functional's."""
    want = s.wf.refcell if s.wf.refcell is not None else s.wf.cell
    have = ef.refcell if ef.refcell is not None else ef.sysdef.cell
    if not np.allclose(np.asarray(want, dtype=float), np.asarray(have, dtype=float),
                       rtol=0, atol=1e-10):
        raise ValueError("the functional's basis is not the sample's (reference cell differs)")
    gf = np.concatenate([sd.gf for sd in s.wf.sds])
    return coefficients(ef, gf)

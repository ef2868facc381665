"""Qbox species files (the fpmd ``species`` XML of test/potentials/*.xml,
read by src/SpeciesReader.cpp and src/SpeciesHandler.cpp) as inputs of the
energy functional.

``read_species(path)`` returns an :class:`NcppSpecies` for a
``norm_conserving_pseudopotential`` (semilocal potentials and
pseudo-wavefunctions per l, Kleinman-Bylander form when nquad = 0); the
library turns it into projectors itself (Species::initialize_ncpp,
src/Species.cpp:71-507).  ``rcps`` is 1.5 bohr as in src/AtomSet.cpp:43-44.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

RCPS = 1.5


@dataclass
class NcppSpecies:
    symbol: str
    zval: int
    lmax: int
    llocal: int
    deltar: float
    vps: np.ndarray            # (lmax+1, nmesh)
    phi: np.ndarray            # (lmax+1, nmesh); zero rows where absent
    nquad: int = 0
    rquad: float = 0.0
    mass: float = 0.0
    atomic_number: int = 0
    rcps: float = RCPS
    description: str = field(default="", repr=False)
    nlcc: np.ndarray = field(default=None, repr=False)   # <core_density>, or None

    kind = "ncpp"

    @property
    def nmesh(self) -> int:
        return self.vps.shape[1]

    def core_density(self):
        """The core density on the radial mesh (Species nlcc_), or None."""
        if self.nlcc is None or not self.nlcc.size:
            return None
        out = np.zeros(self.nmesh)
        n = min(self.nmesh, self.nlcc.size)
        out[:n] = self.nlcc[:n]
        return out


def _local(tag):
    return tag.rsplit("}", 1)[-1]


def _numbers(text):
    return np.array([float(x) for x in text.split()], dtype=np.float64)


def parse_species(text: str) -> NcppSpecies:
    """Parse the XML text of a species file (fpmd species, version 1.0)."""
    return parse_species_element(ET.fromstring(text))


def parse_species_element(root) -> NcppSpecies:
    """A species element: the root of a species file or a <species> of a
    sample's atomset (SpeciesReader, src/SpeciesHandler.cpp)."""
    vals = {}
    ncpp = None
    for el in root.iter():
        name = _local(el.tag)
        if name in ("symbol", "atomic_number", "mass", "description"):
            vals[name] = (el.text or "").strip()
        if name == "norm_conserving_pseudopotential":
            ncpp = el
        if name == "norm_conserving_semilocal_pseudopotential":
            raise ValueError("semilocal (SLPP) species files are given as configs.SpeciesDef")
    if ncpp is None:
        raise ValueError("no norm_conserving_pseudopotential element")
    p = {}
    proj = {}
    nlcc = None
    for el in ncpp:
        name = _local(el.tag)
        if name == "core_density":          # src/SpeciesHandler.cpp:214-220
            nlcc = _numbers(el.text or "")[:int(el.attrib.get("size", "0"))]
        elif name == "projector":
            l = int(el.attrib["l"])
            size = int(el.attrib["size"])
            d = {"size": size}
            for sub in el:
                d[_local(sub.tag)] = _numbers(sub.text or "")
            proj[l] = d
        else:
            p[name] = (el.text or "").strip()
    lmax = int(p["lmax"])
    llocal = int(p["llocal"])
    nquad = int(p.get("nquad", "0"))
    sizes = {d["size"] for d in proj.values()}
    if len(sizes) != 1:
        raise ValueError("projectors of different sizes")
    n = sizes.pop()
    vps = np.zeros((lmax + 1, n))
    phi = np.zeros((lmax + 1, n))
    for l in range(lmax + 1):
        if l not in proj:
            raise ValueError(f"missing projector l={l}")
        vps[l] = proj[l]["radial_potential"][:n]
        if proj[l].get("radial_function") is not None and proj[l]["radial_function"].size:
            phi[l] = proj[l]["radial_function"][:n]
        elif l != llocal and nquad == 0:
            raise ValueError(f"projector l={l} has no radial_function")
    return NcppSpecies(symbol=vals.get("symbol", ""), zval=int(p["valence_charge"]),
                       lmax=lmax, llocal=llocal, deltar=float(p["mesh_spacing"]), vps=vps,
                       phi=phi, nquad=nquad, rquad=float(p.get("rquad", "0")),
                       mass=float(vals.get("mass", "0") or 0),
                       atomic_number=int(vals.get("atomic_number", "0") or 0),
                       description=vals.get("description", ""), nlcc=nlcc)


def read_species(path) -> NcppSpecies:
    with open(path, "r") as f:
        return parse_species(f.read())


def species_xml(sp: NcppSpecies, name: str) -> str:
    """The expanded <species> element of a norm-conserving species, as
    Species::print writes it (src/Species.cpp:1228-1268)."""
    out = [f'<species name="{name}">',
           f"<description>{sp.description}</description>",
           f"<symbol>{sp.symbol}</symbol>",
           f"<atomic_number>{sp.atomic_number}</atomic_number>",
           f"<mass>{sp.mass:.12e}</mass>",
           "<norm_conserving_pseudopotential>",
           f"<valence_charge>{sp.zval}</valence_charge>",
           f"<lmax>{sp.lmax}</lmax>",
           f"<llocal>{sp.llocal}</llocal>",
           f"<nquad>{sp.nquad}</nquad>",
           f"<rquad>{sp.rquad:.12e}</rquad>",
           f"<mesh_spacing>{sp.deltar:.12e}</mesh_spacing>"]
    if sp.core_density() is not None:               # src/Species.cpp:1310-1316
        out.append(f'<core_density size="{sp.nlcc.size}">')
        out += [repr(float(x)) for x in sp.nlcc]
        out.append("</core_density>")
    for l in range(sp.lmax + 1):
        out.append(f'<projector l="{l}" size="{sp.nmesh}">')
        out.append("<radial_potential>")
        out += [repr(float(x)) for x in sp.vps[l]]
        out.append("</radial_potential>")
        out.append("<radial_function>")
        if np.any(sp.phi[l]):
            out += [repr(float(x)) for x in sp.phi[l]]
        out.append("</radial_function>")
        out.append("</projector>")
    out += ["</norm_conserving_pseudopotential>", "</species>"]
    return "\n".join(out) + "\n"

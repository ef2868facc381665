"""The C ABI without a GPU: libqbg.so loads with every symbol resolved, exports
every function include/qbg.h declares, and the Python binding declares each of
them with the header's number of parameters.  Only calls that touch no device
are made (null-handle argument checks)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from qbg import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "qbg.h"


def declared():
    """name -> parameter count, from the prototypes in the header."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"\b(qbg_[a-z0-9_]+)\s*\(([^()]*)\)\s*;", text):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIBPATH.exists():
        pytest.skip("libqbg.so not built")
    # ctypes opens with RTLD_NOW: an unresolved symbol fails here
    return C.CDLL(str(_lib.LIBPATH))


def test_header_parses():
    d = declared()
    assert len(d) >= 40
    for name in ("qbg_init", "qbg_ef_create", "qbg_ef_energy", "qbg_gram",
                 "qbg_graph_begin", "qbg_fft_backward_pair", "qbg_last_error"):
        assert name in d


def test_every_declared_symbol_exported(lib):
    nm = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIBPATH)],
                        capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in nm.splitlines() if ln.strip()}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for n in declared():
        assert getattr(lib, n) is not None


def test_no_unresolved_library_symbols():
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(_lib.LIBPATH)],
                        capture_output=True, text=True, check=True).stdout
    # only the HIP runtime, the C/C++ runtimes and libm may be left undefined;
    # a symbol of this library's own namespace means a missing definition
    own = [ln for ln in nm.splitlines() if "qbg" in ln]
    assert not own, own


def test_binding_covers_header():
    d = declared()
    missing = [n for n in d if n not in _lib.SIGNATURES]
    assert not missing, missing
    for n, k in d.items():
        assert len(_lib.SIGNATURES[n]) == k, (n, len(_lib.SIGNATURES[n]), k)


def test_null_handles_rejected(lib):
    lib.qbg_last_error.restype = C.c_char_p
    lib.qbg_last_error.argtypes = [C.c_void_p]
    assert lib.qbg_last_error(None) == b"null context"
    m = re.search(r"#define\s+QBG_E_INVAL\s+\((-\d+)\)", HEADER.read_text())
    inval = int(m.group(1))
    for name in ("qbg_sync", "qbg_finalize", "qbg_grid_destroy", "qbg_ef_destroy",
                 "qbg_graph_launch", "qbg_graph_destroy", "qbg_graph_begin"):
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = [C.c_void_p]
        assert f(None) == inval, name

"""CPU: the C-ABI library loads and exports every function include/*.h declares
(no compute calls: there is no GPU here).  include/ca3dmm_mpi.h belongs to
libca3dmm_amd_mpi.so (built where an MPI is installed); every other header to
libca3dmm_amd.so."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


MPI_HEADER = "ca3dmm_mpi.h"


def declared_functions(headers=None):
    names = set()
    pat = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?[\s\*])([A-Za-z_]\w*)\s*\(", re.M)
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h") or (headers is None and fn == MPI_HEADER):
            continue
        if headers is not None and fn not in headers:
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        body = text
        for m in pat.finditer(body):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof", "defined"):
                continue
            names.add(name)
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "hip_dgemm" in names and "ca3dmm_engine_init" in names or "hip_dgemm" in names
    assert len(names) > 20


def test_library_exports_every_declared_symbol():
    from ca3dmm_amd import LIB_PATH
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_mpi_library_exports_the_mpi_header():
    path = os.path.join(ROOT, "ca3dmm_amd", "lib", "libca3dmm_amd_mpi.so")
    if not os.path.exists(path):
        pytest.skip("libca3dmm_amd_mpi.so is built only where an MPI is installed")
    lib = ctypes.CDLL(path)
    names = declared_functions([MPI_HEADER])
    assert "ca3dmm_engine_init_mpi" in names and "mat_redist_engine_init_mpi" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_signatures_cover_exports():
    from ca3dmm_amd._lib import SIGNATURES
    from ca3dmm_amd.engine import _SIGS
    names = set(declared_functions())
    known = set(SIGNATURES) | set(_SIGS)
    assert names <= known, sorted(names - known)


def test_no_oracle_in_product():
    """The product library never links or names the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "ca3dmm_amd")):
        for f in files:
            if f.endswith((".hip", ".cpp", ".h", ".py")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "liboracle" not in txt and "from oracle" not in txt and "import oracle" not in txt, f

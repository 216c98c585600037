"""The C-ABI library builds, loads and exports what its header declares.

CPU only: symbols are resolved, nothing is called that would touch a GPU.
"""
import ctypes
import os
import re

import pytest

import fteik_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fteik", "fteik_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"defined", "if", "while", "return", "sizeof"}
    return sorted(set(n for n in names if n not in skip))


def test_header_declares_reference_api():
    names = declared_functions()
    assert len(names) > 80
    for must in ("fteik_solver3d_initialize", "fteik_solver3d_solveSourceLSM",
                 "fteik_solver2d_solveLSM", "locate_locateEvent",
                 "locate_setTravelTimeField64f", "fteik_graph3d_getIJKVF",
                 "fteik_analytic_getTravelTimeField64f"):
        assert must in names


@pytest.fixture(scope="module")
def lib():
    if not os.path.isfile(fteik_amd.library_path()):
        pytest.fail("libfteik_shared.so is not built: run "
                    "`python -m fteik_amd.build`")
    return ctypes.CDLL(fteik_amd.library_path())


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_python_binding_table_resolves():
    """fteik_amd.load_library sets argtypes for its whole table; this fails
    on any name the library lacks."""
    L = fteik_amd.load_library()
    assert L.fteik_amd_version().startswith(b"fteik_amd")


def test_product_has_no_cpu_fallback():
    """The package never imports the oracle (the CPU checker)."""
    pkg = os.path.join(ROOT, "fteik_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src, f


def test_index_functions_closed_form(lib):
    """The pure index helpers (no GPU involved): model.F90:771-780,
    :827-836 and the 0-based shim fwrapper.c:8-14."""
    lib.fteik_model_grid2indexF.restype = ctypes.c_int
    lib.fteik_model_velGrid2indexF.restype = ctypes.c_int
    lib.fteik_model_grid2index.restype = ctypes.c_int
    nz, nx = 34, 10
    for (i, j, k) in ((1, 1, 1), (34, 10, 21), (7, 3, 11)):
        g = lib.fteik_model_grid2indexF(i, j, k, nz, nz * nx)
        assert g == (k - 1) * nz * nx + (j - 1) * nz + i
        assert lib.fteik_model_grid2index(i - 1, j - 1, k - 1, nz,
                                          nz * nx) == g - 1
        c = lib.fteik_model_velGrid2indexF(i, j, k, nz - 1,
                                           (nz - 1) * (nx - 1))
        assert c == (k - 1) * (nz - 1) * (nx - 1) + (j - 1) * (nz - 1) + i

"""The C-ABI library loads and exports every entry point include/cedar_hip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from cedar_amd import _abi

HEADER = _abi.HEADER


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cdr_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 30
    assert "cdr_hip_residual2d_f64" in names and "cdr_hip_solver2d_solve" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_the_header():
    assert set(declared_functions()) == set(_abi.PROTOTYPES)


def test_status_strings_and_version():
    lib = _abi.load()
    assert lib.cdr_hip_status_string(0).decode() == "ok"
    assert "gfx950" in lib.cdr_hip_version().decode()


def test_argument_errors_without_gpu():
    # argument checks run before any device work: null pointers are rejected with CDR_ERR_ARG
    lib = _abi.load()
    assert lib.cdr_hip_residual2d_f64(None, 5, None, None, None, 10, 10, None) == 1
    assert lib.cdr_hip_solver2d_create(None, 5, 10, 10, None, None, None) == 1
    assert lib.cdr_hip_norm_work2d(2, 2) == 0

"""The C-ABI library loads and exports every entry point include/fvs.h
declares; argument errors are reported without touching a GPU (CPU only)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from finitevolumesolver_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        if h.name == "fvs_io.h":  # host output library, checked below
            continue
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(fvs_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "fvs_numeric_fluxes" in names and "fvs_split_sweep" in names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    missing = [n for n in declared_symbols() if n not in L.SIGNATURES]
    assert not missing, missing


def test_create_rejects_bad_dimension_without_gpu():
    lib = L.load()
    d = L.MethodDesc()
    d.dim = 4
    d.gamma = 1.4
    h = C.c_void_p()
    rc = lib.fvs_create(C.byref(d), None, C.byref(h))
    assert rc == L.FVS_BAD_ARGUMENT
    assert b"dim" in lib.fvs_last_error()
    assert not h.value


def test_create_rejects_bad_gamma_and_enum():
    lib = L.load()
    d = L.MethodDesc()
    d.dim = 2
    d.gamma = 0.5
    h = C.c_void_p()
    assert lib.fvs_create(C.byref(d), None, C.byref(h)) == L.FVS_BAD_ARGUMENT
    d.gamma = 1.4
    d.limiter = 9
    assert lib.fvs_create(C.byref(d), None, C.byref(h)) == L.FVS_BAD_ARGUMENT


def test_check_raises_typed_errors():
    with pytest.raises(L.TimeStepTooLarge):
        L.check(L.FVS_TIME_STEP_TOO_LARGE)
    with pytest.raises(L.FvsError):
        L.check(L.FVS_BAD_ARGUMENT)


def test_hdf5_writer_exports_its_header():
    import ctypes
    from finitevolumesolver_amd import output
    if not output.LIB_PATH.exists():
        import pytest
        pytest.skip("libfvs_hdf5.so not built (no HDF5 headers)")
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "fvs_io.h").read_text(), flags=re.S)
    names = re.findall(r"\b(fvsio_[a-z0-9_]+)\s*\(", text)
    lib = ctypes.CDLL(str(output.LIB_PATH))
    assert names and all(hasattr(lib, n) for n in names)

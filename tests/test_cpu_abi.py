"""The C-ABI library loads and exports every entry point include/hes.h declares
(no GPU needed: nothing is called that touches a device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hes.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hes_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from hyperelasticsolver_amd import _lib
    assert set(declared_symbols()) == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    from hyperelasticsolver_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    L.hes_version.restype = ctypes.c_int
    assert L.hes_version() == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """No CPU fallback: without the library every numeric call raises."""
    import importlib
    from hyperelasticsolver_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    from hyperelasticsolver_amd import EquationsOfState as E
    with pytest.raises(_lib.HesLibraryMissing):
        E.energy(E.Barton2009(), 0.0, [1, 0, 0, 0, 1, 0, 0, 0, 1])

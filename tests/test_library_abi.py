"""The C-ABI library loads and exports every entry point of include/namics_scf.h.

CPU-only: loading the library does not touch a GPU; no compute call is made.
"""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(REPO, "include", "namics_scf.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scf_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    from namics_amd import scf
    path = scf.library_path()
    if not os.path.exists(path):
        from namics_amd import build
        build.build()
    return path


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("scf_system_create", "scf_residual", "scf_solve", "scf_solve_batch", "scf_get_results",
              "scf_get_profiles", "scf_propagate_chain"):
        assert n in names


def test_library_exports_every_declared_symbol(libpath):
    L = ctypes.CDLL(libpath)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_fails_loudly_without_library(tmp_path, monkeypatch):
    from namics_amd import scf
    monkeypatch.setattr(scf, "_LIB", str(tmp_path / "missing.so"))
    monkeypatch.setattr(scf, "_lib", None)
    with pytest.raises(ImportError):
        scf.lib()

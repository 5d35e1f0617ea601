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


def test_ctypes_structs_match_header(tmp_path):
    """The ctypes mirrors in namics_amd.scf have the size and field offsets of
    the header's structs (compiled with the host C compiler)."""
    import shutil
    import subprocess
    from namics_amd import scf
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    pairs = {"scf_solve_info": scf.SolveInfo, "scf_results": scf.Results, "scf_tune_report": scf.TuneReport,
             "scf_batch_tune_report": scf.BatchTuneReport, "scf_job_result": scf.JobResult,
             "scf_served": scf.Served}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "namics_scf.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "sizes"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([cc, "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for cname, py in pairs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[cname + "." + f]) == getattr(py, f).offset, (cname, f)


def test_lbfgs_input_keys():
    """`newton : method : LBFGS` is accepted with the reference's defaults
    (Solve_scf::CheckInput src/solve_scf.cpp:170-175,246-250): m = 6
    correction pairs unless given, out-of-range m falls back to 6."""
    from namics_amd.scf import describe_input
    from tests.golden.cases import CFG1
    d = describe_input(CFG1 + "\nnewton : isaac : method : LBFGS\n")
    assert d["solver"]["method"] == 3 and d["solver"]["m"] == 6
    d = describe_input(CFG1 + "\nnewton : isaac : method : LBFGS\nnewton : isaac : m : 40\n")
    assert d["solver"]["m"] == 40
    d = describe_input(CFG1 + "\nnewton : isaac : method : LBFGS\nnewton : isaac : m : 5000\n")
    assert d["solver"]["m"] == 6
    d = describe_input(CFG1 + "\nnewton : isaac : method : DIIS\n")
    assert d["solver"]["method"] == 0 and d["solver"]["m"] == 10


def test_brr_input_keys():
    """`newton : method : BRR` is accepted with the reference's defaults
    (Solve_scf::CheckInput src/solve_scf.cpp:251-255): m = 10 rank-one terms
    unless given, out of 0..1000 it falls back to 10."""
    from namics_amd.scf import describe_input
    from tests.golden.cases import CFG1
    d = describe_input(CFG1 + "\nnewton : isaac : method : BRR\n")
    assert d["solver"]["method"] == 4 and d["solver"]["m"] == 10
    d = describe_input(CFG1 + "\nnewton : isaac : method : BRR\nnewton : isaac : m : 300\n")
    assert d["solver"]["m"] == 300
    d = describe_input(CFG1 + "\nnewton : isaac : method : BRR\nnewton : isaac : m : 5000\n")
    assert d["solver"]["m"] == 10

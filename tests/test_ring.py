"""Ring molecules against the reference's own outputs.

A ring (`mol : X : ring : true`, Molecule::CheckInput
src/molecule.cpp:584-625) merges the first and last one-segment blocks of a
linear chain: for every site where the first segment may sit a chain starts
there and ends there (mol_linear::ComputePhi src/mol_linear.cpp:35-111),
GN sums over those sites, and the merged segment is counted once
(Molecule::fraction src/molecule.cpp:2686-2697).  The oracle restates that
and matches the reference fixtures (tests/golden/make_ring_fixtures.py): F,
Omega, every molecule's GN, n, theta and mu, every density profile and the
DIIS iteration counts; the library matches the same fixtures on the GPU.
"""
import json
import os

import numpy as np
import pytest

from tests.golden.ring_cases import RING
from tests.helpers import rel_err, scaled_err
from tests.oracle_lib import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(name):
    with open(os.path.join(HERE, "golden", "ring", name + ".json")) as f:
        return json.load(f)


def _solver_text(fx, o, tol):
    return fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : %r\n" \
        "newton : %s : iterationlimit : 5000\n" % (o.newton_name, o.newton_name, tol, o.newton_name)


def _check(o, r, ref, tol):
    assert rel_err(r["free_energy"], ref["free_energy"]) < tol
    assert rel_err(r["grand_potential"], ref["grand_potential"]) < tol
    for m, d in ref["mol"].items():
        for k in ("GN", "n", "theta", "phibulk", "Mu"):
            assert abs(r["mol"][m][k] - d[k]) <= tol * max(1e-300, abs(d[k])), (m, k, r["mol"][m][k], d[k])
    for s, p in ref["seg_phi"].items():
        assert scaled_err(o.interior(r["seg_phi"][s]), p) < tol, s


@pytest.mark.parametrize("case", list(RING))
def test_oracle_ring_matches_reference(case):
    fx = _fixture(case)
    o = Oracle(fx["text"])
    assert o.nvar == len(fx["evaluation"][0]["x"])
    for ev in fx["evaluation"]:
        g, res = o.residual(np.array(ev["x"]))
        assert rel_err(res, ev["residual"]) < 1e-12
        _check(o, o.results(), ev, 1e-12)
    for cv in fx["converged"]:
        o2 = Oracle(_solver_text(fx, o, cv["tolerance"]))
        x, it, res, st = o2.solve(np.zeros(o2.nvar))
        assert st == 0 and it == int(cv["iterations"])
        _check(o2, o2.results(), cv, 1e-10)



@pytest.mark.gpu
@pytest.mark.parametrize("case", list(RING))
def test_gpu_ring_matches_reference(case):
    from namics_amd import ScfSystem
    from tests.test_gpu_parity import TOL_ORC, TOL_REF, check_results
    fx = _fixture(case)
    o = Oracle(fx["text"])
    print("ring case", case, flush=True)
    s = ScfSystem(fx["text"])
    for ev in fx["evaluation"]:
        x = np.array(ev["x"])
        g, res = s.residual(x)
        check_results(s, s.results(), ev, TOL_REF)
        assert rel_err(res, ev["residual"]) < TOL_REF
        go, _ = o.residual(x)
        assert scaled_err(g, go) < TOL_ORC
    s.close()
    for cv in fx["converged"]:
        s = ScfSystem(_solver_text(fx, o, cv["tolerance"]))
        x, it, res, st = s.solve(np.zeros(s.nvar))
        print("ring case", case, "iterations", it, "reference", cv["iterations"], flush=True)
        assert st == 0 and res <= cv["tolerance"]
        check_results(s, s.results(), cv, max(1e-10, 100 * cv["tolerance"]))
        s.close()


def test_ring_refusals_and_fallback():
    """A chain whose ends are not one segment of the same type stays linear,
    as in the reference; rings the library does not run are refused with a
    message naming the limit."""
    base = _fixture("ring1d_planar")["text"]
    lin = base.replace("(A)1(B)8(A)1(B)9(A)1", "(A)2(B)8(A)1(B)9(A)1")
    o = Oracle(lin)
    m = [d for d in o.desc["molecules"] if d["name"] == "ring"][0]
    assert not m["ring"] and m["chainlength"] == 21
    assert [d for d in Oracle(base).desc["molecules"] if d["name"] == "ring"][0]["chainlength"] == 19
    for bad, msg in ((base + "mol : ring : Markov : 2\n", "first-order"),
                     (base.replace("lattice_type : simple_cubic",
                                   "lattice_type : hexagonal\nlat : flat : FJC_choices : 5"), "fjc 1"),
                     (base + "lat : flat : bondlength : 5e-10\nmon : B : valence : -1\nmol : Na : composition : (Na)1\n"
                      "mol : Na : freedom : neutralizer\nmon : Na : valence : 1\nmon : Na : freedom : free\n",
                      "charged rings")):
        with pytest.raises(ValueError, match=msg):
            Oracle(bad)

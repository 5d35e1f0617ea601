"""Branched molecules: the CPU oracle against the reference's own outputs.

The fixtures (tests/golden/branched/*.json) were produced by the reference
program compiled from its sources (tests/golden/make_branched_fixtures.py).
The oracle executes the molecule's row schedule (the reference's
ForwardBra / BackwardBra recursion unrolled by build_branch_plan, shared
with the library) with its own restated lattice steps, so at given fields it
must reproduce the reference to a few ulps, and its DIIS must take the
reference's number of iterations.  The schedule itself is also checked by
hand on one composition.
"""
import json
import os

import numpy as np
import pytest

from tests.helpers import rel_err
from tests.oracle_lib import Oracle
from tests.golden.branched_cases import BRANCHED

HERE = os.path.dirname(os.path.abspath(__file__))
EVAL_TOL = 1e-12


def _fixture(name):
    with open(os.path.join(HERE, "golden", "branched", name + ".json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", list(BRANCHED))
def test_branched_evaluation_matches_reference(case):
    fx = _fixture(case)
    o = Oracle(fx["text"])
    for ev in fx["evaluation"]:
        g, res = o.residual(np.array(ev["x"]))
        r = o.results()
        assert rel_err(res, ev["residual"]) < EVAL_TOL
        assert rel_err(r["free_energy"], ev["free_energy"]) < EVAL_TOL
        assert abs(r["grand_potential"] - ev["grand_potential"]) < EVAL_TOL * (abs(ev["free_energy"]) +
                                                                               abs(ev["grand_potential"]))
        for m, d in ev["mol"].items():
            for k in ("GN", "n", "theta", "phibulk"):
                assert rel_err(r["mol"][m][k], d[k]) < EVAL_TOL, (m, k, r["mol"][m][k], d[k])
            assert abs(r["mol"][m]["Mu"] - d["Mu"]) < EVAL_TOL * max(1.0, abs(d["Mu"])), (m, "Mu")
        for s, p in ev["seg_phi"].items():
            assert rel_err(o.interior(r["seg_phi"][s]), p, 1e-200) < EVAL_TOL, s
        for m, p in ev["mol_phi"].items():
            assert rel_err(o.interior(r["mol_phi"][m]), p, 1e-200) < EVAL_TOL, m


@pytest.mark.parametrize("case", list(BRANCHED))
def test_branched_diis_matches_reference(case):
    fx = _fixture(case)
    for cv in fx["converged"]:
        o0 = Oracle(fx["text"])
        o = Oracle(fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : %r\n"
                   "newton : %s : iterationlimit : 5000\n" % (o0.newton_name, o0.newton_name, cv["tolerance"],
                                                              o0.newton_name))
        x0 = np.array(cv["x0"]) if "x0" in cv else np.zeros(o.nvar)
        x, it, res, st = o.solve(x0)
        assert st == 0
        assert it == int(cv["iterations"])
        r = o.results()
        assert rel_err(r["free_energy"], cv["free_energy"]) < 1e-10
        for s, p in cv["seg_phi"].items():
            assert rel_err(o.interior(r["seg_phi"][s]), p, 1e-200) < 1e-10


def test_branched_parse_by_hand():
    """(A)3[(B)2](A)1(A)3: the A block in front of the side group gives its
    last segment to a one-segment branch point, and the molecule is turned
    around (blocks from the far end of the main chain); generation 1 (the
    B side chain) covers segments 4..5 of the turned chain.  The schedule
    puts the branch point on row 6: one step of the junction row 3 with the
    A weight, times one unweighted step of the side chain's end row 5."""
    from namics_amd.scf import describe_input
    text = BRANCHED["branch3d_box"]["text"]
    b = [m for m in describe_input(text)["molecules"] if m["name"] == "br"][0]["branch"]
    A, B = 0, 1
    assert b["block_seg"] == [A, A, B, A, A]
    assert b["block_n"] == [3, 1, 2, 1, 2]
    assert b["gnr"] == [0, 0, 1, 0, 0]
    assert b["first_b"] == [0, 2] and b["last_b"] == [4, 2]
    assert b["first_s"] == [0, 4] and b["last_s"] == [8, 5]
    assert b["gn_row"] == 8
    assert all(o[5] == 1 for o in b["forward"] + b["backward"])  # no degeneracy weights
    fwd = [o[:5] for o in b["forward"]]
    INIT, PROP, COPY, TIMES = 0, 1, 2, 3
    # main chain rows 0..3, then the side chain rows 4, 5
    assert fwd[:4] == [[INIT, 0, 0, 0, A], [PROP, 1, 0, 0, A], [PROP, 2, 1, 0, A], [PROP, 3, 2, 0, A]]
    g0 = fwd[4][1]
    assert fwd[4] == [COPY, g0, 3, 0, 0]
    assert fwd[5:7] == [[INIT, 4, 0, 0, B], [PROP, 5, 4, 0, B]]
    g2 = fwd[7][1]
    assert fwd[7] == [PROP, g2, g0, 0, A]
    g1 = fwd[9][1]
    assert fwd[8:12] == [[COPY, g0, 5, 0, 0], [PROP, g1, g0, 0, -1], [TIMES, g2, g2, g1, 0], [COPY, 6, g2, 0, 0]]
    assert fwd[12:] == [[PROP, 7, 6, 0, A], [PROP, 8, 7, 0, A]]


def test_branched_refusals():
    from namics_amd.scf import ScfError, describe_input
    base = BRANCHED["branch1d_comb"]["text"]
    comp = "(A)4[(B)3](A)1(A)4[(B)3](A)1(A)3"
    for bad in ("(A)4[(B)3]", "(A)4[(B)3](A)1]", "(A)4[[(B)3](A)1", "(A)4[(Q)3](A)1", "(A)4[(B)0](A)1"):
        with pytest.raises(ScfError):
            describe_input(base.replace(comp, bad))


def test_dendrimer_parse_by_hand():
    """@dend(A,(B)3,3;A,(B)2(C)1,2): the path from the core is A, (B)3, A,
    (B)2, (C)1; the core counts once, the first spacer three times, the
    second generation's branch point three times and its spacer 3 x 2 = 6
    times: chain length 1 + 9 + 3 + 18 = 31.  Path indices: generation 0's
    spacer 1..3, generation 1's 5..7."""
    from namics_amd.scf import describe_input
    text = BRANCHED["dend1d_gen2"]["text"]
    m = [m for m in describe_input(text)["molecules"] if m["name"] == "dend"][0]
    b = m["branch"]
    A, B, C = 1, 2, 3  # S is segment 0
    assert m["chainlength"] == 31 and b["dendrimer"]
    assert b["block_seg"] == [A, B, A, B, C] and b["block_n"] == [1, 3, 1, 2, 1]
    assert b["block_deg"] == [1, 3, 3, 6, 6] and b["n_arm"] == [3, 2]
    assert b["first_s"] == [1, 5] and b["last_s"] == [3, 7]
    assert b["first_b"] == [1, 3] and b["last_b"] == [1, 4]
    assert b["gn_row"] == 0
    # the compositions of the spacer rows carry the blocks' degeneracies
    weights = sorted({o[5] for o in b["backward"] if o[0] == 4})
    assert weights == [1, 3, 6]


def test_dendrimer_refusals():
    from namics_amd.scf import ScfError, describe_input
    base = BRANCHED["dend1d_star"]["text"]
    for bad in ("@dend(A,(B)8,4,(B)3)", "@dend(A,(B)8)", "@dend(Q,(B)8,4)", "@dend(A,(B)8,0)", "@comb(A,(B)8,4)",
                "@water(W)"):
        with pytest.raises(ScfError):
            describe_input(base.replace("@dend(A,(B)8,4)", bad))


def test_comb_parse_by_hand():
    """@comb((A)3;A,(B)3,(A)2,3;(A)4): head (A)3 (path 0..2), the tooth
    (B)3 once with degeneracy 3 (path 3..5), then three times branch point A
    (path 6, 9, 12) and spacer (A)2 (7..8, 10..11, 13..14), and the tail
    (A)4 (15..18): 3 + 3 x 3 + 3 x (1 + 2) + 4 = 25 segments."""
    from namics_amd.scf import describe_input
    text = BRANCHED["comb1d_wall"]["text"]
    m = [m for m in describe_input(text)["molecules"] if m["name"] == "comb"][0]
    b = m["branch"]
    A, B = 1, 2
    assert m["chainlength"] == 25 and b["comb"]
    assert b["block_seg"] == [A, B, A, A, A, A, A, A, A] and b["block_n"] == [3, 3, 1, 2, 1, 2, 1, 2, 4]
    assert b["block_deg"] == [1, 3, 1, 1, 1, 1, 1, 1, 1] and b["n_arm"] == [3]
    assert b["first_s"] == [0, 3, 7, 10, 13, 15] and b["last_s"] == [2, 5, 8, 11, 14, 18]
    assert b["gn_row"] == 0
    # each branch point's forward row becomes a tooth's backward start:
    # one division by the tooth factor and one by the segment weight per tooth
    assert sum(1 for o in b["backward"] if o[0] == 7) == 3 and sum(1 for o in b["backward"] if o[0] == 8) == 3


def test_comb_refusals():
    from namics_amd.scf import ScfError, describe_input
    base = BRANCHED["comb1d_wall"]["text"]
    comp = "@comb((A)3;A,(B)3,(A)2,3;(A)4)"
    for bad in ("@comb((A)3;A,(B)3,(A)2;(A)4)", "@comb((A)3;A,(B)3,(A)2,0;(A)4)", "@comb((A)3;(A)4)",
                "@comb((A)3;Q,(B)3,(A)2,3;(A)4)"):
        with pytest.raises(ScfError):
            describe_input(base.replace(comp, bad))


def test_asymmetric_dendrimer_parse_by_hand():
    """@dend(A,(B)2,2,(C)3,1;A,(B)2,2): generation 0 has two arm types (two
    (B)2 arms, one (C)3 arm), generation 1 one ((B)2, twice per branch point).
    The core counts once; its arms 2 x 2 + 1 x 3; the three branch points of
    generation 1 (one per arm of generation 0) three times; their arms
    3 x 2 x 2 = 12: chain length 1 + 7 + 3 + 12 = 23."""
    from namics_amd.scf import describe_input
    text = BRANCHED["adend1d_gen2"]["text"]
    m = [m for m in describe_input(text)["molecules"] if m["name"] == "dend"][0]
    b = m["branch"]
    A, B, C = 1, 2, 3
    assert m["chainlength"] == 23 and b["dendrimer"]
    assert b["block_seg"] == [A, B, C, A, B] and b["block_n"] == [1, 2, 3, 1, 2]
    assert b["block_deg"] == [1, 2, 1, 3, 6] and b["n_arm"] == [2, 1, 2]
    assert b["first_s"] == [1, 3, 7] and b["last_s"] == [2, 5, 8]

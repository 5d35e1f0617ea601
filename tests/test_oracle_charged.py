"""Charged systems: the CPU oracle against the reference's own outputs.

The fixtures (tests/golden/charged/*.json) were produced by the reference
program compiled from its sources (tests/golden/make_charged_fixtures.py):
a residual evaluation at seeded fields (the potential block included) and a
DIIS solve per case.  The oracle restates the reference's electrostatics in
its array form (PutU's -epsilon EE + valence psi, DoElectrostatics, the
upwind / Jacobi Poisson sweeps of LG1Planar and LGrad1, the neutralizer,
the q psi and eps EE terms of F and Omega), so at given fields it must
reproduce the reference to a few ulps, including the potential, charge and
dielectric profiles, and its DIIS must take the reference's number of
iterations.  Two properties that do not come from the reference are
checked too: the neutralizer makes the bulk neutral, and at convergence psi
solves the discretised Poisson equation the sweep iterates.  The cases with
a fixed surface potential (e.psi0/kT) are pinned at given fields only: the
reference's fixed-potential sweep reads psi0 on the last plate and zero
elsewhere instead of the iterate, and its DIIS runs on them diverge (the
reference stops with GN <= 0), as the oracle's do.
"""
import json
import os

import numpy as np
import pytest

from tests.golden.charged_cases import CHARGED
from tests.helpers import rel_err
from tests.oracle_lib import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
EVAL_TOL = 1e-12


def _fixture(name):
    with open(os.path.join(HERE, "golden", "charged", name + ".json")) as f:
        return json.load(f)


def _solver_text(fx, o, tol):
    return fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : %r\n" \
        "newton : %s : iterationlimit : 5000\n" % (o.newton_name, o.newton_name, tol, o.newton_name)


@pytest.mark.parametrize("case", list(CHARGED))
def test_charged_evaluation_matches_reference(case):
    fx = _fixture(case)
    o = Oracle(fx["text"])
    assert o.desc["charged"] and o.nvar == (len(o.field_names) + 1) * o.M
    for ev in fx["evaluation"]:
        g, res = o.residual(np.array(ev["x"]))
        r = o.results()
        el = o.electro()
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
        for k in ("psi", "q", "eps"):
            assert rel_err(o.interior(el[k]), ev[k], 1e-200) < EVAL_TOL, k


@pytest.mark.parametrize("case", list(CHARGED))
def test_charged_diis_matches_reference(case):
    fx = _fixture(case)
    o0 = Oracle(fx["text"])
    for cv in fx["converged"]:
        o = Oracle(_solver_text(fx, o0, cv["tolerance"]))
        x, it, res, st = o.solve(np.zeros(o.nvar))
        assert st == 0 and res <= cv["tolerance"]
        assert it == int(cv["iterations"]), (it, cv["iterations"])
        r = o.results()
        assert rel_err(r["free_energy"], cv["free_energy"]) < 1e-12
        for k in ("psi", "q", "eps"):
            assert rel_err(o.interior(o.electro()[k]), cv[k], 1e-200) < 1e-10, k


@pytest.mark.parametrize("case", ["el1d_poly", "el1d_cyl", "el1d_stiff"])
def test_neutralizer_makes_the_bulk_neutral(case):
    """sum_i Z_i phibulk_i = 0 over all molecules (Z: valence per segment),
    with the neutralizer's phibulk set by the oracle at every evaluation."""
    fx = _fixture(case)
    o = Oracle(fx["text"])
    o.residual(np.array(fx["evaluation"][0]["x"]))
    r = o.results()
    charge = {}
    val = {s["name"]: s["valence"] for s in o.desc["segments"]}
    for m in o.desc["molecules"]:
        charge[m["name"]] = None
    # net valence per segment from the fixture text's compositions
    comps = {}
    for line in fx["text"].splitlines():
        parts = [p.strip() for p in line.split(":")]
        if len(parts) == 4 and parts[0] == "mol" and parts[2] == "composition":
            comps[parts[1]] = parts[3]
    import re
    for name, c in comps.items():
        blocks = re.findall(r"\(([A-Za-z]+)\)(\d+)", c)
        n = sum(int(k) for _, k in blocks)
        charge[name] = sum(val[s] * int(k) for s, k in blocks) / n
    total = sum(charge[m] * r["mol"][m]["phibulk"] for m in o.mol_names)
    assert abs(total) < 1e-15


def test_converged_potential_solves_poisson():
    """At the converged state of the planar case the swept potential equals
    psi, so psi solves (eps_{x-1}+eps_x)(psi_{x-1}-psi_x) + (eps_x+eps_{x+1})
    (psi_{x+1}-psi_x) = -C q_x with C = 2 e^2 / (eps0 kT b), the lattice form
    of div(eps grad psi) = -rho / eps0 (dimensionless psi = e psi / kT)."""
    fx = _fixture("el1d_poly")
    o0 = Oracle(fx["text"])
    o = Oracle(_solver_text(fx, o0, 1e-11))
    x, it, res, st = o.solve(np.zeros(o.nvar))
    assert st == 0
    o.residual(x)
    el = o.electro()
    e, kT, eps0, b = 1.60217e-19, 1.38065e-23 * 298.15, 8.85418e-12, 3e-10
    C = 2 * e * e / (eps0 * kT * b)
    psi = x[-o.M:].copy()
    psi[0], psi[-1] = psi[1], psi[-2]
    E, q = el["eps"], el["q"]
    lhs = (E[:-2] + E[1:-1]) * (psi[:-2] - psi[1:-1]) + (E[1:-1] + E[2:]) * (psi[2:] - psi[1:-1])
    scale = np.max(np.abs(C * q[1:-1]))
    assert np.max(np.abs(lhs + C * q[1:-1])) < 1e-6 * scale


BASE = """
lat : flat : gradients : 1
lat : flat : n_layers : 20
lat : flat : lattice_type : simple_cubic
lat : flat : bondlength : 3e-10
mon : A : freedom : free
mon : W : freedom : free
mon : Na : freedom : free
mon : A : valence : -1
mon : Na : valence : 1
mol : pol : composition : (A)10
mol : water : composition : (W)1
mol : water : freedom : solvent
mol : Na : composition : (Na)1
"""
RESTRICTED = "mol : pol : freedom : restricted\nmol : pol : theta : 1\n"


def _parse_error(text):
    from namics_amd.scf import ScfError, describe_input
    with pytest.raises(ScfError) as e:
        describe_input(text)
    return str(e.value)


def test_charged_input_refusals():
    from namics_amd.scf import describe_input
    ok = BASE + RESTRICTED + "mol : Na : freedom : neutralizer\n"
    d = describe_input(ok)
    assert d["charged"] and d["neutralizer"] == 2 and d["psi_block"] == len(d["fields"])
    assert d["nvar"] == (len(d["fields"]) + 1) * d["M"]
    # a charged restricted molecule without a neutralizer
    assert "neutralizer" in _parse_error(BASE + RESTRICTED + "mol : Na : freedom : free\nmol : Na : phibulk : 0.01\n")
    # a free salt whose bulk is not neutral, without a neutralizer
    assert "neutralizer" in _parse_error(BASE.replace("(A)10", "(W)10") + RESTRICTED +
                                         "mol : Na : freedom : free\nmol : Na : phibulk : 0.01\n")
    # charges without a bond length
    assert "bondlength" in _parse_error(ok.replace("lat : flat : bondlength : 3e-10\n", ""))
    # charged refined 2-gradient lattices are outside this build
    two = ok.replace("lat : flat : gradients : 1\nlat : flat : n_layers : 20",
                     "lat : flat : gradients : 2\nlat : flat : n_layers_x : 10\nlat : flat : n_layers_y : 10\n"
                     "lat : flat : FJC_choices : 5")
    assert "fjc = 1" in _parse_error(two.replace("simple_cubic", "hexagonal"))
    # a fixed surface potential on a frozen segment; on a free one it is refused
    wall = ok + "mon : S : freedom : frozen\nmon : S : frozen_range : 1;1\nmon : S : e.psi0/kT : 2\n"
    assert describe_input(wall)["fixed_psi0"]
    assert "frozen" in _parse_error(ok + "mon : W : e.psi0/kT : 2\n")
    assert "-25..25" in _parse_error(wall.replace("e.psi0/kT : 2", "e.psi0/kT : 30"))
    # valences that sum to zero in every molecule leave the system uncharged
    d = describe_input(BASE.replace("(A)10", "(A)5(Na)5").replace("(Na)1", "(A)1(Na)1") + RESTRICTED +
                       "mol : Na : freedom : free\nmol : Na : phibulk : 0.01\n")
    assert not d["charged"] and d["nvar"] == len(d["fields"]) * d["M"]

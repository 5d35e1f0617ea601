"""FP32 propagators on the 3-gradient grid path (mixed precision).

While the residual is above the switch level the 6-point box steps run with
FP32 end points under a per-molecule gauge shift (composition sums, GN,
residual, DIIS and thermodynamics in FP64); then the solve continues in
FP64 with a fresh DIIS history, so the converged state is an FP64 one.
Checked against the FP64 solve on the small periodic diblock box at
||g|| <= 1e-9 (F within 1e-8 relative, profiles within 1e-6) and, at the
96^3 benchmark size, against the reference's own converged state of the box
(tests/golden/box96_solve.json) at the input's tolerance 1e-7 (F within
1e-8 relative).  The hipGraph replay and kernel-by-kernel launches give the
same solve.
"""
import json
import os

import numpy as np
import pytest

from tests.helpers import rel_err, scaled_err
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _solve(text, x0, mixed, switch=1e-4, graph=True):
    from namics_amd import ScfSystem, solve_resident_info
    s = ScfSystem(text)
    s.set_graph(graph)
    s.set_precision(mixed, switch)
    s.upload(x0)
    info = solve_resident_info([s])[0]
    return info, s.results(), s


def test_mixed_grid_reaches_the_fp64_state(golden):
    fx = golden["diblock3d_periodic"]
    o0 = Oracle(fx["text"])
    text = fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : 1e-09\n" \
        "newton : %s : iterationlimit : 5000\n" % (o0.newton_name, o0.newton_name, o0.newton_name)
    x0 = np.random.default_rng(7).uniform(-0.3, 0.3, o0.nvar)
    print("diblock3d_periodic fp64", flush=True)
    i64, r64, s64 = _solve(text, x0, False)
    print("diblock3d_periodic mixed", flush=True)
    imx, rmx, smx = _solve(text, x0, True)
    print("iterations", i64["iterations"], imx["iterations"], "fp32", imx["fp32_iterations"], flush=True)
    assert i64["status"] == 0 and imx["status"] == 0 and imx["residual"] <= 1e-9
    assert i64["fp32_iterations"] == 0 and imx["fp32_iterations"] > 0
    assert rel_err(rmx["free_energy"], r64["free_energy"]) < 1e-8
    assert abs(rmx["grand_potential"] - r64["grand_potential"]) < 1e-8 * (abs(r64["free_energy"]) +
                                                                         abs(r64["grand_potential"]))
    for sn in r64["seg_phi"]:
        assert scaled_err(smx.interior(rmx["seg_phi"][sn]), s64.interior(r64["seg_phi"][sn])) < 1e-6
    # kernel by kernel: the same solve as the graph replay
    print("diblock3d_periodic mixed without graph", flush=True)
    ing, rng_, sng = _solve(text, x0, True, graph=False)
    assert ing["iterations"] == imx["iterations"] and ing["fp32_iterations"] == imx["fp32_iterations"]
    assert rng_["free_energy"] == rmx["free_energy"]
    for s in (s64, smx, sng):
        s.close()


def test_mixed_grid_box96_reaches_the_reference_state():
    from namics_amd.workloads import seeded_box_fields
    with open(os.path.join(HERE, "golden", "box96_eval.json")) as f:
        ev = json.load(f)
    with open(os.path.join(HERE, "golden", "box96_solve.json")) as f:
        fx = json.load(f)
    from namics_amd import ScfSystem
    names = ScfSystem(ev["text"])
    x0 = seeded_box_fields(names.field_names, fx["n"], fx["seed"])
    names.close()
    print("box96 mixed", flush=True)
    info, r, s = _solve(ev["text"], x0, True)
    print("box96 mixed iterations", info["iterations"], "fp32", info["fp32_iterations"], "F", r["free_energy"],
          fx["free_energy"], flush=True)
    assert info["status"] == 0 and info["residual"] <= 1e-7
    assert info["fp32_iterations"] > 0
    assert rel_err(r["free_energy"], fx["free_energy"]) < 1e-8
    for m, d in fx["mol"].items():
        for k in ("GN", "n", "theta", "phibulk"):
            assert rel_err(r["mol"][m][k], d[k]) < 1e-8, (m, k)
    s.close()


def test_mixed_grid_refused_where_it_does_not_apply(golden):
    from namics_amd import ScfSystem, scf
    s = ScfSystem(golden["hex3d_particle"]["text"])
    with pytest.raises(scf.ScfError):
        s.set_precision(True, 1e-4)
    s.close()

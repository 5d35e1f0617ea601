"""`mol : <name> : save_memory : true` from the input (the reference's
checkpointed end-point storage, src/molecule.cpp:174-191,2300-2446).

The library keeps about sqrt(2N) of the N end-point rows (checkpoints on a
uniform interval) and recomputes the others in the backward pass, on the
grid path (every first-order chain, 1 to 3 gradients) and on the 2-gradient
strip path.  The recomputation repeats the forward arithmetic, so residuals,
densities and solves are bit-identical to full storage on the same layout
(the reference's save_memory is also output-identical to its full storage),
and they match the reference's own fixtures to the parity bar.  The arena
shrinks accordingly.  Second-order chains refuse the key."""
import re

import numpy as np
import pytest

from tests.golden.cases import CASES
from tests.helpers import rel_err
from tests.test_gpu_parity import TOL_REF, check_results

pytestmark = pytest.mark.gpu

FIRST_ORDER = [c for c, v in CASES.items() if "Markov : 2" not in v["text"]]
SECOND_ORDER = [c for c, v in CASES.items() if "Markov : 2" in v["text"]]


def with_save_memory(text):
    mols = re.findall(r"mol : (\S+) : composition", text)
    solvents = set(re.findall(r"mol : (\S+) : freedom : solvent", text))
    return text + "\n" + "".join("mol : %s : save_memory : true\n" % m for m in mols if m not in solvents)


def _profiles_equal(a, b):
    ra, rb = a.results(), b.results()
    for k in ("free_energy", "grand_potential"):
        assert ra[k] == rb[k], k
    for m in ra["mol"]:
        for k, v in ra["mol"][m].items():
            assert v == rb["mol"][m][k], (m, k)
    for key in ("seg_phi", "mol_phi"):
        for n in ra[key]:
            assert np.array_equal(ra[key][n], rb[key][n]), (key, n)


@pytest.mark.parametrize("case", FIRST_ORDER)
def test_save_memory_from_input_is_bit_identical_to_full_storage(golden, case):
    from namics_amd import ScfSystem, scf
    fx = golden[case]
    newton = re.findall(r"newton : (\S+) :", fx["text"])
    text = fx["text"] + "\nnewton : %s : method : DIIS\n" % (newton[0] if newton else "isaac")
    sm = ScfSystem(with_save_memory(text))
    full = ScfSystem(text)
    lay = sm.layout()
    mode = scf.MODE_GRID if lay["mode"] == "grid" else scf.MODE_FUSED
    full.set_mode(mode)
    assert full.layout()["strip_sx"] == lay["strip_sx"]
    print(case, "layout", lay, flush=True)
    assert sm.device_bytes < full.device_bytes
    ev = fx["evaluation"][0]
    x = np.asarray(ev["x"]) if "x" in ev else np.zeros(sm.nvar)
    g1, r1 = sm.residual(x)
    g0, r0 = full.residual(x)
    assert np.array_equal(g1, g0) and r1 == r0
    _profiles_equal(sm, full)
    # the reference's own evaluation at the same fields (its save_memory is
    # output-identical to full storage)
    check_results(sm, sm.results(), ev, TOL_REF)
    assert rel_err(r1, ev["residual"]) < TOL_REF
    # a bounded DIIS run follows the same trajectory to the last bit
    for s in (sm, full):
        s.set_solver(tolerance=1e-9, iterationlimit=25)
    x1, it1, res1, st1 = sm.solve(x)
    x0, it0, res0, st0 = full.solve(x)
    assert (it1, st1) == (it0, st0) and res1 == res0
    assert np.array_equal(x1, x0)
    sm.close()
    full.close()


@pytest.mark.parametrize("case", SECOND_ORDER[:2])
def test_save_memory_refused_for_second_order_chains(golden, case):
    from namics_amd import ScfError, ScfSystem
    with pytest.raises(ScfError) as e:
        ScfSystem(with_save_memory(golden[case]["text"]))
    assert e.value.code == -12 and "save_memory" in str(e.value)


def test_box96_save_memory_stores_a_subset_and_matches_full_storage():
    """The 96^3 diblock box of the 3-gradient configuration asks for
    save_memory: its 20-segment chain keeps 5 of 20 end-point rows
    (interval 4) plus a 3-row buffer on the grid path, and one residual
    evaluation at the seeded fields equals full storage bit for bit."""
    from namics_amd import ScfSystem
    from namics_amd.workloads import diblock_box, seeded_box_fields
    sm = ScfSystem(diblock_box(96, save_memory=True))
    full = ScfSystem(diblock_box(96, save_memory=False))
    M = sm.M
    # rows of 8-byte doubles saved: 20 - (5 + 3)
    assert full.device_bytes - sm.device_bytes >= 12 * M * 8
    x = seeded_box_fields(sm.field_names, 96, 41)
    print("box96 save_memory residual", flush=True)
    g1, r1 = sm.residual(x)
    g0, r0 = full.residual(x)
    assert np.array_equal(g1, g0) and r1 == r0
    _profiles_equal(sm, full)
    sm.close()
    full.close()


def test_checkpoint_change_after_a_graph_solve_replays_the_new_schedule(golden):
    """The grid path captures one DIIS iteration into a hipGraph on its first
    solve, with the end-point storage schedule of that moment.  Changing the
    checkpoint interval afterwards drops the captured graph: the next solve
    gives exactly what a fresh system with that interval gives."""
    from namics_amd import ScfSystem
    text = golden["diblock3d_periodic"]["text"] + "\nnewton : isaac : method : DIIS\n" \
        "newton : isaac : tolerance : 1e-09\nnewton : isaac : iterationlimit : 40\n"
    x0 = np.random.default_rng(12).uniform(-0.1, 0.1, ScfSystem(text).nvar)
    a = ScfSystem(text)
    assert a.layout()["mode"] == "grid"
    a.solve(x0)                      # captures the iteration with every row stored
    a.set_checkpoint(3)
    xa, ita, ra, sta = a.solve(x0)   # replays a new capture
    b = ScfSystem(text)
    b.set_checkpoint(3)
    b.set_graph(False)
    xb, itb, rb, stb = b.solve(x0)
    assert (ita, sta) == (itb, stb) and ra == rb and np.array_equal(xa, xb)
    a.close()
    b.close()

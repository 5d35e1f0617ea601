"""Generate golden fixtures from the reference Namics program.

Runs the reference, compiled from its own sources into oracle/_ref/namics by
oracle/ref.mk, on small SCF systems and records what it reports.  Run here
(where the reference sources exist); the GPU host only reads the committed
JSON files.

Two kinds of fixture:

* evaluation: seeded fields x are written to a guess file
  (Lattice::ReadGuess format, src/lattice.cpp:926-968), the reference is run
  with `initial_guess : file` and `iterationlimit : 0`, so every reported
  value (segment densities, F, Omega, mu, GN, n, theta, phibulk, ||g||) is the
  residual evaluation at exactly those fields;
* converged: the reference solves to a tight tolerance (pseudohessian in 1-D,
  DIIS elsewhere) or to the path tolerance 1e-7 with DIIS, and the iteration
  count plus the converged values are recorded.

Usage: python tests/golden/make_fixtures.py [case ...]   (writes tests/golden/<case>.json;
all cases when none is named)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "namics")
sys.path.insert(0, REPO)

from tests.golden.cases import CASES  # noqa: E402
from namics_amd.workloads import write_guess  # noqa: E402


def oracle_layout(text):
    """Field-block names and lattice sizes, from the oracle's parser."""
    from tests.oracle_lib import Oracle
    o = Oracle(text)
    return o


def read_kal(path):
    with open(path) as f:
        lines = [l.rstrip("\n") for l in f if l.strip()]
    keys = lines[0].split("\t")
    out = []
    for l in lines[1:]:
        vals = l.split("\t")
        row = {}
        for k, v in zip(keys, vals):
            try:
                row[k] = float(v)
            except ValueError:
                row[k] = v
        out.append(row)
    return out


def read_pro(path, ncoord):
    with open(path) as f:
        lines = [l.rstrip("\n").rstrip("\t") for l in f if l.strip()]
    keys = lines[0].split("\t")[ncoord:]
    data = np.array([[float(v) for v in l.split("\t")[ncoord:]] for l in lines[1:]])
    return {k: data[:, i].tolist() for i, k in enumerate(keys)}


def run_ref(workdir, name, text):
    os.makedirs(os.path.join(workdir, "output"), exist_ok=True)
    inp = os.path.join(workdir, name + ".in")
    with open(inp, "w") as f:
        f.write(text)
    r = subprocess.run([REF, name + ".in"], cwd=workdir, capture_output=True, text=True, timeout=600)
    return r


def output_lines(o, extra_kal=()):
    segs = o.seg_names
    mols = o.mol_names
    kal = ["output : kal : append : false",
           "kal : sys : %s : free_energy" % o.sys_name,
           "kal : sys : %s : grand_potential" % o.sys_name,
           "kal : newton : %s : residual" % o.newton_name,
           "kal : newton : %s : iterations" % o.newton_name]
    for m in mols:
        for p in ("GN", "n", "theta", "phibulk", "Mu"):
            kal.append("kal : mol : %s : %s" % (m, p))
    kal += list(extra_kal)
    pro = ["output : pro : append : false"]
    for s in segs:
        pro.append("pro : mon : %s : phi" % s)
    for m in mols:
        pro.append("pro : mol : %s : phi" % m)
    if o.desc["charged"]:
        for p in ("psi", "q", "eps"):
            pro.append("pro : sys : %s : %s" % (o.sys_name, p))
    return "\n".join(kal + pro) + "\n"


def collect(o, kal, pro):
    row = kal[-1]
    res = {
        "free_energy": row["sys_%s_free_energy" % o.sys_name],
        "grand_potential": row["sys_%s_grand_potential" % o.sys_name],
        "residual": row["newton_%s_residual" % o.newton_name],
        "iterations": row["newton_%s_iterations" % o.newton_name],
        "mol": {},
        "seg_phi": {},
        "mol_phi": {},
    }
    for m in o.mol_names:
        res["mol"][m] = {p: row["mol_%s_%s" % (m, p)] for p in ("GN", "n", "theta", "phibulk", "Mu")}
    for s in o.seg_names:
        res["seg_phi"][s] = pro["mon_%s_phi" % s]
    for m in o.mol_names:
        res["mol_phi"][m] = pro["mol_%s_phi" % m]
    if o.desc["charged"]:
        for p in ("psi", "q", "eps"):
            res[p] = pro["sys_%s_%s" % (o.sys_name, p)]
    return res


def evaluation_fixture(case, seed, amp):
    """Reference evaluation at seeded uniform fields in [-amp, amp]; amp = 0
    evaluates at zero fields without a guess file (the reference misreads
    guess files of refined 3-gradient lattices with several field blocks:
    Solve_scf::Guess strides them as unrefined)."""
    o = oracle_layout(case["text"])
    rng = np.random.default_rng(seed)
    x = rng.uniform(-amp, amp, o.nvar) if amp > 0 else np.zeros(o.nvar)
    with tempfile.TemporaryDirectory() as d:
        guess = ""
        if amp > 0:
            write_guess(os.path.join(d, "x.ig"), o.field_names, o.dims, x, fjc=o.fjc, charged=o.desc["charged"])
            guess = ("sys : %s : initial_guess : file\n"
                     "sys : %s : guess_inputfile : x.ig\n") % (o.sys_name, o.sys_name)
        text = case["text"] + "\n" + output_lines(o) + guess + (
            "newton : %s : method : DIIS\n"
            "newton : %s : iterationlimit : 0\n"
            "newton : %s : e_info : false\n"
            "start\n") % (o.newton_name, o.newton_name, o.newton_name)
        r = run_ref(d, "ev", text)
        if r.returncode != 0:
            raise RuntimeError("reference failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        kal = read_kal(os.path.join(d, "output", "ev.kal"))
        pro = read_pro(os.path.join(d, "output", "ev.pro"), o.gradients)
    out = collect(o, kal, pro)
    out["x"] = x.tolist()
    out["seed"] = seed
    out["amp"] = amp
    return out


def converged_fixture(case, method, tol, itlim=5000, seed=None, amp=0.0):
    o = oracle_layout(case["text"])
    with tempfile.TemporaryDirectory() as d:
        guess = ""
        x = None
        if seed is not None:
            rng = np.random.default_rng(seed)
            x = rng.uniform(-amp, amp, o.nvar)
            write_guess(os.path.join(d, "x.ig"), o.field_names, o.dims, x, fjc=o.fjc, charged=o.desc["charged"])
            guess = ("sys : %s : initial_guess : file\n"
                     "sys : %s : guess_inputfile : x.ig\n") % (o.sys_name, o.sys_name)
        text = case["text"] + "\n" + output_lines(o) + guess + (
            "newton : %s : method : %s\n"
            "newton : %s : tolerance : %r\n"
            "newton : %s : iterationlimit : %d\n"
            "newton : %s : e_info : false\n"
            "start\n") % (o.newton_name, method, o.newton_name, tol, o.newton_name, itlim, o.newton_name)
        r = run_ref(d, "cv", text)
        if r.returncode != 0:
            raise RuntimeError("reference failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        kal = read_kal(os.path.join(d, "output", "cv.kal"))
        pro = read_pro(os.path.join(d, "output", "cv.pro"), o.gradients)
    out = collect(o, kal, pro)
    out["method"] = method
    out["tolerance"] = tol
    if x is not None:
        out["x0"] = x.tolist()
        out["seed"] = seed
        out["amp"] = amp
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    only = set(sys.argv[1:])
    for name, case in CASES.items():
        if only and name not in only:
            continue
        fx = {"case": name, "text": case["text"], "evaluation": [], "converged": []}
        for seed, amp in case.get("eval", [(1, 0.3)]):
            fx["evaluation"].append(evaluation_fixture(case, seed, amp))
        for conv in case.get("converged", []):
            fx["converged"].append(converged_fixture(case, **conv))
        path = os.path.join(HERE, "%s.json" % name)
        with open(path, "w") as f:
            json.dump(fx, f)
        print("wrote", path, flush=True)


if __name__ == "__main__":
    main()

"""Reference fixtures at the benchmark configurations' full sizes.

Runs the reference program (oracle/_ref/namics, built from the reference's
sources by oracle/ref.mk) here, where the reference exists, and records what
it reports, compactly enough to commit:

* brush64x128_eval.json -- the 2-gradient brush (64 x 128, (X)1(A)199,
  chi 0.8) evaluated at seeded fields: F, Omega, per-molecule GN, n, theta,
  phibulk, mu, ||g||, and moments of every density profile;
* box96_eval.json -- the 3-gradient diblock box (96^3, (A)10(B)10,
  theta = 0.05 * 96^3) evaluated at seeded fields plus a Gaussian seed, the
  same quantities;
* box96_solve.json -- the reference's DIIS solve of that box from the same
  fields (tolerance 1e-7): iterations, residual, F, Omega, per-molecule
  results;
* batch1d_sample.json -- the first draws of the batched 1-gradient sweep
  (seed 12345) solved by the reference's DIIS to 1e-7 from zero fields
  (iteration limit 1000): iterations, residual, F, Omega;
* serve_cfg5.json -- a 56 x 182 cylindrical brush (N = 211) and a 54^3
  diblock box drawn by the request-stream generator (seed 11): evaluated at
  the request's starting fields (as above) and solved by the reference's
  DIIS from them (tolerance 1e-7, iteration limit 1000).

Profile moments (per profile p over the interior sites in x-major order,
i = 0, 1, ...): sum p, sum p^2, sum i p / sum p, min, max, and p at the first,
middle and last site.  The starting fields are regenerated from the seeds by
the tests (numpy's PCG64 is reproducible).

Usage: python tests/golden/make_fullsize.py [brush|box|box_solve|batch|serve ...]
"""
import json
import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.make_fixtures import (REF, collect, oracle_layout, output_lines, read_kal,  # noqa: E402
                                        read_pro, run_ref)
from namics_amd.workloads import (batch_draws, brush_cylinder, diblock_box, homopolymer_adsorption,  # noqa: E402
                                  seeded_box_fields, write_guess)


def moments(p):
    p = np.asarray(p, dtype=np.float64)
    i = np.arange(p.size)
    s = float(p.sum())
    return {"sum": s, "sum2": float((p * p).sum()), "centroid": float((i * p).sum() / s) if s != 0 else 0.0,
            "min": float(p.min()), "max": float(p.max()), "first": float(p[0]), "mid": float(p[p.size // 2]),
            "last": float(p[-1])}


def evaluate_at(text, x):
    o = oracle_layout(text)
    with tempfile.TemporaryDirectory() as d:
        write_guess(os.path.join(d, "x.ig"), o.field_names, o.dims, x)
        full = text + "\n" + output_lines(o) + (
            "sys : %s : initial_guess : file\n"
            "sys : %s : guess_inputfile : x.ig\n"
            "newton : %s : method : DIIS\n"
            "newton : %s : iterationlimit : 0\n"
            "newton : %s : e_info : false\n"
            "start\n") % (o.sys_name, o.sys_name, o.newton_name, o.newton_name, o.newton_name)
        r = run_ref(d, "ev", full)
        if r.returncode != 0:
            raise RuntimeError("reference failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        kal = read_kal(os.path.join(d, "output", "ev.kal"))
        pro = read_pro(os.path.join(d, "output", "ev.pro"), o.gradients)
    out = collect(o, kal, pro)
    out["seg_phi"] = {k: moments(v) for k, v in out["seg_phi"].items()}
    out["mol_phi"] = {k: moments(v) for k, v in out["mol_phi"].items()}
    return out


def brush():
    text = brush_cylinder(0.8)
    o = oracle_layout(text)
    seed, amp = 31, 0.05
    x = np.random.default_rng(seed).uniform(-amp, amp, o.nvar)
    out = evaluate_at(text, x)
    out.update(text=text, seed=seed, amp=amp)
    return out


def box():
    n = 96
    text = diblock_box(n)
    o = oracle_layout(text)
    seed = 41
    x = seeded_box_fields(o.field_names, n, seed)
    out = evaluate_at(text, x)
    out.update(text=text, seed=seed, n=n)
    return out


def box_solve():
    """The reference's DIIS solve of the 96^3 box from the same seeded fields
    (the input's tolerance 1e-7 and iteration limit; about 8 minutes here)."""
    n = 96
    text = diblock_box(n)
    o = oracle_layout(text)
    seed = 41
    out = solve_from(text, seeded_box_fields(o.field_names, n, seed))
    out.update(seed=seed, n=n)
    print("box solve", out["iterations"], out["residual"], out["free_energy"], flush=True)
    return out


def batch(count=12):
    draws = batch_draws(512, seed=12345)[:count]
    res = []
    for k, d in enumerate(draws):
        text = homopolymer_adsorption(d["layers"], d["N"], d["chi"], d["chi_s"], d["phib"], markov=d["markov"],
                                      tol=1e-7)
        o = oracle_layout(text)
        with tempfile.TemporaryDirectory() as dd:
            full = text + "\n" + output_lines(o) + (
                "newton : %s : iterationlimit : 1000\n"
                "newton : %s : e_info : false\n"
                "start\n") % (o.newton_name, o.newton_name)
            r = run_ref(dd, "cv", full)
            if r.returncode != 0:
                res.append({"draw": k, "params": d, "error": r.returncode})
                continue
            kal = read_kal(os.path.join(dd, "output", "cv.kal"))
        row = kal[-1]
        res.append({"draw": k, "params": d,
                    "iterations": int(row["newton_%s_iterations" % o.newton_name]),
                    "residual": row["newton_%s_residual" % o.newton_name],
                    "free_energy": row["sys_%s_free_energy" % o.sys_name],
                    "grand_potential": row["sys_%s_grand_potential" % o.sys_name]})
        print("draw", k, res[-1].get("iterations"), res[-1].get("residual"), flush=True)
    return {"draws": res, "seed": 12345, "tolerance": 1e-7, "iterationlimit": 1000}


SERVE_SEED, SERVE_COUNT = 11, 15   # the request-stream test's mix: 12 1-D systems, 2 brushes, 1 box
SERVE_PINNED = (12, 14)           # the 56 x 182 brush (N = 211) and the 54^3 diblock box


def solve_from(text, x, iterationlimit=None):
    """The reference's DIIS solve from fields x with the input's own
    tolerance and iteration limit (or `iterationlimit`): iterations,
    residual, F, Omega and the per-molecule results."""
    if iterationlimit is not None:
        text = re.sub(r"iterationlimit : \d+", "iterationlimit : %d" % iterationlimit, text)
    o = oracle_layout(text)
    with tempfile.TemporaryDirectory() as d:
        write_guess(os.path.join(d, "x.ig"), o.field_names, o.dims, x)
        full = text + "\n" + output_lines(o) + (
            "sys : %s : initial_guess : file\n"
            "sys : %s : guess_inputfile : x.ig\n"
            "newton : %s : e_info : false\n"
            "start\n") % (o.sys_name, o.sys_name, o.newton_name)
        r = run_ref(d, "cv", full)
        if r.returncode != 0:
            raise RuntimeError("reference failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        kal = read_kal(os.path.join(d, "output", "cv.kal"))
    row = kal[-1]
    return {"iterations": int(row["newton_%s_iterations" % o.newton_name]),
            "residual": row["newton_%s_residual" % o.newton_name],
            "free_energy": row["sys_%s_free_energy" % o.sys_name],
            "grand_potential": row["sys_%s_grand_potential" % o.sys_name],
            "mol": {m: {p: row["mol_%s_%s" % (m, p)] for p in ("GN", "n", "theta", "phibulk", "Mu")}
                    for m in o.mol_names}}


def serve():
    """Two requests of the request-stream mix at their full sizes (a
    2-gradient brush and a 3-gradient box): the reference's evaluation at the
    request's starting fields and its DIIS solve from them."""
    from namics_amd.bench_serve import fields_for, make_requests
    reqs = make_requests(SERVE_COUNT, SERVE_SEED)
    out = {"seed": SERVE_SEED, "count": SERVE_COUNT, "requests": {}}
    for k in SERVE_PINNED:
        cls, text, spec = reqs[k]
        o = oracle_layout(text)
        x = fields_for(spec, {"nvar": o.nvar, "fields": o.field_names})
        ev = evaluate_at(text, x)
        cv = solve_from(text, x)
        print("request", k, cls, "iterations", cv["iterations"], "residual", cv["residual"], flush=True)
        out["requests"][str(k)] = {"class": cls, "evaluation": ev, "converged": cv}
        if cv["residual"] > 1e-7:
            # the request's own limit ends the reference's DIIS before it
            # converges: the converged state from a longer run of it
            lv = solve_from(text, x, iterationlimit=4 * cv["iterations"])
            print("request", k, "long run", lv["iterations"], lv["residual"], flush=True)
            out["requests"][str(k)]["converged_long"] = lv
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    which = sys.argv[1:] or ["brush", "box", "box_solve", "batch", "serve"]
    for w in which:
        out = {"brush": brush, "box": box, "box_solve": box_solve, "batch": batch, "serve": serve}[w]()
        name = {"brush": "brush64x128_eval", "box": "box96_eval", "box_solve": "box96_solve",
                "batch": "batch1d_sample", "serve": "serve_cfg5"}[w]
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f)
        print("wrote", name, flush=True)


if __name__ == "__main__":
    main()

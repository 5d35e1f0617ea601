"""The request server on the request-stream configuration at its own sizes.

Fifteen requests drawn by the serving benchmark's generator (seed 11): twelve
1-gradient adsorption systems as in the batched sweep, a 56 x 182 cylindrical
brush with N = 211 and an 85 x 47 one with N = 240 (2-gradient, 32-128 x
32-256 layers, N in [50, 400]), and a 54^3 diblock box (3-gradient, 48^3 -
96^3).  Served together through a pool that holds only part of them at once,
every request gets exactly the iterations, status and free energy of its
solve alone.  Two of them are pinned to the reference program's own runs
(tests/golden/serve_cfg5.json): at its starting fields the brush and the box
evaluate to the reference's F, Omega, mu, GN, ||g|| and density moments
(1e-10); from them the brush converges, as the reference's DIIS does, in
about as many iterations to the reference's free energy.  The box is one
the reference's DIIS converges only after 1901 iterations, past the
request's limit of 1000; where the solve here converges within the limit,
it reaches the reference's converged free energy (1e-8)."""
import json
import os

import numpy as np
import pytest

from tests.helpers import rel_err
from tests.test_gpu_parity import TOL_REF, _moment_check, serve_mix

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    with open(os.path.join(HERE, "golden", "serve_cfg5.json")) as f:
        return json.load(f)


def test_cfg5_mix_served_equals_alone_and_matches_reference():
    from namics_amd import ScfSystem, Server
    from namics_amd.bench_serve import fields_for, make_requests
    from namics_amd.scf import describe_input
    fx = _fixture()
    reqs = make_requests(fx["count"], fx["seed"])
    mix = []
    for k, (cls, text, spec) in enumerate(reqs):
        x0 = fields_for(spec, describe_input(text)) if spec is not None else None
        mix.append((cls, text, x0))
    assert sorted(c for c, _, _ in mix).count("2d") == 2 and [c for c, _, _ in mix].count("3d") == 1
    alone, sizes = [], []
    for k, (cls, text, x0) in enumerate(mix):
        s = ScfSystem(text)
        print("alone", k, cls, s.dims, flush=True)
        x, it, res, st = s.solve(x0)
        alone.append((it, st, s.results()["free_energy"]))
        sizes.append(s.device_bytes)
        ref = fx["requests"].get(str(k))
        if ref is not None:
            ev, cv = ref["evaluation"], ref["converged"]
            g, r = s.residual(x0)
            out = s.results()
            assert rel_err(r, ev["residual"]) < TOL_REF
            assert rel_err(out["free_energy"], ev["free_energy"]) < TOL_REF
            assert abs(out["grand_potential"] - ev["grand_potential"]) < TOL_REF * (
                abs(ev["free_energy"]) + abs(ev["grand_potential"]))
            for m, d in ev["mol"].items():
                for q in ("GN", "n", "theta", "phibulk"):
                    assert rel_err(out["mol"][m][q], d[q]) < TOL_REF, (k, m, q)
            for sn, mom in ev["seg_phi"].items():
                _moment_check(s.interior(out["seg_phi"][sn]), mom, TOL_REF)
            print("request", k, "iterations", it, "status", st, "reference", cv["iterations"], cv["residual"],
                  flush=True)
            if cv["residual"] <= 1e-7:
                # converged within the request's limit in the reference: here too,
                # in about as many iterations, to the same state
                assert st == 0, (k, st)
                assert rel_err(alone[-1][2], cv["free_energy"]) < 1e-8, (k, alone[-1][2], cv["free_energy"])
                assert abs(it - cv["iterations"]) <= max(3, cv["iterations"] // 10), (k, it, cv["iterations"])
            else:
                # the reference's DIIS wanders for longer than the request's
                # limit before it converges (at ||g|| ~ 0.4-3, where last-bit
                # differences in the small least-squares solve decide the path);
                # a solve that converges here within the limit reaches the state
                # the reference converges to in its longer run
                lv = ref["converged_long"]
                assert lv["residual"] <= 1e-7
                if st == 0:
                    assert rel_err(alone[-1][2], lv["free_energy"]) < 1e-8, (k, alone[-1][2], lv["free_energy"])
                else:
                    assert it == cv["iterations"]
        s.close()
    # the largest request fills most of the pool: the others wait for room
    pool = int(1.5 * max(sizes))
    got = serve_mix(Server, pool, 20.0, mix)
    assert len(got) == len(mix)
    for k, (it, st, F) in enumerate(alone):
        r = got[k]
        print("served", k, mix[k][0], r["iterations"], it, flush=True)
        assert (r["iterations"], r["status"]) == (it, st), k
        assert r["free_energy"] == F, k
    assert any(got[k]["queue_ms"] > 0.5 for k in got)

"""The job queue's entry point that allocates nothing (scf_solve_jobs_noalloc)
against the allocating one (scf_solve_jobs): same kernels, so every solve's
iterations, status, residual, free energy and grand potential are equal to
the last bit -- on two calls in a row, the second with fewer jobs and slots,
over the scratch the first call left behind -- and a call whose scratch
would exceed the fixed array is refused before anything runs."""
import numpy as np
import pytest

from tests.golden.cases import CFG1

FIELDS = ("iterations", "status", "fp32_iterations", "residual", "free_energy", "grand_potential")


def _texts():
    base = CFG1 + "\nnewton : isaac : method : DIIS\nnewton : isaac : tolerance : 1e-08\n" \
        "newton : isaac : iterationlimit : 2000\n"
    return [base, base + "\nmon : A : chi_W : 0.7\n"]


def _same(a, b):
    assert len(a) == len(b)
    for ja, jb in zip(a, b):
        for ra, rb in zip(ja, jb):
            for f in FIELDS:
                assert ra[f] == rb[f], (f, ra[f], rb[f])


@pytest.mark.gpu
def test_noalloc_job_queue_equals_allocating_one():
    from namics_amd import DeviceBuffer, ScfError, ScfSystem, solve_jobs
    texts = _texts()
    slots = [[ScfSystem(t) for t in texts] for _ in range(3)]
    nvar = slots[0][0].nvar
    seeds = np.random.default_rng(5).uniform(-0.1, 0.1, (7, nvar))
    dev = DeviceBuffer(seeds)
    try:
        print("noalloc: 7 jobs over 3 slots", flush=True)
        ref7, _ = solve_jobs(slots, dev, 7)
        got7, ms = solve_jobs(slots, dev, 7, noalloc=True)
        assert ms > 0
        _same(got7, ref7)
        assert sum(r["status"] == 0 for job in got7 for r in job) >= 7
        print("noalloc: 3 jobs over 2 slots", flush=True)
        ref3, _ = solve_jobs(slots[:2], dev, 3)
        got3, _ = solve_jobs(slots[:2], dev, 3, noalloc=True)
        _same(got3, ref3)
        _same(got3, ref7[:3])
        # 30000 jobs x 2 solves of records exceed the fixed scratch: refused
        # before any launch (the seed buffer is not read)
        with pytest.raises(ScfError) as e:
            solve_jobs(slots, dev, 30000, noalloc=True)
        assert e.value.code == -10 and "scratch" in str(e.value)
    finally:
        dev.close()
        for row in slots:
            for s in row:
                s.close()

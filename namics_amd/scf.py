"""ctypes binding of libnamics_scf.so (include/namics_scf.h).

Mirrors the reference's entry points for the SCF hot path: a system is
created from `.in` text (the reference's input format), `residual(x)` is
Solve_scf::residuals / System::Classical_residual, `solve(x)` is
Solve_scf::Solve with DIIS, `results()` and `profiles()` are what
System::CheckResults and the `kal`/`pro` outputs report.
"""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libnamics_scf.so")
_lib = None
_D = ctypes.POINTER(ctypes.c_double)

SCF_OK = 0
SCF_NOT_CONVERGED = 1
MODE_AUTO, MODE_FUSED, MODE_GRID = 0, 1, 2
MAX_MOL = 8
MAX_SEG = 8


class ScfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("scf error %d: %s" % (code, msg))
        self.code = code


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int), ("residual", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("fp32_iterations", ctypes.c_int)]


class TuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int), ("chosen", ctypes.c_int), ("mode", ctypes.c_int * 4),
                ("strip_sx", ctypes.c_int * 4), ("ms_per_eval", ctypes.c_double * 4)]


class BatchTuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int), ("chosen", ctypes.c_int), ("checkpoint", ctypes.c_int * 4),
                ("ms_per_launch", ctypes.c_double * 4)]


class JobResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int), ("fp32_iterations", ctypes.c_int),
                ("residual", ctypes.c_double), ("free_energy", ctypes.c_double),
                ("grand_potential", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("phase_ms", ctypes.c_double * 4), ("start_ms", ctypes.c_double)]


class Served(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("gradients", ctypes.c_int), ("residual", ctypes.c_double), ("free_energy", ctypes.c_double),
                ("grand_potential", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("queue_ms", ctypes.c_double), ("solve_ms", ctypes.c_double), ("device_bytes", ctypes.c_size_t)]


class Results(ctypes.Structure):
    _fields_ = [("free_energy", ctypes.c_double), ("grand_potential", ctypes.c_double),
                ("residual", ctypes.c_double), ("n_mol", ctypes.c_int), ("n_seg", ctypes.c_int),
                ("GN", ctypes.c_double * MAX_MOL), ("n", ctypes.c_double * MAX_MOL),
                ("theta", ctypes.c_double * MAX_MOL), ("phibulk", ctypes.c_double * MAX_MOL),
                ("mu", ctypes.c_double * MAX_MOL), ("seg_phibulk", ctypes.c_double * MAX_SEG)]


def library_path():
    return _LIB


def lib():
    """Load the HIP library; raise if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError("namics_amd/libnamics_scf.so is not built; run `python -m namics_amd.build`")
        L = ctypes.CDLL(_LIB)
        vp = ctypes.c_void_p
        L.scf_system_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp), ctypes.c_char_p,
                                        ctypes.c_size_t]
        L.scf_system_create.restype = ctypes.c_int
        L.scf_system_destroy.argtypes = [vp]
        L.scf_system_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.scf_system_nvar.argtypes = [vp]
        L.scf_system_nvar.restype = ctypes.c_int64
        L.scf_system_sites.argtypes = [vp]
        L.scf_system_sites.restype = ctypes.c_int64
        L.scf_system_device_bytes.argtypes = [vp]
        L.scf_system_device_bytes.restype = ctypes.c_size_t
        L.scf_system_set_mode.argtypes = [vp, ctypes.c_int]
        L.scf_system_set_solver.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.scf_residual.argtypes = [vp, _D, _D, _D]
        L.scf_solve.argtypes = [vp, _D, ctypes.POINTER(SolveInfo)]
        L.scf_solve_batch.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(_D),
                                      ctypes.POINTER(SolveInfo)]
        L.scf_upload_fields.argtypes = [vp, _D]
        L.scf_download_fields.argtypes = [vp, _D]
        L.scf_carry_fields.argtypes = [vp, vp]
        L.scf_zero_fields.argtypes = [vp]
        L.scf_solve_resident.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(SolveInfo)]
        L.scf_get_results.argtypes = [vp, ctypes.POINTER(Results)]
        L.scf_get_profiles.argtypes = [vp, _D, _D]
        L.scf_get_electrostatics.argtypes = [vp, _D, _D, _D]
        L.scf_propagate_chain.argtypes = [vp, _D, ctypes.c_int, _D]
        L.scf_get_phase_ms.argtypes = [vp, _D]
        L.scf_system_set_precision.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.scf_system_set_checkpoint.argtypes = [vp, ctypes.c_int]
        L.scf_system_set_graph.argtypes = [vp, ctypes.c_int]
        L.scf_system_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(TuneReport)]
        L.scf_system_layout.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
        L.scf_write_kal.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(SolveInfo)]
        L.scf_write_pro.argtypes = [vp, ctypes.c_char_p]
        L.scf_store_guess.argtypes = [vp, ctypes.c_char_p]
        L.scf_read_guess.argtypes = [vp, ctypes.c_char_p]
        L.scf_run_input_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.scf_autotune_batch.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(BatchTuneReport)]
        L.scf_autotune_forget.argtypes = []
        L.scf_autotune_forget.restype = None
        L.scf_describe_input.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.scf_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(vp)]
        L.scf_server_set_slice.argtypes = [vp, ctypes.c_double]
        L.scf_server_submit.argtypes = [vp, ctypes.c_char_p, _D, ctypes.c_size_t, ctypes.POINTER(ctypes.c_longlong)]
        L.scf_server_progress.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.scf_server_results.argtypes = [vp, ctypes.POINTER(Served), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.scf_server_destroy.argtypes = [vp]
        L.scf_server_destroy.restype = None
        L.scf_buffer_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
        L.scf_buffer_upload.argtypes = [vp, vp, ctypes.c_size_t]
        L.scf_buffer_free.argtypes = [vp]
        L.scf_device_info.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]
        L.scf_device_synchronize.argtypes = []
        L.scf_device_copy.argtypes = [vp, vp, ctypes.c_size_t]
        L.scf_solve_jobs.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, vp, ctypes.c_int,
                                     ctypes.POINTER(JobResult), ctypes.POINTER(ctypes.c_double)]
        L.scf_solve_jobs_noalloc.argtypes = L.scf_solve_jobs.argtypes
        L.scf_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _check(rc):
    if rc < 0:
        raise ScfError(rc, lib().scf_last_error().decode())
    return rc


class ScfSystem:
    """One SCF system resident on the current GPU."""

    def __init__(self, text, start=1, mode=MODE_AUTO):
        L = lib()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        rc = L.scf_system_create(text.encode(), start, ctypes.byref(h), err, 1024)
        if rc != 0:
            raise ScfError(rc, err.value.decode())
        self.h = h
        n = L.scf_system_describe(h, None, 0)
        buf = ctypes.create_string_buffer(n)
        L.scf_system_describe(h, buf, n)
        self.desc = json.loads(buf.value.decode())
        self.M = self.desc["M"]
        self.nvar = self.desc["nvar"]
        self.gradients = self.desc["gradients"]
        self.dims = (self.desc["MX"], self.desc["MY"], self.desc["MZ"])
        self.fjc = self.desc.get("fjc", 1)
        self.seg_names = [s["name"] for s in self.desc["segments"]]
        self.mol_names = [m["name"] for m in self.desc["molecules"]]
        self.field_names = list(self.desc["fields"])
        if mode != MODE_AUTO:
            self.set_mode(mode)

    def close(self):
        if getattr(self, "h", None):
            lib().scf_system_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def device_bytes(self):
        return lib().scf_system_device_bytes(self.h)

    def set_mode(self, mode):
        _check(lib().scf_system_set_mode(self.h, mode))

    def set_solver(self, tolerance=0.0, deltamax=0.0, m=0, iterationlimit=-1):
        _check(lib().scf_system_set_solver(self.h, tolerance, deltamax, m, iterationlimit))

    def residual(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.nvar
        g = np.zeros(self.nvar)
        r = ctypes.c_double(0)
        _check(lib().scf_residual(self.h, _p(x), _p(g), ctypes.byref(r)))
        return g, r.value

    def solve(self, x=None):
        """Solve from fields x (zeros by default) with the input's method:
        DIIS on the device, or Newton-Krylov over device residuals for
        `pseudohessian` / `hessian`.  Returns (x, iterations, residual, status)."""
        x = np.zeros(self.nvar) if x is None else np.array(x, dtype=np.float64)
        info = SolveInfo()
        _check(lib().scf_solve(self.h, _p(x), ctypes.byref(info)))
        return x, info.iterations, info.residual, info.status

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.nvar
        _check(lib().scf_upload_fields(self.h, _p(x)))

    def download(self):
        x = np.zeros(self.nvar)
        _check(lib().scf_download_fields(self.h, _p(x)))
        return x

    def zero_fields(self):
        _check(lib().scf_zero_fields(self.h))

    def carry_from(self, other):
        """Start from `other`'s fields (by segment name), device to device."""
        _check(lib().scf_carry_fields(self.h, other.h))

    def results(self):
        r = Results()
        _check(lib().scf_get_results(self.h, ctypes.byref(r)))
        out = {"free_energy": r.free_energy, "grand_potential": r.grand_potential, "residual": r.residual,
               "mol": {}}
        for i, m in enumerate(self.mol_names):
            out["mol"][m] = dict(GN=r.GN[i], n=r.n[i], theta=r.theta[i], phibulk=r.phibulk[i], Mu=r.mu[i])
        phi = np.zeros((len(self.seg_names), self.M))
        mphi = np.zeros((len(self.mol_names), self.M))
        _check(lib().scf_get_profiles(self.h, _p(phi), _p(mphi)))
        out["seg_phi"] = {s: phi[i] for i, s in enumerate(self.seg_names)}
        out["mol_phi"] = {m: mphi[i] for i, m in enumerate(self.mol_names)}
        if self.desc.get("charged"):
            el = np.zeros((3, self.M))
            _check(lib().scf_get_electrostatics(self.h, _p(el[0]), _p(el[1]), _p(el[2])))
            out["psi"], out["q"], out["eps"] = el[0], el[1], el[2]
        return out

    def set_precision(self, mixed, switch_residual=1e-4):
        """FP64 throughout (mixed=False) or FP32 propagators until the
        residual falls below `switch_residual`, then FP64 (mixed=True)."""
        _check(lib().scf_system_set_precision(self.h, 1 if mixed else 0, float(switch_residual)))

    def autotune(self, evals=5):
        """Time every applicable launch layout and keep the fastest (also for
        systems of the same lattice shape built later in this process)."""
        r = TuneReport()
        _check(lib().scf_system_autotune(self.h, int(evals), ctypes.byref(r)))
        names = {MODE_FUSED: "fused", MODE_GRID: "grid"}
        cands = [dict(mode=names.get(r.mode[i], r.mode[i]), strip_sx=r.strip_sx[i], ms_per_eval=r.ms_per_eval[i])
                 for i in range(r.candidates)]
        return dict(candidates=cands, chosen=r.chosen)

    def layout(self):
        m, sx, th = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().scf_system_layout(self.h, ctypes.byref(m), ctypes.byref(sx), ctypes.byref(th)))
        return dict(mode={MODE_FUSED: "fused", MODE_GRID: "grid"}.get(m.value, m.value), strip_sx=sx.value,
                    threads=th.value)

    def checkpoint_interval(self):
        """Current end-point checkpoint interval (0: every row kept)."""
        L = lib()
        n = L.scf_system_describe(self.h, None, 0)
        buf = ctypes.create_string_buffer(n)
        L.scf_system_describe(self.h, buf, n)
        return json.loads(buf.value.decode())["checkpoint"]

    def set_checkpoint(self, interval):
        """Keep every `interval`-th end-point distribution and recompute the
        rest in the backward pass (save_memory); 0 stores every row."""
        _check(lib().scf_system_set_checkpoint(self.h, int(interval)))

    def set_graph(self, on):
        """Replay the grid-path iteration from a captured hipGraph (default)
        or launch its kernels one by one."""
        _check(lib().scf_system_set_graph(self.h, 1 if on else 0))

    def write_kal(self, path, append=True, iterations=None, residual=None):
        """Append the kal row the input requests (reference format)."""
        info = None
        if iterations is not None or residual is not None:
            info = SolveInfo()
            info.iterations = int(iterations or 0)
            info.residual = float(residual if residual is not None else 0.0)
        _check(lib().scf_write_kal(self.h, str(path).encode(), 1 if append else 0,
                                   ctypes.byref(info) if info is not None else None))

    def write_pro(self, path):
        """Write the profiles the input requests (reference format)."""
        _check(lib().scf_write_pro(self.h, str(path).encode()))

    def store_guess(self, path):
        """Write the current fields as a reference guess file."""
        _check(lib().scf_store_guess(self.h, str(path).encode()))

    def read_guess(self, path):
        """Upload the field blocks a reference guess file holds."""
        _check(lib().scf_read_guess(self.h, str(path).encode()))

    def phase_ms(self):
        """Device time of the last fused residual / solve by phase (ms):
        prepare, chains, densities+residual, rest (DIIS, thermodynamics)."""
        out = np.zeros(4)
        _check(lib().scf_get_phase_ms(self.h, _p(out)))
        return dict(zip(("prepare", "chains", "densities_residual", "diis_other"), out.tolist()))

    def propagate_chain(self, G1, N):
        G1 = np.ascontiguousarray(G1, dtype=np.float64)
        out = np.zeros((N, self.M))
        _check(lib().scf_propagate_chain(self.h, _p(G1), N, _p(out)))
        return out

    def interior(self, a):
        mx, my, mz = self.dims
        a = np.asarray(a)
        if self.gradients == 1:
            return a[self.fjc:mx + self.fjc]
        if self.gradients == 2:
            f = self.fjc
            return a.reshape(mx + 2 * f, my + 2 * f)[f:mx + f, f:my + f].ravel()
        f = self.fjc
        return a.reshape(mx + 2 * f, my + 2 * f, mz + 2 * f)[f:mx + f, f:my + f, f:mz + f].ravel()


def solve_batch(systems, xs=None):
    """Solve several systems in one launch (one workgroup per system)."""
    n = len(systems)
    if xs is None:
        xs = [np.zeros(s.nvar) for s in systems]
    xs = [np.array(x, dtype=np.float64) for x in xs]
    hs = (ctypes.c_void_p * n)(*[s.h.value for s in systems])
    ps = (_D * n)(*[_p(x) for x in xs])
    infos = (SolveInfo * n)()
    _check(lib().scf_solve_batch(hs, n, ps, infos))
    return xs, [(i.iterations, i.residual, i.status) for i in infos]


def describe_input(text, start=1):
    """Layout of an input (fields, sizes, nvar) without creating a system."""
    n = lib().scf_describe_input(text.encode(), start, None, 0)
    if n < 0:
        raise ScfError(n, lib().scf_last_error().decode())
    buf = ctypes.create_string_buffer(n)
    lib().scf_describe_input(text.encode(), start, buf, n)
    return json.loads(buf.value.decode())


class Server:
    """Request server (continuous batching, HBM-footprint admission)."""

    def __init__(self, pool_bytes, streams=4):
        self.h = ctypes.c_void_p()
        _check(lib().scf_server_create(int(pool_bytes), int(streams), ctypes.byref(self.h)))

    def set_slice(self, ms):
        _check(lib().scf_server_set_slice(self.h, float(ms)))

    def submit(self, text, x0=None):
        rid = ctypes.c_longlong(0)
        if x0 is None:
            _check(lib().scf_server_submit(self.h, text.encode(), None, 0, ctypes.byref(rid)))
        else:
            x = np.ascontiguousarray(x0, dtype=np.float64)
            _check(lib().scf_server_submit(self.h, text.encode(), _p(x), x.size, ctypes.byref(rid)))
        return rid.value

    def progress(self):
        q, r = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().scf_server_progress(self.h, ctypes.byref(q), ctypes.byref(r)))
        return q.value, r.value

    def results(self, max_n=4096):
        buf = (Served * max_n)()
        n = ctypes.c_int(0)
        _check(lib().scf_server_results(self.h, buf, max_n, ctypes.byref(n)))
        return [{f: getattr(buf[i], f) for f, _ in Served._fields_} for i in range(n.value)]

    def close(self):
        if self.h:
            lib().scf_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory owned by the library (e.g. a job queue's seed fields)."""

    def __init__(self, array):
        a = np.ascontiguousarray(array, dtype=np.float64)
        self.nbytes = a.nbytes
        self.ptr = ctypes.c_void_p()
        _check(lib().scf_buffer_alloc(self.nbytes, ctypes.byref(self.ptr)))
        _check(lib().scf_buffer_upload(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def close(self):
        if self.ptr:
            lib().scf_buffer_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_info():
    """(compute units, HBM bytes) of the current device."""
    cu = ctypes.c_int(0)
    hb = ctypes.c_size_t(0)
    _check(lib().scf_device_info(ctypes.byref(cu), ctypes.byref(hb)))
    return cu.value, hb.value


def device_copy(dst, src, nbytes):
    """Streaming copy kernel between two DeviceBuffers."""
    _check(lib().scf_device_copy(dst.ptr, src.ptr, int(nbytes)))


def synchronize():
    _check(lib().scf_device_synchronize())


def solve_jobs(slots, seeds, njobs, noalloc=False):
    """Persistent job queue: `slots` is a list of equal chains (lists of
    systems), `seeds` a DeviceBuffer holding njobs x nvar seed fields of chain
    position 0.  Returns (results as a list of per-job lists of dicts,
    kernel_ms).  noalloc: the entry point that takes its scratch from the
    library's fixed device array instead of allocating it."""
    nslots, chain = len(slots), len(slots[0])
    flat = [s for row in slots for s in row]
    hs = (ctypes.c_void_p * len(flat))(*[s.h.value for s in flat])
    res = (JobResult * (njobs * chain))()
    ms = ctypes.c_double(0.0)
    fn = lib().scf_solve_jobs_noalloc if noalloc else lib().scf_solve_jobs
    _check(fn(hs, nslots, chain, seeds.ptr, njobs, res, ctypes.byref(ms)))
    def conv(r):
        d = {f: getattr(r, f) for f, _ in JobResult._fields_}
        d["phase_ms"] = list(d["phase_ms"])
        return d
    out = [[conv(res[j * chain + c]) for c in range(chain)] for j in range(njobs)]
    return out, ms.value


def solve_resident_info(systems):
    """Solve from device-resident fields in one launch; returns one dict per
    system with every field of scf_solve_info."""
    n = len(systems)
    hs = (ctypes.c_void_p * n)(*[s.h.value for s in systems])
    infos = (SolveInfo * n)()
    _check(lib().scf_solve_resident(hs, n, infos))
    return [{f: getattr(i, f) for f, _ in SolveInfo._fields_} for i in infos]


def solve_resident(systems, with_solve_ms=False):
    """Solve from device-resident fields; returns [(iterations, residual, status, kernel_ms)]
    (plus each system's own device solve time in ms with with_solve_ms)."""
    n = len(systems)
    hs = (ctypes.c_void_p * n)(*[s.h.value for s in systems])
    infos = (SolveInfo * n)()
    _check(lib().scf_solve_resident(hs, n, infos))
    if with_solve_ms:
        return [(i.iterations, i.residual, i.status, i.kernel_ms, i.solve_ms) for i in infos]
    return [(i.iterations, i.residual, i.status, i.kernel_ms) for i in infos]


def run_input_file(path):
    """Run every start block of a reference input file (guesses, solves,
    kal / pro / guess files in the output folder, relative to the working
    directory); `var` scans solve every scan point.  Returns (solves,
    converged solves)."""
    n, c = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().scf_run_input_file(str(path).encode(), ctypes.byref(n), ctypes.byref(c)))
    return n.value, c.value


def autotune_batch(systems, evals=3):
    """Time one residual launch of all `systems` together for each end-point
    storage that applies (every row; every other row, odd rows recomputed on
    chip), apply the fastest to all of them and report the timings."""
    n = len(systems)
    hs = (ctypes.c_void_p * n)(*[s.h.value for s in systems])
    rep = BatchTuneReport()
    _check(lib().scf_autotune_batch(hs, n, int(evals), ctypes.byref(rep)))
    return {"candidates": [{"checkpoint": rep.checkpoint[i], "ms_per_launch": rep.ms_per_launch[i]}
                           for i in range(rep.candidates)],
            "chosen": rep.checkpoint[rep.chosen]}


def autotune_forget():
    """Forget the launch layouts remembered per lattice shape."""
    lib().scf_autotune_forget()

"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker for the HIP path and is never
part of the product.  Only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(os.path.dirname(_HERE), "oracle", "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_describe.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.oracle_describe.restype = ctypes.c_int
        L.oracle_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_read_guess.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _D, ctypes.c_char_p, ctypes.c_int]
        L.oracle_residual.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.oracle_solve.argtypes = [ctypes.c_void_p, _D, ctypes.POINTER(ctypes.c_int), _D]
        L.oracle_scalars.argtypes = [ctypes.c_void_p, _D]
        L.oracle_seg_phi.argtypes = [ctypes.c_void_p, _D]
        L.oracle_seg_g1.argtypes = [ctypes.c_void_p, _D]
        L.oracle_mol_phi.argtypes = [ctypes.c_void_p, _D]
        L.oracle_propagate_chain.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D]
        L.oracle_solve_ones.argtypes = [_D, ctypes.c_int, _D]
        L.oracle_apply.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_int]
        L.oracle_apply_markov2.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_int, ctypes.c_int]
        L.oracle_electro.argtypes = [ctypes.c_void_p, _D]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


class Oracle:
    """One parsed SCF system evaluated on the CPU in FP64."""

    def __init__(self, text, start=1):
        err = ctypes.create_string_buffer(1024)
        self.h = lib().oracle_create(text.encode(), start, None, err, 1024)
        if not self.h:
            raise ValueError(err.value.decode())
        n = lib().oracle_describe(self.h, None, 0)
        buf = ctypes.create_string_buffer(n)
        lib().oracle_describe(self.h, buf, n)
        self.desc = json.loads(buf.value.decode())
        d = self.desc
        self.M = d["M"]
        self.nvar = d["nvar"]
        self.gradients = d["gradients"]
        self.dims = (d["MX"], d["MY"], d["MZ"])
        self.fjc = d.get("fjc", 1)
        self.seg_names = [s["name"] for s in d["segments"]]
        self.mol_names = [m["name"] for m in d["molecules"]]
        self.field_names = list(d["fields"])
        self.sys_name = d["sys_name"]
        self.newton_name = d["newton_name"]

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def read_guess(self, text):
        x = np.zeros(self.nvar)
        err = ctypes.create_string_buffer(512)
        if lib().oracle_read_guess(self.h, text.encode(), _p(x), err, 512) != 0:
            raise ValueError(err.value.decode())
        return x

    def residual(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros(self.nvar)
        r = ctypes.c_double(0)
        st = lib().oracle_residual(self.h, _p(x), _p(g), ctypes.byref(r))
        if st != 0:
            raise RuntimeError("oracle residual failed with status %d" % st)
        return g, r.value

    def solve(self, x):
        x = np.array(x, dtype=np.float64)
        it = ctypes.c_int(0)
        r = ctypes.c_double(0)
        st = lib().oracle_solve(self.h, _p(x), ctypes.byref(it), ctypes.byref(r))
        return x, it.value, r.value, st

    def results(self):
        nmol = len(self.mol_names)
        sc = np.zeros(2 + 6 * nmol)
        lib().oracle_scalars(self.h, _p(sc))
        out = {"free_energy": sc[0], "grand_potential": sc[1], "mol": {}}
        for i, m in enumerate(self.mol_names):
            p = sc[2 + 6 * i: 8 + 6 * i]
            out["mol"][m] = dict(GN=p[0], n=p[1], theta=p[2], phibulk=p[3], Mu=p[4], norm=p[5])
        phi = np.zeros((len(self.seg_names), self.M))
        lib().oracle_seg_phi(self.h, _p(phi))
        mphi = np.zeros((nmol, self.M))
        lib().oracle_mol_phi(self.h, _p(mphi))
        out["seg_phi"] = {s: phi[i] for i, s in enumerate(self.seg_names)}
        out["mol_phi"] = {m: mphi[i] for i, m in enumerate(self.mol_names)}
        return out

    def electro(self):
        """psi, q, eps, EE of the last residual (charged systems), M each."""
        a = np.zeros((4, self.M))
        lib().oracle_electro(self.h, _p(a))
        return dict(psi=a[0], q=a[1], eps=a[2], EE=a[3])

    def g1(self):
        a = np.zeros((len(self.seg_names), self.M))
        lib().oracle_seg_g1(self.h, _p(a))
        return a

    def propagate_chain(self, G1, N):
        G1 = np.ascontiguousarray(G1, dtype=np.float64)
        out = np.zeros((N, self.M))
        lib().oracle_propagate_chain(self.h, _p(G1), N, _p(out))
        return out

    def apply(self, X, side=False):
        """One lattice-operator application (propagator stencil with G1=1, or Side)."""
        X = np.array(X, dtype=np.float64)
        out = np.zeros(self.M)
        lib().oracle_apply(self.h, _p(X), _p(out), 1 if side else 0)
        return out

    def apply_markov2(self, X, mol=0, forward=True):
        X = np.array(X, dtype=np.float64).ravel()
        out = np.zeros(3 * self.M)
        lib().oracle_apply_markov2(self.h, _p(X), _p(out), mol, 1 if forward else 0)
        return out.reshape(3, self.M)

    def interior(self, a):
        """Interior sites of a lattice array (ghost layers dropped), x-major."""
        mx, my, mz = self.dims
        a = np.asarray(a)
        if self.gradients == 1:
            return a[self.fjc:mx + self.fjc]
        if self.gradients == 2:
            f = self.fjc
            return a.reshape(mx + 2 * f, my + 2 * f)[f:mx + f, f:my + f].ravel()
        f = self.fjc
        return a.reshape(mx + 2 * f, my + 2 * f, mz + 2 * f)[f:mx + f, f:my + f, f:mz + f].ravel()


def solve_ones(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    c = np.zeros(n)
    if lib().oracle_solve_ones(_p(A), n, _p(c)) != 0:
        raise RuntimeError("NaN in SVD solve")
    return c

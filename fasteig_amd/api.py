"""ctypes binding of the solver layer (libfasteig_hip.so, include/fasteig_hip.h).

    from fasteig_amd.api import Problem
    p = Problem(A)                                   # scipy CSR, full symmetric
    r = p.solve(solver="Chebyshev", nev=64, degree=16, tol=1e-8, which="SR")
    r.eigenvalues, r.residual_norms, p.eigenvectors()

Options use the reference's fasteig_options keys (nev, which, solver, maxiter,
degree, blksize, maxblks, tol, seed, verbosity, print_every) plus the HIP
backend's precision ("fp64" | "mixed"), gram ("native" | "bf16x3"), format
("auto" | "csr" | "sell" | "blocked" | "tiled") and device. Everything runs on the GPU; a missing library is
an error, never a CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _fe

HOST_SO = os.path.join(_fe.LIBDIR, "libfasteig_hip.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        _fe.lib()  # libfe_hip.so first (global symbols)
        if not os.path.exists(HOST_SO):
            raise _fe.FeError(f"{HOST_SO} is not built; run `python -m fasteig_amd.build`")
        L = C.CDLL(HOST_SO)
        vp, i64, i32, dp = C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_double)
        L.feh_last_error.restype = C.c_char_p
        L.feh_problem_create.argtypes = [i32, i64, vp, vp, vp, C.POINTER(vp)]
        L.feh_problem_create_from_deck.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.feh_problem_set_mass.argtypes = [vp, vp]
        L.feh_problem_share_executor.argtypes = [vp, vp]
        L.feh_problem_destroy.argtypes = [vp]
        L.feh_problem_n.argtypes = [vp]
        L.feh_problem_solve.argtypes = [vp, C.c_char_p]
        L.feh_problem_nev.argtypes = [vp]
        L.feh_problem_results.argtypes = [vp, vp, vp, vp, C.POINTER(i64), dp]
        L.feh_problem_eigenvectors.argtypes = [vp, vp]
        L.feh_problem_results_text.argtypes = [vp, C.c_char_p, i32]
        _lib = L
    return _lib


class SolveError(RuntimeError):
    pass


def _check(rc):
    if rc < 0:
        msg = lib().feh_last_error().decode()
        if rc == -2:
            raise ValueError(msg)
        raise SolveError(msg)
    return rc


@dataclass
class Result:
    eigenvalues: np.ndarray
    residual_norms: np.ndarray          # ||Ax - lx|| / ((1+|l|) ||x||), the reference's norm
    scaled_residual_norms: np.ndarray   # ||Ax - lx|| / (||A|| |l| ||x||)
    iterations: int
    runtime: float
    status: int


def _optstring(opts: dict) -> bytes:
    return ";".join(f"{k}={v}" for k, v in opts.items()).encode()


class Problem:
    def __init__(self, A=None, mass=None, deck: str | None = None, share_with: "Problem | None" = None):
        h = C.c_void_p()
        if deck is not None:
            _check(lib().feh_problem_create_from_deck(deck.encode(), C.byref(h)))
        else:
            rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
            ci = np.ascontiguousarray(A.indices, dtype=np.int32)
            v = np.ascontiguousarray(A.data, dtype=np.float64)
            n = len(rp) - 1
            _check(lib().feh_problem_create(n, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data, C.byref(h)))
        self.h = h.value
        self.n = lib().feh_problem_n(self.h)
        if mass is not None:
            m = np.ascontiguousarray(mass, dtype=np.float64)
            _check(lib().feh_problem_set_mass(self.h, m.ctypes.data))
        if share_with is not None:
            _check(lib().feh_problem_share_executor(self.h, share_with.h))

    def solve(self, **opts) -> Result:
        status = _check(lib().feh_problem_solve(self.h, _optstring(opts)))
        return self.results(status)

    def results(self, status=0) -> Result:
        k = lib().feh_problem_nev(self.h)
        ev, rr, sr = np.zeros(k), np.zeros(k), np.zeros(k)
        it = C.c_int64()
        rt = C.c_double()
        _check(lib().feh_problem_results(self.h, ev.ctypes.data, rr.ctypes.data, sr.ctypes.data, C.byref(it),
                                         C.byref(rt)))
        return Result(ev, rr, sr, int(it.value), float(rt.value), status)

    def eigenvectors(self) -> np.ndarray:
        k = lib().feh_problem_nev(self.h)
        out = np.zeros((self.n, k), order="F")
        _check(lib().feh_problem_eigenvectors(self.h, out.ctypes.data))
        return out

    def results_text(self) -> str:
        buf = C.create_string_buffer(1 << 20)
        _check(lib().feh_problem_results_text(self.h, buf, len(buf)))
        return buf.value.decode()

    def close(self):
        if getattr(self, "h", None):
            lib().feh_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

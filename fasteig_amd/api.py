"""ctypes binding of the solver layer (libfasteig_hip.so, include/fasteig_hip.h).

    from fasteig_amd.api import Problem
    p = Problem(A)                                   # scipy CSR, full symmetric
    r = p.solve(solver="Chebyshev", nev=64, degree=16, tol=1e-8, which="SR")
    r.eigenvalues, r.residual_norms, p.eigenvectors()

Options use the reference's fasteig_options keys (nev, which, solver, maxiter,
degree, blksize, maxblks, tol, seed, verbosity, print_every) plus the HIP
backend's precision ("fp64" | "mixed"), gram ("auto" | "native" | "bf16x3" | "bf16x6": the
product mode of FP32 blocks, "auto" measured per shape), format
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
        L.feh_problem_product_tuning.argtypes = [vp, C.c_char_p, i32]
        L.feh_core_lanczos_bounds.argtypes = [vp, C.c_double, C.c_uint, dp, dp]
        L.feh_core_thin_qr.argtypes = [vp, i64, i32, vp, i32]
        L.feh_core_orthonormalize_block.argtypes = [vp, i64, i32, vp, i32, vp, C.c_double, vp,
                                                    C.POINTER(i32)]
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

    def product_tuning(self) -> list[dict]:
        """Measured product modes of this problem's device context (gram="auto"):
        per product shape the time (ms) of the native, BF16x3 and BF16x6 modes
        and their errors relative to the FP64 product of the same FP32 data
        (row sample)."""
        buf = C.create_string_buffer(1 << 16)
        _check(lib().feh_problem_product_tuning(self.h, buf, len(buf)))
        out = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            out.append({"kind": f[0], "rows": int(f[1]), "m": int(f[2]), "p": int(f[3]),
                        "ms": dict(zip(("native", "bf16x3", "bf16x6"), map(float, f[4:7]))),
                        "err": dict(zip(("native", "bf16x3", "bf16x6"), map(float, f[7:10])))})
        return out

    # ---- core numerics on this problem's device (the reference's L4 free functions)
    def lanczos_bounds(self, sign: float = 1.0, seed: int = 1) -> tuple[float, float]:
        """20-step Lanczos bounds of sign * A (src/ChebyshevSolver.cpp:26-49)."""
        lo, hi = C.c_double(), C.c_double()
        _check(lib().feh_core_lanczos_bounds(self.h, float(sign), int(seed), C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def thin_qr(self, X, method: str = "householder") -> tuple[np.ndarray, str]:
        """thin_qr_in_place on the device: method "cholesky" returns (X, "failed")
        unchanged when the Gram matrix is not positive definite; "householder"
        (the orthonormal-basis role: CholQR2, shifted CholQR3 when needed)
        reports "shifted" when the shifted path ran."""
        Xf = np.array(X, dtype=np.float64, order="F", copy=True)
        n, k = Xf.shape
        m = {"cholesky": 0, "householder": 1}[method]
        rc = _check(lib().feh_core_thin_qr(self.h, n, k, Xf.ctypes.data, m))
        return Xf, {0: "ok", 1: "failed", 2: "shifted"}[rc]

    def orthonormalize_block(self, basis, src, rank_tol: float = 1e-12) -> np.ndarray:
        """orthonormalize_block (src/core/Orthogonalization.cpp:717-775) on the
        device; basis may be None (no columns)."""
        S = np.array(src, dtype=np.float64, order="F", copy=True)
        n, b = S.shape
        if basis is None:
            B, act = None, 0
        else:
            B = np.array(basis, dtype=np.float64, order="F", copy=True)
            act = B.shape[1]
        out = np.zeros((n, b), order="F")
        r = C.c_int()
        _check(lib().feh_core_orthonormalize_block(self.h, n, act, None if B is None else B.ctypes.data, b,
                                                   S.ctypes.data, float(rank_tol), out.ctypes.data, C.byref(r)))
        return out[:, :r.value].copy(order="F")

    def close(self):
        if getattr(self, "h", None):
            lib().feh_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

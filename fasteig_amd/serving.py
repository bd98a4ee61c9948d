"""ctypes binding of the continuous-batching solve server (feh_server_* in
include/fasteig_hip.h).

    from fasteig_amd.serving import Server
    srv = Server(max_slots=128, degree=16)
    srv.submit(id=7, A=A, nev=16, tol=1e-8)        # scipy CSR, full symmetric
    while srv.pending():
        srv.step()                                  # one cycle for the whole batch
        for r in srv.results():
            r.id, r.eigenvalues, r.residuals, r.eigenvectors

Requests are admitted in arrival order as slots and HBM budget free up; each
step runs one cycle (filter, CholQR, Rayleigh-Ritz) for every active problem
with grouped launches. Results are the smallest nev eigenpairs, residuals
||Ax - lx|| / ((1 + |l|) ||x||). Everything runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import api


def _lib():
    L = api.lib()
    if not getattr(L, "_serve_ready", False):
        vp, i64, i32, dp = C.c_void_p, C.c_int64, C.c_int, C.c_double
        L.feh_serve_footprint.argtypes = [i32, i64, i32]
        L.feh_serve_footprint.restype = C.c_size_t
        L.feh_server_create.argtypes = [i32, C.c_char_p, C.POINTER(vp)]
        L.feh_server_destroy.argtypes = [vp]
        L.feh_server_submit.argtypes = [vp, i64, i32, i64, vp, vp, vp, i32, dp]
        L.feh_server_step.argtypes = [vp]
        L.feh_server_counts.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
        L.feh_server_peek.argtypes = [vp, C.POINTER(i64), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                      C.POINTER(i32), C.POINTER(i32), vp]
        L.feh_server_pop.argtypes = [vp, vp, vp, vp]
        L.feh_server_stats.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64),
                                       C.POINTER(dp), C.POINTER(C.c_size_t)]
        L._serve_ready = True
    return L


def footprint(n: int, nnz: int, nev: int) -> int:
    """Device bytes a request holds while resident (its admission unit)."""
    return int(_lib().feh_serve_footprint(n, nnz, nev))


@dataclass
class ServeResult:
    id: int
    status: int              # 0 converged, 1 maxiter, 2 rejected
    n: int
    nev: int
    cycles: int
    fp64_cycles: int
    queued: float            # server clock, seconds
    admitted: float
    finished: float
    eigenvalues: np.ndarray = field(repr=False)
    residuals: np.ndarray = field(repr=False)
    eigenvectors: np.ndarray | None = field(repr=False, default=None)


class Server:
    def __init__(self, device: int = 0, **options):
        h = C.c_void_p()
        opts = ";".join(f"{k}={int(v) if isinstance(v, bool) else v}" for k, v in options.items()).encode()
        api._check(_lib().feh_server_create(device, opts, C.byref(h)))
        self.h = h.value
        self.keep_vectors = bool(options.get("keep_vectors", True))
        self._pending = 0

    def submit(self, id: int, A, nev: int, tol: float = 1e-8):
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        n = len(rp) - 1
        api._check(_lib().feh_server_submit(self.h, int(id), n, len(ci), rp.ctypes.data, ci.ctypes.data,
                                            v.ctypes.data, int(nev), float(tol)))
        self._pending += 1

    def step(self) -> int:
        return api._check(_lib().feh_server_step(self.h))

    def counts(self):
        a, q, r = C.c_int(), C.c_int(), C.c_int()
        api._check(_lib().feh_server_counts(self.h, C.byref(a), C.byref(q), C.byref(r)))
        return a.value, q.value, r.value

    def pending(self) -> int:
        """Requests submitted whose results have not been taken yet."""
        return self._pending

    def results(self) -> list[ServeResult]:
        out = []
        L = _lib()
        while True:
            rid, st, n, nev, cy, c64 = (C.c_int64(), C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int())
            times = (C.c_double * 3)()
            if api._check(L.feh_server_peek(self.h, C.byref(rid), C.byref(st), C.byref(n), C.byref(nev),
                                            C.byref(cy), C.byref(c64), times)) == 0:
                break
            ev = np.zeros(nev.value)
            rs = np.zeros(nev.value)
            vec = np.zeros((n.value, nev.value), order="F") if (self.keep_vectors and st.value != 2) else None
            api._check(L.feh_server_pop(self.h, ev.ctypes.data, rs.ctypes.data,
                                        vec.ctypes.data if vec is not None else None))
            self._pending -= 1
            out.append(ServeResult(rid.value, st.value, n.value, nev.value, cy.value, c64.value, times[0],
                                   times[1], times[2], ev, rs, vec))
        return out

    def stats(self) -> dict:
        cy, ad, rt, pk = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        secs, pb = C.c_double(), C.c_size_t()
        api._check(_lib().feh_server_stats(self.h, C.byref(cy), C.byref(ad), C.byref(rt), C.byref(pk),
                                           C.byref(secs), C.byref(pb)))
        return {"cycles": cy.value, "admitted": ad.value, "retired": rt.value, "peak_active": pk.value,
                "cycle_seconds": secs.value, "peak_bytes": pb.value}

    def close(self):
        if self.h:
            _lib().feh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

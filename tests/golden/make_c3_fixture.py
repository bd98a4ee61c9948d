"""Writes tests/golden/c3_eigs.json: the 256 lowest eigenvalues of the full-size
C3 Hamiltonian (H = -0.5 * isotropic 27-point Laplacian + potential on a 96^3
grid, n = 884,736; fasteig_amd.generators.hamiltonian_27pt(96, seed=4)) from the
CPU oracle's FP64 Chebyshev solver run to a relative residual of 1e-12
(adaptive filter bounds, degree 20, block 512, seed 1).

The oracle is the checker, run here on the build machine's cores (about
30 GB of host memory, tens of minutes); the GPU tests and the bench read only
the JSON. Run from the repository root:

    python tests/golden/make_c3_fixture.py [threads]
"""
import hashlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fasteig_amd import generators as G  # noqa: E402
from oracle import oracle as orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_eigs.json")


def matrix_digest(A):
    h = hashlib.sha256()
    for a in (A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else orc.max_threads()
    orc.set_threads(threads)
    A = G.hamiltonian_27pt(96, seed=4)
    print(f"C3: n={A.shape[0]} nnz={A.nnz}; oracle Chebyshev FP64, nev 256, degree 20, tol 1e-12, "
          f"{threads} threads", flush=True)
    t0 = time.time()
    r = orc.chebyshev(A, 256, degree=20, maxiter=80, tol=1e-12, which="sr", seed=1, bounds="ritz",
                      want_vectors=False)
    dt = time.time() - t0
    print(f"status {r['status']} cycles {r['iterations']} max relres {r['residuals'].max():.3e} "
          f"in {dt:.0f} s", flush=True)
    if r["residuals"].max() > 1e-12:
        raise SystemExit("oracle did not reach 1e-12")
    json.dump({
        "workload": "C3: H = -0.5 L27 + V on 96^3 (generators.hamiltonian_27pt(96, seed=4)), lowest 256",
        "n": int(A.shape[0]), "nnz": int(A.nnz), "matrix_sha256_16": matrix_digest(A),
        "solver": "oracle Chebyshev FP64, bounds=ritz, degree 20, block 512, seed 1, tol 1e-12",
        "cycles": int(r["iterations"]), "seconds": round(dt, 1), "threads": threads,
        "host": platform.processor() or platform.machine(),
        "eigenvalues": [float(v) for v in r["eigenvalues"]],
        "relres": [float(v) for v in r["residuals"]],
    }, open(OUT, "w"), indent=1)
    print("wrote", OUT, flush=True)


if __name__ == "__main__":
    main()

"""The full-size C3 configuration against the oracle's answer.

C3: H = -0.5 * (isotropic 27-point Laplacian) + potential on a 96^3 grid
(n = 884,736, 23.4 M nonzeros; generators.hamiltonian_27pt(96, seed=4)),
lowest 256 eigenpairs, Chebyshev degree 20 with adaptive bounds. The oracle's
FP64 eigenvalues (run to a relative residual of 1e-12) are in
tests/golden/c3_eigs.json (written by make_c3_fixture.py; 27 minutes of CPU).

Bars (the target's): every eigenvalue of the mixed FP32/FP64 solve and of the
plain FP64 solve within 1e-10 relative of the oracle's, and every reported
residual ||Ax - lx|| / ((1 + |l|) ||x||) at or below 1e-8.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from fasteig_amd import generators as G
from fasteig_amd.api import Problem

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(__file__), "golden", "c3_eigs.json")


@pytest.fixture(scope="module", autouse=True)
def _require_gpu(dev):
    yield


def digest(A):
    h = hashlib.sha256()
    for a in (A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


@pytest.fixture(scope="module")
def c3():
    ref = json.load(open(FIX))
    A = G.hamiltonian_27pt(96, seed=4)
    assert A.shape[0] == ref["n"] and A.nnz == ref["nnz"] and digest(A) == ref["matrix_sha256_16"]
    p = Problem(A)
    yield p, np.array(ref["eigenvalues"])
    p.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_c3_full_size_matches_oracle(c3, precision):
    p, ev = c3
    print(f"C3 full size, precision={precision}", flush=True)
    r = p.solve(solver="Chebyshev", nev=256, degree=20, tol=1e-8, maxiter=60, which="SR", seed=1,
                cheb_bounds="ritz", precision=precision)
    assert r.status == 0
    err = np.abs(r.eigenvalues - ev) / np.abs(ev)
    assert err.max() <= 1e-10, (precision, err.max(), int(err.argmax()))
    assert r.residual_norms.max() <= 1e-8, (precision, r.residual_norms.max())

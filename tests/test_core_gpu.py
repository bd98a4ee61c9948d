"""Direct GPU parity of the block solvers' core numerics (through the C ABI of
libfasteig_hip.so) with the oracle: the seeded 20-step Lanczos bound estimate,
Cholesky QR and the shifted CholQR3 that takes the place of Householder QR, and
the rank-revealing orthonormalize_block. These run inside whole solves too;
here each is compared on its own, where a wrong bound, a wrong rank decision
or a lost direction would otherwise only change a convergence rate.

Bars: bounds 1e-12 relative to the spectral radius; Q^T Q = I to 1e-13;
span(Q) = span(X) to 1e-12 relative; kept rank equal to the oracle's
(Householder + Jacobi SVD, src/core/Orthogonalization.cpp:717-775).
"""
import numpy as np
import pytest

from fasteig_amd import generators as G
from fasteig_amd.api import Problem
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _require_gpu(dev):
    yield


@pytest.fixture(scope="module")
def holder():
    """A problem whose executor runs the core calls (its matrix is unused by
    the QR / orthonormalization tests)."""
    p = Problem(G.laplacian_2d(8))
    yield p
    p.close()


@pytest.mark.parametrize("mat", ["lap2d", "powlaw", "h27", "c2"])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("seed", [1, 7])
def test_lanczos_bounds_match_oracle(mat, sign, seed):
    A = {"lap2d": lambda: G.laplacian_2d(40, 31),
         "powlaw": lambda: G.power_law(5000, seed=2, dmax=200),
         "h27": lambda: G.hamiltonian_27pt(12, seed=4, wells=6),
         "c2": lambda: G.conductivity_3d(14, seed=2, mass_seed=3)[0]}[mat]()
    lo_o, hi_o = orc.lanczos_bounds(A, sign, seed)
    p = Problem(A)
    lo, hi = p.lanczos_bounds(sign, seed)
    scale = max(abs(lo_o), abs(hi_o))
    assert abs(lo - lo_o) <= 1e-12 * scale, (lo, lo_o)
    assert abs(hi - hi_o) <= 1e-12 * scale, (hi, hi_o)
    # and they are bounds: the extreme eigenvalues of sign*A lie inside
    if A.shape[0] <= 3000:
        w = np.linalg.eigvalsh(sign * A.toarray())
        assert lo <= w[0] + 1e-8 * scale and hi >= w[-1] - 1e-8 * scale
    p.close()


def test_cholesky_qr_matches_oracle_and_fails_on_singular(holder):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((700, 24))
    Q, st = holder.thin_qr(X, "cholesky")
    info, Qo, Ro = orc.cholqr(X)
    assert st == "ok" and info == 0
    # Cholesky QR is unique (positive diagonal of R): same Q as the oracle
    assert np.max(np.abs(Q - Qo)) <= 1e-12
    assert np.max(np.abs(Q.T @ Q - np.eye(24))) <= 1e-12
    # singular Gram: the reference's -1 status, X unchanged
    Xs = X.copy()
    Xs[:, 7] = 0.0
    Q2, st2 = holder.thin_qr(Xs, "cholesky")
    assert st2 == "failed" and orc.cholqr(Xs)[0] == -1
    assert np.array_equal(Q2, Xs)


def graded_block(n, sigmas, seed, avoid=None):
    """U diag(sigmas) V^T with random orthonormal U (orthogonal to `avoid`)."""
    rng = np.random.default_rng(seed)
    U0 = rng.standard_normal((n, len(sigmas)))
    if avoid is not None:
        U0 -= avoid @ (avoid.T @ U0)
    U = np.linalg.qr(U0)[0]
    V = np.linalg.qr(rng.standard_normal((len(sigmas), len(sigmas))))[0]
    return (U * np.asarray(sigmas)) @ V.T


@pytest.mark.parametrize("kappa", [1e3, 1e9, 1e12])
def test_orthonormal_qr_shifted_cholqr3_on_ill_conditioned_block(holder, kappa):
    """The Householder role: orthonormal Q with span(Q) = span(X). Past
    kappa ~ 1e8 the plain Cholesky factorization of the Gram fails and the
    shifted path must run (and still give an orthonormal Q of the same span)."""
    n, k = 3000, 20
    X = graded_block(n, np.logspace(0, -np.log10(kappa), k), seed=int(np.log10(kappa)))
    Q, st = holder.thin_qr(X, "householder")
    assert st == ("shifted" if kappa >= 1e9 else "ok")
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) <= 1e-13
    # span(Q) contains X: the projection residual is at rounding level
    assert np.linalg.norm(X - Q @ (Q.T @ X)) <= 1e-12 * np.linalg.norm(X)
    # same subspace as the oracle's Householder QR, up to the perturbation bound:
    # a rounding-level change of X turns the direction of its smallest singular
    # value by about eps * kappa, for either method
    Qh, _ = orc.householder_qr(X)
    sin_max = np.linalg.norm(Q - Qh @ (Qh.T @ Q), 2)
    assert sin_max <= max(1e-12, 10 * np.finfo(float).eps * kappa), sin_max


def test_orthonormalize_block_rank_drop_matches_oracle(holder):
    """The oracle's rank-drop case: one source column in span(basis), one a
    sum of two others; two directions survive."""
    rng = np.random.default_rng(3)
    n = 200
    B, _ = np.linalg.qr(rng.standard_normal((n, 6)))
    S = rng.standard_normal((n, 4))
    S[:, 2] = B[:, 1] * 3.0
    S[:, 3] = S[:, 0] + S[:, 1]
    Qo = orc.orthonormalize_block(B, 6, S)
    Q = holder.orthonormalize_block(B, S)
    assert Q.shape[1] == Qo.shape[1] == 2
    assert np.max(np.abs(Q.T @ Q - np.eye(2))) <= 1e-13
    assert np.max(np.abs(B.T @ Q)) <= 1e-13
    # same kept subspace as the oracle
    s = np.linalg.svd(Qo.T @ Q, compute_uv=False)
    assert np.max(np.abs(s - 1.0)) <= 1e-10


@pytest.mark.parametrize("with_basis", [False, True])
def test_orthonormalize_block_graded_singular_values_keeps_oracle_rank(holder, with_basis):
    """Singular values from 1 down to 1e-15 (none within 30x of the 1e-12
    cutoff): the kept rank equals the oracle's, including the directions
    between 1e-7 and 1e-12 that a single Gram eigen-decomposition cannot
    resolve."""
    n = 1500
    sig = [1.0, 3e-2, 1e-3, 2e-5, 1e-7, 3e-9, 1e-10, 3e-14, 1e-15]
    rng = np.random.default_rng(10)
    if with_basis:
        # the part outside span(B) has exactly these singular values
        B = np.linalg.qr(rng.standard_normal((n, 5)))[0]
        S = graded_block(n, sig, seed=9, avoid=B) + 0.5 * B @ rng.standard_normal((5, len(sig)))
    else:
        B = None
        S = graded_block(n, sig, seed=9)
    Qo = orc.orthonormalize_block(B if with_basis else np.zeros((n, 1)), 5 if with_basis else 0, S)
    Q = holder.orthonormalize_block(B, S)
    assert Q.shape[1] == Qo.shape[1]
    assert Q.shape[1] == 7  # sigma > 1e-12 * max column norm (about 1)
    assert np.max(np.abs(Q.T @ Q - np.eye(Q.shape[1]))) <= 1e-12
    if with_basis:
        assert np.max(np.abs(B.T @ Q)) <= 1e-12
    # the same 7-dimensional subspace (its weakest direction, sigma 1e-10, is
    # fixed only to about eps / 1e-10 by either method)
    assert np.linalg.norm(Q - Qo @ (Qo.T @ Q), 2) <= 1e-5


def test_orthonormalize_block_nothing_survives(holder):
    rng = np.random.default_rng(4)
    n = 100
    B, _ = np.linalg.qr(rng.standard_normal((n, 5)))
    S = B @ rng.standard_normal((5, 3))
    assert orc.orthonormalize_block(B, 5, S).shape[1] == 0
    assert holder.orthonormalize_block(B, S).shape[1] == 0
    with pytest.raises(ValueError):
        holder.orthonormalize_block(B, S, rank_tol=1.5)

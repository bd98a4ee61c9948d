"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each oracle function is checked against something that fixes its answer
independently of the oracle's own formulas: numpy/LAPACK, closed-form spectra,
invariants, the C++ standard's required mt19937 output, and the reference's own
golden spectrum (tests/reference/SiO.out.121mpi, kept as tests/golden/sio_eigs.json).
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from numpy.polynomial import chebyshev as npcheb

from fasteig_amd import generators as G
from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def mt19937_uniform(seed, count):
    """libstdc++ uniform_real_distribution<double>(-1,1) over mt19937(seed):
    generate_canonical<double,53> draws two 32-bit words per value."""
    rs = np.random.RandomState(seed)  # init_genrand(seed), the std::mt19937 seeding
    raw = np.frombuffer(rs.bytes(8 * count), dtype=np.uint32).astype(np.float64)
    u = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return -1.0 + 2.0 * u


def test_mt19937_standard_value():
    # C++ [rand.predef]: the 10000th output of default-seeded mt19937 is 4123659995
    raw = np.frombuffer(np.random.RandomState(5489).bytes(4 * 10000), dtype=np.uint32)
    assert raw[-1] == 4123659995


@pytest.mark.parametrize("seed", [1, 2, 12345])
def test_random_block_column_major_fill(seed):
    n, k = 37, 5
    X = orc.random_block(n, k, seed)
    assert np.array_equal(X.ravel(order="F"), mt19937_uniform(seed, n * k))


@pytest.mark.parametrize("m", [1, 2, 7, 40, 120])
def test_syev_matches_lapack(m):
    rng = np.random.default_rng(m)
    A = rng.standard_normal((m, m))
    A = A + A.T
    w, V = orc.syev(A)
    wr = np.linalg.eigvalsh(A)
    s = max(1.0, np.abs(wr).max())
    assert np.max(np.abs(w - wr)) < 1e-13 * s * m
    assert np.max(np.abs(V.T @ V - np.eye(m))) < 1e-12 * m
    assert np.max(np.abs(A @ V - V * w)) < 1e-12 * s * m


def test_syev_wilkinson_and_degenerate():
    W = np.diag(np.abs(np.arange(-10, 11.0))) + np.diag(np.ones(20), 1) + np.diag(np.ones(20), -1)
    w, _ = orc.syev(W)
    assert np.max(np.abs(w - np.linalg.eigvalsh(W))) < 1e-13 * 10
    # W21+ has famously close top pairs: 10.746194182903393 (twice to 14 digits)
    assert abs(w[-1] - w[-2]) < 1e-13 and abs(w[-1] - 10.746194182903393) < 1e-12
    L = G.laplacian_2d(8).toarray()
    w, V = orc.syev(L)
    assert np.max(np.abs(w - G.laplacian_2d_eigenvalues(8))) < 1e-13
    assert np.max(np.abs(V.T @ V - np.eye(64))) < 1e-12


def test_householder_and_cholesky_qr_invariants():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((300, 17))
    Q, R = orc.householder_qr(X)
    assert np.max(np.abs(Q.T @ Q - np.eye(17))) < 1e-14 * 30
    assert np.max(np.abs(Q @ R - X)) < 1e-13
    assert np.allclose(R, np.triu(R))
    info, Q2, R2 = orc.cholqr(X)
    assert info == 0
    assert np.max(np.abs(Q2.T @ Q2 - np.eye(17))) < 1e-12
    assert np.max(np.abs(Q2 @ R2 - X)) < 1e-12
    assert np.all(np.diag(R2) > 0)
    # same column space as Householder: Q2 = Q D with D = diag(+-1)
    D = Q.T @ Q2
    assert np.max(np.abs(np.abs(D) - np.eye(17))) < 1e-10
    # singular Gram (zero column: pivot exactly 0) -> -1, X unchanged
    Xs = X.copy()
    Xs[:, 3] = 0.0
    info, Q3, _ = orc.cholqr(Xs)
    assert info == -1 and np.array_equal(Q3, Xs)


def test_jacobi_svd_matches_lapack():
    rng = np.random.default_rng(2)
    R = np.triu(rng.standard_normal((12, 12)))
    R[:, 5] = R[:, 4] * 1e-3
    s, U = orc.jacobi_svd(R)
    assert np.max(np.abs(s - np.linalg.svd(R, compute_uv=False))) < 1e-13
    assert np.max(np.abs(U.T @ U - np.eye(12))) < 1e-10


def test_orthonormalize_block_drops_dependent_directions():
    rng = np.random.default_rng(3)
    n = 200
    B, _ = np.linalg.qr(rng.standard_normal((n, 6)))
    S = rng.standard_normal((n, 4))
    S[:, 2] = B[:, 1] * 3.0          # lies in span(B): vanishes after projection
    S[:, 3] = S[:, 0] + S[:, 1]      # dependent on other source columns
    Q = orc.orthonormalize_block(B, 6, S)
    assert Q.shape[1] == 2
    assert np.max(np.abs(Q.T @ Q - np.eye(2))) < 1e-13
    assert np.max(np.abs(B.T @ Q)) < 1e-13


def test_residual_norms_match_numpy():
    """orc_residual_norms against numpy on the reference's definition
    (src/fasteig_solver.cpp:89-129): ||A x - l x|| / ((1 + |l|) ||x||), and
    ||A x - l x|| alone when that denominator is 0; unnormalized columns,
    negative and zero l, a zero column, and an exact eigenpair (residual at
    the rounding level). A dropped (1 + |l|), a missing ||x|| or a sign error
    in l x each fail one of these."""
    rng = np.random.default_rng(5)
    n = 60
    A = sp.random(n, n, density=0.1, random_state=7, format="csr")
    A = (A + A.T).tocsr()
    X = rng.standard_normal((n, 5)) * np.array([1.0, 3.0, 0.2, 1.0, 1.0])
    X[:, 3] = 0.0
    lam = np.array([0.7, -2.5, 0.0, 4.0, -1.0])
    w, V = np.linalg.eigh(A.toarray())
    X[:, 4] = 2.0 * V[:, 0]
    lam[4] = w[0]
    Ad = A.toarray()
    ref = np.empty(5)
    for j in range(5):
        num = np.linalg.norm(Ad @ X[:, j] - lam[j] * X[:, j])
        den = (1.0 + abs(lam[j])) * np.linalg.norm(X[:, j])
        ref[j] = num / den if den > 0 else num
    got = orc.residual_norms(A, lam, X)
    assert np.allclose(got[:4], ref[:4], rtol=1e-13, atol=0)
    assert got[3] == 0.0 and got[4] <= 1e-14


@pytest.mark.parametrize("n,ka,kb", [(1, 1, 1), (70, 5, 3), (5000, 37, 20), (70001, 16, 64)])
def test_dense_products_match_numpy(n, ka, kb):
    rng = np.random.default_rng(n + ka)
    X = rng.standard_normal((n, ka))
    Y = rng.standard_normal((n, kb))
    Cm = rng.standard_normal((ka, kb))
    assert np.max(np.abs(orc.gram(X, Y) - X.T @ Y)) <= 1e-12 * np.sqrt(n)
    assert np.max(np.abs(orc.right_multiply(X, Cm) - X @ Cm)) <= 1e-12 * np.sqrt(ka)


def test_spmm_matches_dense():
    A = G.power_law(400, seed=1, dmax=50)
    X = np.random.default_rng(4).standard_normal((400, 6))
    Y0 = np.random.default_rng(5).standard_normal((400, 6))
    Y = orc.spmm(A, X, 2.0, -0.5, Y0)
    assert np.max(np.abs(Y - (2.0 * (A.toarray() @ X) - 0.5 * Y0))) < 1e-12


@pytest.mark.parametrize("degree", [1, 2, 8, 20])
def test_chebyshev_filter_is_chebyshev_polynomial(degree):
    lam = np.linspace(-1.0, 5.0, 23)
    A = sp.diags(lam, format="csr")
    lower, upper = -1.0, 5.0
    Y = orc.cheb_filter(A, np.eye(23), degree, lower, upper)
    c, d = orc.cheb_scalars(degree, lower, upper)
    # closed-form scalars (reference ChebyshevSolver.cpp:232-236)
    xlb = np.cosh(np.arccosh(1e3) / degree)
    cut = (2 * lower + upper * (xlb - 1)) / (1 + xlb)
    assert np.isclose(c, 2 / (upper - cut)) and np.isclose(d, -(upper + cut) / (upper - cut))
    coef = np.zeros(degree + 1)
    coef[-1] = 1.0
    expect = npcheb.chebval(c * lam + d, coef)
    assert np.max(np.abs(np.diag(Y) - expect)) <= 1e-12 * max(1.0, np.abs(expect).max())
    # the lower bound maps to -xlb, so |T_d| = cosh(d acosh(xlb)) = 1e3 there
    assert np.isclose(c * lower + d, -xlb, rtol=1e-12)
    assert np.isclose(abs(npcheb.chebval(c * lower + d, coef)), 1e3, rtol=1e-9)


@pytest.mark.parametrize("degree", [1, 2, 7, 20])
def test_scaled_chebyshev_filter_is_normalised_polynomial(degree):
    """The adaptive-bound filter (Zhou & Saad 2007, Alg. 4.1) applied to a
    diagonal matrix must equal T_d((lam - c)/e) / T_d((a0 - c)/e) at every
    diagonal entry: 1 at a0, bounded by 1/|T_d((a0-c)/e)| on [cutoff, upper]."""
    lam = np.concatenate([np.linspace(-2.0, 0.9, 9), np.linspace(1.0, 6.0, 17)])
    A = sp.diags(lam, format="csr")
    a0, cutoff, upper = -2.0, 1.0, 6.0
    Y = orc.cheb_filter_scaled(A, np.eye(lam.size), degree, a0, cutoff, upper)
    e, c = 0.5 * (upper - cutoff), 0.5 * (upper + cutoff)
    coef = np.zeros(degree + 1)
    coef[-1] = 1.0
    expect = npcheb.chebval((lam - c) / e, coef) / npcheb.chebval((a0 - c) / e, coef)
    assert np.max(np.abs(np.diag(Y) - expect)) <= 1e-12 * max(1.0, np.abs(expect).max())
    # off-diagonal entries stay exactly zero (a diagonal operator)
    assert np.array_equal(Y - np.diag(np.diag(Y)), np.zeros_like(Y))
    assert np.isclose(Y[0, 0], 1.0, rtol=1e-12)
    damped = np.abs(np.diag(Y)[lam >= cutoff])
    assert damped.max() <= 1.0 / abs(npcheb.chebval((a0 - c) / e, coef)) * (1 + 1e-12)


def test_scaled_chebyshev_filter_general_matrix_matches_eigh():
    """On a non-diagonal symmetric matrix the filter is V p(Lam) V^T X."""
    A = G.hamiltonian_27pt(6, wells=3)
    w, V = np.linalg.eigh(A.toarray())
    X = np.random.default_rng(11).standard_normal((A.shape[0], 5))
    a0, cutoff, upper = w[0], w[40], w[-1] * 1.01
    d = 9
    e, c = 0.5 * (upper - cutoff), 0.5 * (upper + cutoff)
    coef = np.zeros(d + 1)
    coef[-1] = 1.0
    p = npcheb.chebval((w - c) / e, coef) / npcheb.chebval((a0 - c) / e, coef)
    ref = V @ (p[:, None] * (V.T @ X))
    Y = orc.cheb_filter_scaled(A, X, d, a0, cutoff, upper)
    assert np.max(np.abs(Y - ref)) <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("which", ["sr", "lr"])
def test_chebyshev_ritz_bounds_closed_form_laplacian(which):
    """The adaptive-bound Chebyshev mode against the 2D Laplacian closed form."""
    A = G.laplacian_2d(24, 19)
    exact = G.laplacian_2d_eigenvalues(24, 19)
    nev = 6
    r = orc.chebyshev(A, nev, degree=12, maxiter=3000, tol=1e-10, which=which, bounds="ritz")
    ref = exact[:nev] if which == "sr" else exact[::-1][:nev]
    assert np.max(np.abs(r["eigenvalues"] - ref) / np.abs(ref)) < 1e-10
    assert r["residuals"].max() <= 1e-10
    assert orc.residual_norms(A, r["eigenvalues"], r["eigenvectors"]).max() <= 1.0001e-10
    # the adaptive mode converges in fewer cycles than the fixed amplification cutoff
    r0 = orc.chebyshev(A, nev, degree=12, maxiter=3000, tol=1e-10, which=which)
    assert r["iterations"] < r0["iterations"]


@pytest.mark.parametrize("which", ["sr", "lr"])
def test_chebyshev_ritz_bounds_hamiltonian_matches_eigh(which):
    """A small 27-point Hamiltonian (the C3 family) in the adaptive mode against
    LAPACK: eigenvalues to 1e-10 relative and vectors by subspace angle."""
    A = G.hamiltonian_27pt(9, wells=5)
    w, V = np.linalg.eigh(A.toarray())
    nev = 16
    r = orc.chebyshev(A, nev, degree=10, maxiter=2000, tol=1e-11, which=which, bounds="ritz")
    ref = w[:nev] if which == "sr" else w[::-1][:nev]
    Vr = V[:, :nev] if which == "sr" else V[:, ::-1][:, :nev]
    scale = np.abs(w).max()
    assert np.max(np.abs(r["eigenvalues"] - ref)) <= 1e-10 * scale
    # sine of the largest principal angle between the computed and exact
    # nev-subspaces: ||(I - Vr Vr^T) X||_2 for the orthonormal X
    X = np.linalg.qr(r["eigenvectors"])[0]
    assert np.linalg.norm(X - Vr @ (Vr.T @ X), 2) <= 1e-8


@pytest.mark.parametrize("mat", ["lap2d", "powlaw", "h27"])
def test_lanczos_bounds_enclose_spectrum(mat):
    A = {"lap2d": G.laplacian_2d(20), "powlaw": G.power_law(600, seed=2, dmax=60),
         "h27": G.hamiltonian_27pt(8, wells=4)}[mat]
    w = np.linalg.eigvalsh(A.toarray())
    lo, hi = orc.lanczos_bounds(A, 1.0, 1)
    assert lo <= w[0] + 1e-8 * abs(w[0]) + 1e-10 and hi >= w[-1] - 1e-8 * abs(w[-1])
    lo2, hi2 = orc.lanczos_bounds(A, -1.0, 1)
    assert lo2 <= -w[-1] + 1e-8 * abs(w[-1]) and hi2 >= -w[0] - 1e-8


@pytest.mark.parametrize("solver", ["lobpcg", "chebyshev", "davidson"])
@pytest.mark.parametrize("which", ["sr", "lr"])
def test_solvers_closed_form_laplacian(solver, which):
    A = G.laplacian_2d(24, 19)
    exact = G.laplacian_2d_eigenvalues(24, 19)
    nev = 6
    fn = {"lobpcg": lambda: orc.lobpcg(A, nev, block=8, maxiter=3000, tol=1e-9, which=which),
          "chebyshev": lambda: orc.chebyshev(A, nev, degree=12, maxiter=3000, tol=1e-9, which=which),
          "davidson": lambda: orc.davidson(A, nev, maxblks=6, maxiter=3000, tol=1e-9, which=which)}[solver]
    r = fn()
    ref = exact[:nev] if which == "sr" else exact[::-1][:nev]
    assert np.max(np.abs(r["eigenvalues"] - ref) / np.abs(ref)) < 1e-10
    assert r["residuals"].max() <= 1e-9
    rn = orc.residual_norms(A, r["eigenvalues"], r["eigenvectors"])
    assert rn.max() <= 1.0001e-9


def load_sio():
    d = np.load(os.path.join(GOLD, "SiO.npz"))
    n = int(d["n"])
    T = sp.coo_matrix((d["vals"], (d["rows"], d["cols"])), shape=(n, n)).tocsr()
    full = T + T.T - sp.diags(T.diagonal())
    return G._finish(full)


def test_sio_golden_spectrum_lobpcg():
    """The reference's CTest case reference_cpu_SiO_LOBPCG_SR (k=5, tol 1e-8,
    seed 1) against the ScaLAPACK spectrum, with its acceptance rule."""
    A = load_sio()
    ref = json.load(open(os.path.join(GOLD, "sio_eigs.json")))["eigenvalues"][:5]
    r = orc.lobpcg(A, 5, maxiter=2000, tol=1e-8, which="sr", seed=1, want_vectors=True)
    ev = r["eigenvalues"]
    assert np.all(np.abs(ev - ref) <= 1e-10 + 1e-10 * np.abs(ref))
    rn = orc.residual_norms(A, ev, r["eigenvectors"])
    assert rn.max() <= 1e-8


def test_sio_golden_spectrum_block_krylov_schur():
    """The reference's CTest case reference_cpu_SiO_BlockKrylovSchur_SR (k=5,
    block 10, maxdim 40, tol 1e-8, seed 1; tests/scripts/run_case.py defaults)
    against the ScaLAPACK spectrum."""
    A = load_sio()
    ref = json.load(open(os.path.join(GOLD, "sio_eigs.json")))["eigenvalues"][:5]
    r = orc.block_krylov_schur(A, 5, block=10, maxdim=40, maxiter=2000, tol=1e-8, which="sr", seed=1)
    assert r["status"] == 0
    assert np.all(np.abs(r["eigenvalues"] - ref) <= 1e-10 + 1e-10 * np.abs(ref))
    assert orc.residual_norms(A, r["eigenvalues"], r["eigenvectors"]).max() <= 1e-8


@pytest.mark.parametrize("which", ["sr", "lr"])
def test_block_krylov_schur_closed_form(which):
    A = G.laplacian_2d(24, 20)
    ex = np.sort(G.laplacian_2d_eigenvalues(24, 20))
    ex = ex[:6] if which == "sr" else ex[::-1][:6]
    r = orc.block_krylov_schur(A, 6, block=4, maxdim=32, maxiter=3000, tol=1e-10, which=which, seed=3)
    assert r["status"] == 0
    assert np.all(np.abs(r["eigenvalues"] - ex) <= 1e-9 * np.abs(ex))


def test_block_krylov_schur_full_dimension_partial_block():
    """maxdim >= n: the basis covers the whole space with a partial last block
    (n = 25, block 4) and the Ritz values are the exact spectrum."""
    rng = np.random.default_rng(8)
    M = rng.standard_normal((25, 25))
    A = sp.csr_matrix(M + M.T)
    w = np.linalg.eigvalsh(A.toarray())
    r = orc.block_krylov_schur(A, 5, block=4, maxdim=40, maxiter=5, tol=1e-10, which="sr", seed=1)
    assert r["iterations"] == 1
    assert np.all(np.abs(r["eigenvalues"] - w[:5]) <= 1e-11 * np.abs(w).max())
    assert orc.residual_norms(A, r["eigenvalues"], r["eigenvectors"]).max() <= 1e-12


def test_block_krylov_schur_invariant_start_replenishes():
    """A = diag(1 x 10, 2 x 10, 3 x 10): the Krylov space of a 3-column start
    block has dimension at most 9 (3 directions per distinct eigenvalue), so
    the block QR must complete the basis with seeded random directions to find
    four copies of the eigenvalue 1."""
    d = np.r_[np.ones(10), 2.0 * np.ones(10), 3.0 * np.ones(10)]
    A = sp.diags(d).tocsr()
    r = orc.block_krylov_schur(A, 4, block=3, maxdim=15, maxiter=200, tol=1e-10, which="sr", seed=2)
    assert r["status"] == 0
    assert np.all(np.abs(r["eigenvalues"] - 1.0) <= 1e-12)
    X = r["eigenvectors"]
    assert np.max(np.abs(X.T @ X - np.eye(4))) <= 1e-10


def test_sio_golden_spectrum_krylov_schur():
    """The reference's CTest case reference_cpu_SiO_KrylovSchur_SR (k=5, maxdim
    40, tol 1e-8, seed 1; tests/scripts/run_case.py defaults) against the
    ScaLAPACK spectrum."""
    A = load_sio()
    ref = json.load(open(os.path.join(GOLD, "sio_eigs.json")))["eigenvalues"][:5]
    r = orc.krylov_schur(A, 5, maxdim=40, maxiter=2000, tol=1e-8, which="sr", seed=1)
    assert r["status"] == 0
    assert np.all(np.abs(r["eigenvalues"] - ref) <= 1e-10 + 1e-10 * np.abs(ref))
    assert orc.residual_norms(A, r["eigenvalues"], r["eigenvectors"]).max() <= 1e-8


def test_sio_golden_spectrum_lanczos():
    """reference_cpu_SiO_Lanczos_SR: maxiter min(2000, n), k=5, tol 1e-8."""
    A = load_sio()
    ref = json.load(open(os.path.join(GOLD, "sio_eigs.json")))["eigenvalues"][:5]
    r = orc.lanczos(A, 5, maxiter=min(2000, A.shape[0]), tol=1e-8, which="sr", seed=1)
    assert r["status"] == 0 and r["iterations"] % 5 == 0
    assert np.all(np.abs(r["eigenvalues"] - ref) <= 1e-10 + 1e-10 * np.abs(ref))
    assert orc.residual_norms(A, r["eigenvalues"], r["eigenvectors"]).max() <= 1e-8


@pytest.mark.parametrize("which", ["sr", "lr", "be"])
def test_scalar_krylov_closed_form(which):
    """30 x 17 Laplacian (simple spectrum): both scalar solvers return the
    closed-form ends in the reference's order (BE alternates largest,
    smallest)."""
    A = G.laplacian_2d(30, 17)
    ex = np.sort(G.laplacian_2d_eigenvalues(30, 17))
    want = {"sr": ex[:6], "lr": ex[::-1][:6], "be": ex[[-1, 0, -2, 1, -3, 2]]}[which]
    for r in (orc.lanczos(A, 6, maxiter=300, tol=1e-9, which=which, seed=1),
              orc.krylov_schur(A, 6, maxdim=30, maxiter=300, tol=1e-9, which=which, seed=1)):
        assert r["status"] == 0
        assert np.all(np.abs(r["eigenvalues"] - want) <= 1e-9 * np.abs(want))


def test_lanczos_full_dimension_is_exact():
    """m = n steps: the Lanczos basis spans the space and the Ritz values are
    the eigenvalues (dense eigh)."""
    A = G.laplacian_2d(7, 5)
    w = np.linalg.eigvalsh(A.toarray())
    r = orc.lanczos(A, 5, maxiter=35, tol=1e-300, which="sr", seed=1)
    assert r["iterations"] == 35
    assert np.all(np.abs(r["eigenvalues"] - w[:5]) <= 1e-12 * np.abs(w).max())


def test_scalar_krylov_invariant_subspace_restart():
    """diag(1 x 10, 2 x 10, 3 x 10): one start vector spans a 3-dimensional
    Krylov space, so the chain breaks down and continues from seeded random
    vectors. Krylov-Schur finds four copies of 1 with orthonormal vectors;
    Lanczos stops at its first check (m = 10) with the values the broken
    chain holds, as the reference's scalar recurrence does."""
    A = sp.diags(np.r_[np.ones(10), 2.0 * np.ones(10), 3.0 * np.ones(10)]).tocsr()
    r = orc.krylov_schur(A, 4, maxdim=12, maxiter=200, tol=1e-10, which="sr", seed=2)
    assert r["status"] == 0 and np.all(np.abs(r["eigenvalues"] - 1.0) <= 1e-12)
    X = r["eigenvectors"]
    assert np.max(np.abs(X.T @ X - np.eye(4))) <= 1e-10
    r = orc.lanczos(A, 4, maxiter=30, tol=1e-10, which="sr", seed=2)
    assert r["iterations"] == 10
    assert np.allclose(r["eigenvalues"], [1.0, 1.0, 1.0, 2.0], atol=1e-12)


def test_scalar_krylov_rejects_bad_arguments():
    A = G.laplacian_2d(6)
    with pytest.raises(ValueError, match="maxiter must be >= nev"):
        orc.lanczos(A, 4, maxiter=3)
    with pytest.raises(ValueError, match="maxiter must be <= n"):
        orc.lanczos(A, 4, maxiter=37)
    with pytest.raises(ValueError, match="invalid which"):
        orc.krylov_schur(A, 4, which="xx")


@pytest.mark.parametrize("which", ["sr", "lr", "be"])
def test_block_lanczos_closed_form(which):
    A = G.laplacian_2d(30, 17)
    ex = np.sort(G.laplacian_2d_eigenvalues(30, 17))
    want = {"sr": ex[:6], "lr": ex[::-1][:6], "be": ex[[-1, 0, -2, 1, -3, 2]]}[which]
    r = orc.block_lanczos(A, 6, block=4, maxiter=200, tol=1e-9, which=which, seed=1)
    assert r["status"] == 0
    assert np.all(np.abs(r["eigenvalues"] - want) <= 1e-9 * np.abs(want))
    assert r["residuals"].max() <= 1e-8


def test_block_lanczos_full_space_is_exact():
    """35 = 8 blocks of 4 + a partial block: after 9 blocks the basis spans the
    space and the Ritz values are the eigenvalues."""
    A = G.laplacian_2d(7, 5)
    w = np.linalg.eigvalsh(A.toarray())
    r = orc.block_lanczos(A, 5, block=4, maxiter=100, tol=1e-300, which="sr", seed=1)
    assert r["iterations"] == 9
    assert np.all(np.abs(r["eigenvalues"] - w[:5]) <= 1e-12 * np.abs(w).max())


def test_block_lanczos_narrow_block_misses_multiplicity():
    """diag(1 x 10, 2 x 10, 3 x 10) with a 3-column block: the block Krylov
    space closes after three blocks holding three copies of each eigenvalue,
    so the converged answer is (1, 1, 1, 2) -- the limitation the reference
    documents for blocks narrower than nev (src/BlockLanczosSolver.cpp:129-131)."""
    A = sp.diags(np.r_[np.ones(10), 2.0 * np.ones(10), 3.0 * np.ones(10)]).tocsr()
    r = orc.block_lanczos(A, 4, block=3, maxiter=50, tol=1e-10, which="sr", seed=2)
    assert r["iterations"] == 3
    assert np.allclose(r["eigenvalues"], [1.0, 1.0, 1.0, 2.0], atol=1e-12)
    assert r["residuals"].max() <= 1e-12

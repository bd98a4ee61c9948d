"""Seeded generators for the benchmark matrices (full symmetric CSR, int32 indices).

Every generator returns a ``scipy.sparse.csr_matrix`` with sorted column
indices, both triangles stored, float64 values. The workloads:

* ``laplacian_2d(nx, ny)``: 5-point Dirichlet Laplacian, diag 4, off-diag -1.
  Closed-form spectrum ``4 - 2cos(i*pi/(nx+1)) - 2cos(j*pi/(ny+1))``.
* ``conductivity_3d(g, seed)``: 7-point stencil on a g^3 grid with per-edge
  conductivities U[0.5, 2.0]; K_ii is the sum of the conductivities of the six
  incident edges, including the edges to the Dirichlet boundary. With
  ``mass_seed`` also returns the diagonal mass M ~ U[0.5, 1.5].
* ``hamiltonian_27pt(g, seed)``: H = -0.5 * L27 + V on a g^3 grid, where L27 is
  the isotropic 27-point Laplacian (weights centre -128, faces 14, edges 3,
  corners 1, all over 30; Dirichlet) and V_i = 2*N(0,1) minus 200 Gaussian wells
  of depth 5 and width 2 grid points at seeded centres.
* ``power_law(n, seed)``: random symmetric SPD matrix whose row degrees follow a
  truncated discrete power law (alpha 2.1 on [4, 2000]); N(0,1) values,
  symmetrized as (A + A^T)/2 over the pattern union, diagonal = row abs-sum + 1.
* ``stencil_with_potential(kind, n_target, seed)``: 2D 5-point or 3D 7-point
  Laplacian plus an N(0,1) diagonal (serving requests).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _finish(A) -> sp.csr_matrix:
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A


def laplacian_1d(n: int) -> sp.csr_matrix:
    return sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")


def laplacian_2d(nx: int, ny: int | None = None) -> sp.csr_matrix:
    ny = nx if ny is None else ny
    A = sp.kron(sp.identity(ny), laplacian_1d(nx)) + sp.kron(laplacian_1d(ny), sp.identity(nx))
    return _finish(A)


def laplacian_2d_eigenvalues(nx: int, ny: int | None = None) -> np.ndarray:
    ny = nx if ny is None else ny
    i = np.arange(1, nx + 1)[:, None]
    j = np.arange(1, ny + 1)[None, :]
    lam = 4.0 - 2.0 * np.cos(i * np.pi / (nx + 1)) - 2.0 * np.cos(j * np.pi / (ny + 1))
    return np.sort(lam.ravel())


def laplacian_3d(g: int) -> sp.csr_matrix:
    I = sp.identity(g)
    L = laplacian_1d(g)
    A = sp.kron(sp.kron(I, I), L) + sp.kron(sp.kron(I, L), I) + sp.kron(sp.kron(L, I), I)
    return _finish(A)


def conductivity_3d(g: int, seed: int = 2, mass_seed: int | None = None):
    """7-point SPD stencil with random per-edge conductivities (Dirichlet)."""
    rng = np.random.default_rng(seed)
    n = g ** 3
    idx = np.arange(n).reshape(g, g, g)  # idx[z, y, x]
    diag = np.zeros(n)
    rows, cols, vals = [], [], []
    for axis in range(3):
        # g+1 edge layers along `axis`: layer 0 and g touch the boundary.
        shape = [g, g, g]
        shape[axis] = g + 1
        cond = rng.uniform(0.5, 2.0, size=shape)
        # every node gets the conductivity of its two edges along this axis
        lo = np.take(cond, np.arange(g), axis=axis)
        hi = np.take(cond, np.arange(1, g + 1), axis=axis)
        diag += (lo + hi).ravel()
        # interior edges couple neighbours
        inner = np.take(cond, np.arange(1, g), axis=axis)
        a = np.take(idx, np.arange(g - 1), axis=axis).ravel()
        b = np.take(idx, np.arange(1, g), axis=axis).ravel()
        w = inner.ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [-w, -w]
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    K = _finish(sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                              shape=(n, n)))
    if mass_seed is None:
        return K
    M = np.random.default_rng(mass_seed).uniform(0.5, 1.5, size=n)
    return K, M


def _stencil_27(g: int, weights) -> sp.csr_matrix:
    """Assemble a 27-point constant-coefficient stencil (Dirichlet) from
    weights[dz+1][dy+1][dx+1]."""
    n = g ** 3
    idx = np.arange(n).reshape(g, g, g)
    rows, cols, vals = [], [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                w = weights[dz + 1][dy + 1][dx + 1]
                zs = slice(max(0, -dz), g - max(0, dz))
                ys = slice(max(0, -dy), g - max(0, dy))
                xs = slice(max(0, -dx), g - max(0, dx))
                src = idx[zs, ys, xs]
                dst = idx[slice(zs.start + dz, zs.stop + dz), slice(ys.start + dy, ys.stop + dy),
                          slice(xs.start + dx, xs.stop + dx)]
                rows.append(src.ravel())
                cols.append(dst.ravel())
                vals.append(np.full(src.size, w))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr)
    A = sp.csr_matrix((v, c.astype(np.int32), indptr.astype(np.int32)), shape=(n, n))
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A


def isotropic_27pt_weights():
    """Isotropic 27-point Laplacian weights (h = 1): centre -128/30, faces 14/30,
    edges 3/30, corners 1/30 (row sums zero)."""
    W = np.zeros((3, 3, 3))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dz) + abs(dy) + abs(dx)
                W[dz + 1, dy + 1, dx + 1] = {0: -128.0, 1: 14.0, 2: 3.0, 3: 1.0}[k] / 30.0
    return W


def hamiltonian_27pt(g: int = 96, seed: int = 4, wells: int = 200, depth: float = 5.0,
                     width: float = 2.0) -> sp.csr_matrix:
    rng = np.random.default_rng(seed)
    n = g ** 3
    H = _stencil_27(g, -0.5 * isotropic_27pt_weights())
    V = 2.0 * rng.standard_normal(n)
    centres = rng.uniform(0, g - 1, size=(wells, 3))
    Vg = V.reshape(g, g, g)
    reach = int(np.ceil(5 * width))
    for cz, cy, cx in centres:
        z0, z1 = max(0, int(cz) - reach), min(g, int(cz) + reach + 2)
        y0, y1 = max(0, int(cy) - reach), min(g, int(cy) + reach + 2)
        x0, x1 = max(0, int(cx) - reach), min(g, int(cx) + reach + 2)
        z = np.arange(z0, z1)[:, None, None]
        y = np.arange(y0, y1)[None, :, None]
        x = np.arange(x0, x1)[None, None, :]
        r2 = (z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2
        Vg[z0:z1, y0:y1, x0:x1] -= depth * np.exp(-r2 / (2.0 * width * width))
    # add the potential on the (already present) diagonal
    rp, ci = H.indptr, H.indices
    row_of = np.repeat(np.arange(n), np.diff(rp))
    diag_pos = np.nonzero(ci == row_of)[0]
    H.data[diag_pos] += V[row_of[diag_pos]]
    return H


def power_law(n: int = 2_000_000, seed: int = 5, alpha: float = 2.1, dmin: int = 4,
              dmax: int = 2000) -> sp.csr_matrix:
    rng = np.random.default_rng(seed)
    support = np.arange(dmin, dmax + 1, dtype=np.float64)
    pmf = support ** (-alpha)
    pmf /= pmf.sum()
    deg = rng.choice(np.arange(dmin, dmax + 1), size=n, p=pmf)
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    cols = rng.integers(0, n, size=rows.size, dtype=np.int64)
    vals = rng.standard_normal(rows.size)
    keep = rows != cols
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A = (A + A.T) * 0.5
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    absrow = np.asarray(abs(A).sum(axis=1)).ravel()
    A = A + sp.diags(absrow + 1.0)
    return _finish(A)


def stencil_with_potential(kind: str, n_target: int, seed: int) -> sp.csr_matrix:
    """2D 5-point ('2d') or 3D 7-point ('3d') Laplacian of about n_target rows
    plus an N(0,1) diagonal potential."""
    rng = np.random.default_rng(seed)
    if kind == "2d":
        g = max(4, int(round(np.sqrt(n_target))))
        L = laplacian_2d(g)
    else:
        g = max(4, int(round(n_target ** (1.0 / 3.0))))
        L = laplacian_3d(g)
    V = rng.standard_normal(L.shape[0])
    return _finish(L + sp.diags(V))


def congruence(K: sp.csr_matrix, mass: np.ndarray) -> sp.csr_matrix:
    """M^{-1/2} K M^{-1/2} for a diagonal M (K u = lambda M u  <=>  A y = lambda y,
    u = M^{-1/2} y)."""
    d = 1.0 / np.sqrt(mass)
    A = sp.diags(d) @ K @ sp.diags(d)
    return _finish(A)


def serving_stream(count: int = 512, seed: int = 11, rate: float = 200.0, n_min: float = 1e3,
                   n_max: float = 2e5):
    """Request stream of the serving workload: ``count`` requests with Poisson
    arrivals at ``rate`` per second; each a 2D 5-point or 3D 7-point Laplacian
    plus an N(0,1) diagonal with n log-uniform in [n_min, n_max], nev in
    {8, 16, 32}, tol in {1e-6, 1e-8}. Returns dicts (id, arrival, kind, n_target,
    nev, tol, seed); build the matrix with ``stencil_with_potential``."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count)) if rate > 0 else np.zeros(count)
    out = []
    for i in range(count):
        out.append(dict(id=i, arrival=float(t[i]), kind="2d" if rng.random() < 0.5 else "3d",
                        n_target=int(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))),
                        nev=int(rng.choice([8, 16, 32])), tol=float(rng.choice([1e-6, 1e-8])),
                        seed=int(rng.integers(1 << 30))))
    return out

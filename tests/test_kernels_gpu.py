"""Kernel-level parity of the HIP library (C ABI) against the CPU oracle.

Every test calls libfe_hip.so through its C ABI (fasteig_amd._fe) on seeded
inputs and compares with the oracle (oracle/oracle.cpp) or with numpy closed
forms. Tolerances: FP64 kernels 1e-13 relative (different summation order
only); FP32 kernels 1e-5 relative to the operand scale; compensated BF16
products 2^-14 (BF16x3) and twice the native FP32 error (BF16x6).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from fasteig_amd import _fe
from fasteig_amd import generators as G
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
L = _fe.lib


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def mats():
    rng = np.random.default_rng(7)
    out = {
        "lap2d": G.laplacian_2d(37, 29),
        "cond3d": G.conductivity_3d(11, seed=2),
        "h27": G.hamiltonian_27pt(9, seed=4, wells=5),
        "h27big": G.hamiltonian_27pt(23, seed=4, wells=9),
        "powlaw": G.power_law(3000, seed=5, dmax=300),
    }
    # ragged: empty rows and one dense row
    R = sp.random(500, 500, density=0.01, random_state=3, format="csr")
    R = R + R.T
    R = sp.lil_matrix(R)
    R[17, :] = 0.0
    R[:, 17] = 0.0
    R[3, :] = rng.standard_normal(500)
    R[:, 3] = R[3, :].T.toarray()
    out["ragged"] = G._finish(sp.csr_matrix(R))
    return out


@pytest.mark.parametrize("name", ["lap2d", "cond3d", "h27", "powlaw", "ragged"])
@pytest.mark.parametrize("k", [1, 3, 8, 12, 16, 48, 96, 256, 512])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_spmm_fused(dev, mats, name, k, prec):
    A = mats[name]
    n = A.shape[0]
    rng = np.random.default_rng(k + 13 * prec)
    X = rng.uniform(-1, 1, (n, k))
    Z = rng.uniform(-1, 1, (n, k))
    alpha, beta, gamma = 1.7, -0.3, -1.0
    ref = orc.spmm(A, X, alpha, 0.0) + beta * X + gamma * Z
    dt = np.float64 if prec == _fe.F64 else np.float32
    Xd = dev.put(X, dt)
    Yd = dev.put(Z, dt)  # Z aliases Y: read before written
    C = dev.csr(A)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Yd.ptr, k, alpha, beta, Yd.ptr, k, gamma))
    got = Yd.download().astype(np.float64)
    scale = np.max(np.abs(A).sum(axis=1)) + 2.0
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    assert np.max(np.abs(got - ref)) <= tol * scale


@pytest.mark.parametrize("k", [4, 16, 48, 64])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_spmm_sell_matches_csr(dev, mats, k, prec):
    A = mats["powlaw"]
    n = A.shape[0]
    X = np.random.default_rng(1).uniform(-1, 1, (n, k))
    dt = np.float64 if prec == _fe.F64 else np.float32
    C = dev.csr(A)
    Xd = dev.put(X, dt)
    Y1 = dev.buffer((n, k), dt)
    Y2 = dev.buffer((n, k), dt)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y1.ptr, k, 1.0, 0.0, None, 0, 0.0))
    C.build_sell(32, 256)
    C.use(_fe.FMT_SELL)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y2.ptr, k, 1.0, 0.0, None, 0, 0.0))
    ref = orc.spmm(A, X)
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    scale = np.max(np.abs(A).sum(axis=1))
    assert np.max(np.abs(Y1.download() - ref)) <= tol * scale
    assert np.max(np.abs(Y2.download() - ref)) <= tol * scale


@pytest.mark.parametrize("name", ["h27", "h27big"])
@pytest.mark.parametrize("rb", [4, 8])
@pytest.mark.parametrize("k", [4, 12, 32, 48, 100, 256, 512])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_spmm_blocked_matches_csr(dev, mats, name, rb, k, prec):
    """Row-block copy (BCSR rb x 1 with explicit zeros): the same sums in the
    same ascending-column order as the CSR kernel (bitwise), fused epilogue
    included; k not a multiple of the 16-byte run falls back to CSR."""
    A = mats[name]
    n = A.shape[0]
    rng = np.random.default_rng(31 + k)
    X = rng.uniform(-1, 1, (n, k))
    Z = rng.uniform(-1, 1, (n, k))
    dt = np.float64 if prec == _fe.F64 else np.float32
    C = dev.csr(A)
    built, reuse = C.build_blocked(rb)
    assert built and reuse >= 1.5
    Xd = dev.put(X, dt)
    Zd = dev.put(Z, dt)
    Y1 = dev.buffer((n, k), dt)
    Y2 = dev.buffer((n, k), dt)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y1.ptr, k, 0.9, 0.4, Zd.ptr, k, -1.2))
    C.use(_fe.FMT_BLOCKED)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y2.ptr, k, 0.9, 0.4, Zd.ptr, k, -1.2))
    ref = orc.spmm(A, X, 0.9, 0.0) + 0.4 * X - 1.2 * Z
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    scale = np.max(np.abs(A).sum(axis=1)) + 2.0
    assert np.max(np.abs(Y2.download().astype(np.float64) - ref)) <= tol * scale
    assert np.array_equal(Y1.download(), Y2.download())


def test_spmm_blocked_ragged_tail(dev):
    """n not a multiple of the row block, a row with no entries in the last
    block, and a strided (ld > k) aligned block."""
    A = G.hamiltonian_27pt(7, seed=1, wells=3).tolil()
    n = A.shape[0]  # 343 = 4 * 85 + 3
    A[n - 2, :] = 0.0
    A[:, n - 2] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    ld = 48
    X = np.random.default_rng(5).uniform(-1, 1, (n, ld))
    C = dev.csr(A)
    built, _ = C.build_blocked(4)
    assert built
    for k in (40, 37, 3):  # 37, 3: partial last 16-byte run
        for prec, dt, tol in ((_fe.F64, np.float64, 1e-13), (_fe.F32, np.float32, 2e-6)):
            C.use(_fe.FMT_CSR)
            Xd = dev.put(X, dt)
            Y1 = dev.put(np.full((n, ld), 7.0), dt)
            Y2 = dev.put(np.full((n, ld), 7.0), dt)
            _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, ld, Y1.ptr, ld, 1.0, 0.5, Y1.ptr, ld, 0.25))
            C.use(_fe.FMT_BLOCKED)
            _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, ld, Y2.ptr, ld, 1.0, 0.5, Y2.ptr, ld, 0.25))
            got = Y2.download().astype(np.float64)
            Xr = X.astype(dt).astype(np.float64)
            ref = orc.spmm(A, Xr[:, :k]) + 0.5 * Xr[:, :k] + 0.25 * 7.0
            assert np.max(np.abs(got[:, :k] - ref)) <= tol * (np.max(np.abs(A).sum(axis=1)) + 2)
            assert np.all(got[:, k:] == 7.0)
            assert np.array_equal(Y1.download(), Y2.download())


@pytest.mark.parametrize("k", [4, 48, 256, 300, 512])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_spmm_block_schedule_is_bitwise(dev, k, prec):
    """Strip-by-strip row-block order (24^3 grid: line 24, plane 576): the
    automatic order, an explicit 2-line strip and no order give bitwise the
    same products (each row's sum is unchanged)."""
    A = G.hamiltonian_27pt(24, seed=4, wells=9)
    n = A.shape[0]
    rng = np.random.default_rng(k)
    dt = np.float64 if prec == _fe.F64 else np.float32
    Xd = dev.put(rng.uniform(-1, 1, (n, k)), dt)
    Zd = dev.put(rng.uniform(-1, 1, (n, k)), dt)
    C = dev.csr(A)
    assert C.build_blocked(4)[0]
    C.use(_fe.FMT_BLOCKED)
    outs = []
    for plane, strip in ((0, 0), (-1, 0), (576, 48), (576, 576), (0, 0)):
        on = C.block_schedule(plane, strip)
        assert on == (plane != 0)
        Y = dev.buffer((n, k), dt)
        _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y.ptr, k, 0.7, 0.3, Zd.ptr, k, -1.1))
        outs.append(Y.download())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_block_schedule_detection(dev, mats):
    """Only 3D grid stencils whose line is a multiple of the row block get an
    order; a 2D 9-point stencil and a 23-point line do not; bad explicit
    arguments are rejected."""
    C = dev.csr(mats["h27big"])  # 23^3: line 23
    assert C.build_blocked(4)[0]
    assert C.block_schedule() is False
    T = sp.diags([1.0, 1.0, 1.0], [-1, 0, 1], shape=(64, 64))
    C2 = dev.csr(G._finish(sp.kron(T, T).tocsr()))
    assert C2.build_blocked(4)[0]
    assert C2.block_schedule() is False
    with pytest.raises(_fe.FeError):
        C2.block_schedule(530, 8)  # plane not a multiple of the row block
    C3 = dev.csr(mats["h27"])
    with pytest.raises(_fe.FeError):
        C3.block_schedule()  # no row-block copy yet


TILED = {
    # 4 x 4 plane patches, pencils through the planes (line a multiple of 4)
    "p12": lambda: G.hamiltonian_27pt(12, seed=4, wells=5),
    "p20": lambda: G.hamiltonian_27pt(20, seed=6, wells=7),  # 20 planes: pencils of 16 and 4
    "c12": lambda: G.conductivity_3d(12, seed=3),            # 7-point
    # 8 x 4 plane patches, 32-row tiles of 512 threads (line a multiple of 8)
    "w16": lambda: G.hamiltonian_27pt(16, seed=5, wells=6),
    "w24c": lambda: G.conductivity_3d(24, seed=4),           # 7-point, pencils of 24 planes
    # 16 consecutive rows per tile, pencils of 32 tiles (line not a multiple of 4)
    "h27": lambda: G.hamiltonian_27pt(9, seed=4, wells=5),   # 729 rows: a padded last tile
    "h27big": lambda: G.hamiltonian_27pt(23, seed=4, wells=9),
    "c11": lambda: G.conductivity_3d(11, seed=2),
}


@pytest.fixture(scope="module")
def tiled_mats():
    return {name: f() for name, f in TILED.items()}


@pytest.mark.parametrize("name", list(TILED))
@pytest.mark.parametrize("k", [4, 16, 48, 64, 100, 256, 512])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_spmm_tiled_matches_csr(dev, tiled_mats, name, k, prec):
    """LDS-tiled copy: each row's sum over its CSR entries in CSR order, so the
    product is bitwise the CSR kernel's, fused epilogue included; several
    256-byte column slices and a partial last slice (k = 100, 48 in FP64)."""
    A = tiled_mats[name]
    n = A.shape[0]
    rng = np.random.default_rng(41 + k)
    X = rng.uniform(-1, 1, (n, k))
    Z = rng.uniform(-1, 1, (n, k))
    dt = np.float64 if prec == _fe.F64 else np.float32
    C = dev.csr(A)
    built, lpr = C.build_tiled()
    assert built and lpr <= 0.5 * A.nnz / n
    Xd = dev.put(X, dt)
    Zd = dev.put(Z, dt)
    Y1 = dev.buffer((n, k), dt)
    Y2 = dev.buffer((n, k), dt)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y1.ptr, k, 0.9, 0.4, Zd.ptr, k, -1.2))
    C.use(_fe.FMT_TILED)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y2.ptr, k, 0.9, 0.4, Zd.ptr, k, -1.2))
    ref = orc.spmm(A, X, 0.9, 0.0) + 0.4 * X - 1.2 * Z
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    scale = np.max(np.abs(A).sum(axis=1)) + 2.0
    assert np.max(np.abs(Y2.download().astype(np.float64) - ref)) <= tol * scale
    assert np.array_equal(Y1.download(), Y2.download())


@pytest.mark.parametrize("k,prec", [(256, _fe.F32), (128, _fe.F64)])
def test_spmm_tiled_larger_grid(dev, k, prec):
    """48^3 grid (110,592 rows): two pencils of 24 planes per patch column,
    144 pencils spread over the XCDs, several column slices."""
    A = G.hamiltonian_27pt(48, seed=7, wells=20)
    n = A.shape[0]
    rng = np.random.default_rng(k)
    dt = np.float64 if prec == _fe.F64 else np.float32
    X = rng.uniform(-1, 1, (n, k))
    Z = rng.uniform(-1, 1, (n, k))
    C = dev.csr(A)
    built, lpr = C.build_tiled()
    assert built and lpr < 3.0
    Xd, Zd = dev.put(X, dt), dev.put(Z, dt)
    Y1, Y2 = dev.buffer((n, k), dt), dev.buffer((n, k), dt)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y1.ptr, k, 0.3, -0.2, Zd.ptr, k, 0.7))
    C.use(_fe.FMT_TILED)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y2.ptr, k, 0.3, -0.2, Zd.ptr, k, 0.7))
    got = Y2.download()
    assert np.array_equal(Y1.download(), got)
    rows = rng.choice(n, 2000, replace=False)  # oracle on a sample of rows
    Xr, Zr = X.astype(dt).astype(np.float64), Z.astype(dt).astype(np.float64)
    ref = A[rows] @ Xr * 0.3 - 0.2 * Xr[rows] + 0.7 * Zr[rows]
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    assert np.max(np.abs(got[rows].astype(np.float64) - ref)) <= tol * (np.max(np.abs(A).sum(axis=1)) + 2)


def _replay_tile_loads(A, line, plane, ex=4, ey=4, pz=24):
    """X rows a pencil walk loads, replayed from the planning rule: tiles are
    ex x ey patches of a plane, pencils run through pz planes, and a tile loads
    the columns of its rows that the previous tile of its pencil did not hold."""
    A = A.tocsr()
    n = A.shape[0]
    nx, ny, nz = line // ex, plane // line // ey, n // plane
    loads, widest = 0, 0
    for z0 in range(0, nz, pz):
        for py in range(ny):
            for px in range(nx):
                held = set()
                for z in range(z0, min(nz, z0 + pz)):
                    rows = [z * plane + (ey * py + dy) * line + ex * px + dx for dy in range(ey) for dx in range(ex)]
                    need = set(np.concatenate([A.indices[A.indptr[r]:A.indptr[r + 1]] for r in rows]).tolist())
                    widest = max(widest, len(need))
                    loads += len(need - held)
                    held = need
    return loads / n, widest


@pytest.mark.parametrize("g,rows", [(24, None), (24, 16), (28, None), (32, 32)])
def test_tiled_plan_load_count(dev, monkeypatch, g, rows):
    """The host slot plan loads exactly the rows the pencil walk adds: its
    reported X rows per matrix row equal an independent replay of the rule
    (24^3: one pencil per patch column, 8 x 4 patches by default, 4 x 4 when
    16-row tiles are asked for; 28^3: the line is no multiple of 8, so 4 x 4
    patches, a pencil of 24 planes and one of 4; 32^3: 8 x 4 patches, pencils
    of 24 and 8 planes)."""
    monkeypatch.setenv("FE_TILE_STENCIL", "0")  # entry-list tiles, not the stencil form
    if rows:
        monkeypatch.setenv("FE_TILE_ROWS", str(rows))
    else:
        monkeypatch.delenv("FE_TILE_ROWS", raising=False)
    A = G.hamiltonian_27pt(g, seed=2, wells=4)
    C = dev.csr(A)
    built, lpr = C.build_tiled()
    assert built
    ex = 8 if g % 8 == 0 and rows != 16 else 4
    want, widest = _replay_tile_loads(A, g, g * g, ex=ex)
    assert widest <= 3 * (ex + 2) * 6
    assert lpr == pytest.approx(want, rel=0, abs=1e-12)
    if ex == 8:  # fewer loads than the 4 x 4 plan of the same grid
        assert lpr < _replay_tile_loads(A, g, g * g, ex=4)[0] * 0.9


@pytest.mark.parametrize("rows", [16, 32])
@pytest.mark.parametrize("k,prec", [(64, _fe.F32), (100, _fe.F32), (512, _fe.F32), (32, _fe.F64), (48, _fe.F64)])
def test_spmm_tiled_both_heights(dev, monkeypatch, rows, k, prec):
    """Both tile heights, forced, bitwise equal to the CSR kernel: a 27-point
    grid whose line splits into 8 x 4 patches, and a banded matrix without a
    grid (consecutive-row tiles of 16 or 32 rows, padded last tile)."""
    monkeypatch.setenv("FE_TILE_ROWS", str(rows))
    monkeypatch.setenv("FE_TILE_STENCIL", "0")
    rng = np.random.default_rng(rows + k)
    band = sp.diags([rng.uniform(-1, 1, 3001 - abs(o)) for o in range(-9, 10)], list(range(-9, 10)))
    for A in (G.hamiltonian_27pt(16, seed=9, wells=4), sp.csr_matrix(band + band.T)):
        n = A.shape[0]
        dt = np.float64 if prec == _fe.F64 else np.float32
        X = rng.uniform(-1, 1, (n, k))
        Z = rng.uniform(-1, 1, (n, k))
        C = dev.csr(A)
        built, lpr = C.build_tiled()
        assert built and C.tiled_kind() == rows
        Xd, Zd = dev.put(X, dt), dev.put(Z, dt)
        Y1, Y2 = dev.buffer((n, k), dt), dev.buffer((n, k), dt)
        _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y1.ptr, k, 0.7, 0.3, Zd.ptr, k, -0.9))
        C.use(_fe.FMT_TILED)
        _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y2.ptr, k, 0.7, 0.3, Zd.ptr, k, -0.9))
        got = Y2.download()
        assert np.array_equal(Y1.download(), got)
        Xr, Zr = X.astype(dt).astype(np.float64), Z.astype(dt).astype(np.float64)
        ref = orc.spmm(A, Xr, 0.7, 0.0) + 0.3 * Xr - 0.9 * Zr
        tol = 1e-13 if prec == _fe.F64 else 2e-6
        assert np.max(np.abs(got.astype(np.float64) - ref)) <= tol * (np.max(np.abs(A).sum(axis=1)) + 2)


def _box_stencil(nx, ny, nz, W, d):
    """Constant-coefficient operator on an nx x ny x nz grid in natural order
    (Dirichlet: no entries past the faces): off-diagonal W[dz+1][dy+1][dx+1]
    (zeros not stored), diagonal d (zeros not stored)."""
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                w = W[dz + 1][dy + 1][dx + 1]
                if (dz, dy, dx) == (0, 0, 0) or w == 0:
                    continue
                src = idx[max(0, -dz):nz - max(0, dz), max(0, -dy):ny - max(0, dy), max(0, -dx):nx - max(0, dx)]
                rows.append(src.ravel())
                cols.append((src + dz * nx * ny + dy * nx + dx).ravel())
                vals.append(np.full(src.size, w))
    keep = d != 0
    rows.append(np.arange(n)[keep])
    cols.append(np.arange(n)[keep])
    vals.append(d[keep])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sort_indices()
    return A


def _w19():
    W = -0.5 * G.isotropic_27pt_weights()
    for dz in (0, 2):
        for dy in (0, 2):
            for dx in (0, 2):
                W[dz, dy, dx] = 0.0  # no corners: 19-point
    return W


STENCIL = {
    # 8 x 8 patches, one pencil of 16 planes
    "h16": lambda: G.hamiltonian_27pt(16, seed=3, wells=5),
    # ragged patches (30 = 3 x 8 + 6), pencils of 24 and 6 planes
    "h30": lambda: G.hamiltonian_27pt(30, seed=8, wells=9),
    # a box: 13 x 10 x 27 (ragged both ways, pencils of 24 and 3 planes)
    "box": lambda: _box_stencil(13, 10, 27, -0.5 * G.isotropic_27pt_weights(),
                                np.random.default_rng(1).normal(size=13 * 10 * 27)),
    # 7-point constant Laplacian (faces only)
    "lap7": lambda: G.laplacian_3d(12),
    # 19-point (no corners: 27-point form with zero corner coefficients), some
    # rows without a stored diagonal
    "s19": lambda: _box_stencil(11, 9, 10, _w19(),
                                np.where(np.arange(990) % 7 == 0, 0.0,
                                         np.random.default_rng(2).normal(size=990))),
}


@pytest.fixture(scope="module")
def stencil_mats():
    return {name: f() for name, f in STENCIL.items()}


@pytest.mark.parametrize("name", list(STENCIL))
@pytest.mark.parametrize("k,prec", [(4, _fe.F32), (64, _fe.F32), (100, _fe.F32), (512, _fe.F32),
                                    (32, _fe.F64), (48, _fe.F64), (100, _fe.F64)])
def test_spmm_stencil_matches_csr(dev, stencil_mats, name, k, prec):
    """Stencil form of the tiled copy (constant coefficient per offset, any
    diagonal): bitwise the CSR kernel's product, fused epilogue included, and
    within rounding of the oracle; ragged patches and pencils, partial column
    slices, 7-point and 19-point neighbourhoods, rows without a diagonal."""
    A = stencil_mats[name]
    n = A.shape[0]
    rng = np.random.default_rng(7 + k)
    X = rng.uniform(-1, 1, (n, k))
    Z = rng.uniform(-1, 1, (n, k))
    dt = np.float64 if prec == _fe.F64 else np.float32
    C = dev.csr(A)
    built, lpr = C.build_tiled()
    assert built and C.tiled_kind() == 64
    assert lpr <= 3.0
    Xd, Zd = dev.put(X, dt), dev.put(Z, dt)
    Y1, Y2 = dev.buffer((n, k), dt), dev.buffer((n, k), dt)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y1.ptr, k, 0.9, 0.4, Zd.ptr, k, -1.2))
    C.use(_fe.FMT_TILED)
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, k, Y2.ptr, k, 0.9, 0.4, Zd.ptr, k, -1.2))
    got = Y2.download()
    assert np.array_equal(Y1.download(), got)
    Xr, Zr = X.astype(dt).astype(np.float64), Z.astype(dt).astype(np.float64)
    ref = orc.spmm(A, Xr, 0.9, 0.0) + 0.4 * Xr - 1.2 * Zr
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    assert np.max(np.abs(got.astype(np.float64) - ref)) <= tol * (np.max(np.abs(A).sum(axis=1)) + 2)


def test_spmm_stencil_in_place_strided_and_without_z(dev):
    """Stencil form with Z aliasing Y (read before written), a strided block
    (ld > k) with a partial 16-byte run, beta = 0 and no Z term: bitwise the
    CSR kernel's, and the columns past k untouched."""
    A = G.hamiltonian_27pt(12, seed=5, wells=4)
    n = A.shape[0]
    ld = 72
    X = np.random.default_rng(9).uniform(-1, 1, (n, ld))
    C = dev.csr(A)
    assert C.build_tiled()[0] and C.tiled_kind() == 64
    for k in (72, 70, 37, 3):
        for prec, dt in ((_fe.F64, np.float64), (_fe.F32, np.float32)):
            for beta, zterm in ((0.5, True), (0.0, False)):
                outs = []
                for fmt in (_fe.FMT_CSR, _fe.FMT_TILED):
                    C.use(fmt)
                    Xd = dev.put(X, dt)
                    Y = dev.put(np.full((n, ld), 7.0), dt)
                    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, ld, Y.ptr, ld, 1.0, beta,
                                          Y.ptr if zterm else None, ld, 0.25 if zterm else 0.0))
                    outs.append(Y.download())
                assert np.array_equal(outs[0], outs[1])
                assert np.all(outs[1][:, k:] == 7.0)


def test_spmm_stencil_declines_varying_coefficients(dev):
    """No stencil form where one offset's coefficient varies (per-edge
    conductivities; one perturbed pair of a constant stencil): the entry-list
    tiles are built instead, bitwise the CSR product as before."""
    H = G.hamiltonian_27pt(16, seed=3, wells=5).tolil()
    H[1000, 1001] += 1e-3
    H[1001, 1000] += 1e-3
    for A in (G.conductivity_3d(16, seed=1), H.tocsr()):
        C = dev.csr(A)
        assert C.build_tiled()[0]
        assert C.tiled_kind() in (16, 32)


def test_spmm_tiled_ragged_and_in_place(dev):
    """Padded last tile, an empty row, a strided block (ld > k) with a partial
    16-byte run, and Z aliasing Y (read before written)."""
    A = G.hamiltonian_27pt(7, seed=1, wells=3).tolil()
    n = A.shape[0]  # 343 = 16 * 21 + 7
    A[n - 2, :] = 0.0
    A[:, n - 2] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    ld = 72
    X = np.random.default_rng(5).uniform(-1, 1, (n, ld))
    C = dev.csr(A)
    assert C.build_tiled()[0]
    for k in (72, 70, 37, 3):
        for prec, dt, tol in ((_fe.F64, np.float64, 1e-13), (_fe.F32, np.float32, 2e-6)):
            C.use(_fe.FMT_CSR)
            Xd = dev.put(X, dt)
            Y1 = dev.put(np.full((n, ld), 7.0), dt)
            Y2 = dev.put(np.full((n, ld), 7.0), dt)
            _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, ld, Y1.ptr, ld, 1.0, 0.5, Y1.ptr, ld, 0.25))
            C.use(_fe.FMT_TILED)
            _fe.check(L().fe_spmm(dev.ctx, C.h, prec, k, Xd.ptr, ld, Y2.ptr, ld, 1.0, 0.5, Y2.ptr, ld, 0.25))
            got = Y2.download().astype(np.float64)
            Xr = X.astype(dt).astype(np.float64)
            ref = orc.spmm(A, Xr[:, :k]) + 0.5 * Xr[:, :k] + 0.25 * 7.0
            assert np.max(np.abs(got[:, :k] - ref)) <= tol * (np.max(np.abs(A).sum(axis=1)) + 2)
            assert np.all(got[:, k:] == 7.0)
            assert np.array_equal(Y1.download(), Y2.download())


@pytest.mark.parametrize("name", ["powlaw", "lap2d"])
def test_spmm_tiled_rejects_scattered(dev, mats, name):
    """No tiled copy where a tile's rows share few columns with the next
    tile's (power law: unions past 240 rows; 5-point: loads above half the
    nonzeros per row)."""
    C = dev.csr(mats[name])
    built, _ = C.build_tiled()
    assert not built
    with pytest.raises(_fe.FeError):
        C.use(_fe.FMT_TILED)


def test_scale_after_copies_is_rejected(dev):
    """The SELL, row-block and tiled copies hold their own values: scaling
    the matrix after building one of them is an error, not a stale copy."""
    A = G.hamiltonian_27pt(12, seed=4, wells=5)
    d = dev.put(np.full(A.shape[0], 0.5), np.float64)
    for build in (lambda C: C.build_blocked(4), lambda C: C.build_tiled()):
        C = dev.csr(A)
        assert build(C)[0]
        assert L().fe_csr_scale_sym(dev.ctx, C.h, d.ptr) == -2


@pytest.mark.parametrize("name", ["powlaw", "lap2d", "cond3d"])
def test_spmm_blocked_rejects_low_reuse(dev, mats, name):
    """Scattered columns (power law) and thin stencils (5- and 7-point: under
    1.5 nonzeros per union entry of a 4-row block) get no row-block copy."""
    C = dev.csr(mats[name])
    built, reuse = C.build_blocked(4)
    assert not built and reuse < 1.5
    with pytest.raises(_fe.FeError):
        C.use(_fe.FMT_BLOCKED)


@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_spmv_strided_column(dev, mats, prec):
    """k = 1 (the Lanczos product) on one column of a wider block, fused terms."""
    A = mats["ragged"]
    n = A.shape[0]
    dt = np.float64 if prec == _fe.F64 else np.float32
    X = np.random.default_rng(9).uniform(-1, 1, (n, 3)).astype(dt)
    Z = np.random.default_rng(10).uniform(-1, 1, (n, 3)).astype(dt)
    Xd, Yd = dev.put(X, dt), dev.put(Z, dt)
    C = dev.csr(A)
    es = np.dtype(dt).itemsize
    _fe.check(L().fe_spmm(dev.ctx, C.h, prec, 1, Xd.ptr + es, 3, Yd.ptr + 2 * es, 3, 0.5, -2.0, Yd.ptr + 2 * es, 3,
                          0.75))
    got = Yd.download().astype(np.float64)
    x = X[:, 1].astype(np.float64)
    ref = 0.5 * (A @ x) - 2.0 * x + 0.75 * Z[:, 2]
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    assert np.max(np.abs(got[:, 2] - ref)) <= tol * (np.max(np.abs(A).sum(axis=1)) + 3)
    assert np.array_equal(got[:, :2], Z[:, :2].astype(np.float64))


def test_spmm_strided_subblock(dev, mats):
    """Column sub-ranges with ld > k (unaligned) take the scalar path."""
    A = mats["lap2d"]
    n = A.shape[0]
    X = np.random.default_rng(2).uniform(-1, 1, (n, 21))
    Xd = dev.put(X)
    Yd = dev.put(np.zeros((n, 21)))
    C = dev.csr(A)
    off = 8 * 3  # bytes: start at column 3
    _fe.check(L().fe_spmm(dev.ctx, C.h, _fe.F64, 7, Xd.ptr + off, 21, Yd.ptr + off, 21, 1.0, 0.0, None, 0, 0.0))
    got = Yd.download()
    ref = orc.spmm(A, X[:, 3:10])
    assert np.max(np.abs(got[:, 3:10] - ref)) < 1e-13
    assert np.all(got[:, :3] == 0) and np.all(got[:, 10:] == 0)


@pytest.mark.parametrize("n,m,p", [(1000, 16, 16), (5000, 48, 40), (20000, 96, 96), (3000, 200, 70),
                                   (777, 5, 3), (100000, 256, 256), (40000, 512, 512), (200003, 192, 48),
                                   (50000, 144, 144), (3001, 65, 64), (4000, 161, 150), (884736, 20, 1),
                                   (100003, 32, 1), (5000, 1, 1), (777, 9, 1), (3000, 33, 1)])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_gram(dev, n, m, p, prec):
    rng = np.random.default_rng(n + m)
    X = rng.uniform(-1, 1, (n, m))
    Y = X[:, :p] * 0.5 + rng.uniform(-1, 1, (n, p)) * 0.5
    w = rng.uniform(0.5, 1.5, n)
    dt = np.float64 if prec == _fe.F64 else np.float32
    Xd, Yd, wd = dev.put(X, dt), dev.put(Y, dt), dev.put(w)
    Gd = dev.buffer((m, p), np.float64, order="F")
    _fe.check(L().fe_gram(dev.ctx, 0, prec, n, m, p, Xd.ptr, m, Yd.ptr, p, wd.ptr, Gd.ptr, m))
    Xr = X.astype(dt).astype(np.float64)
    Yr = Y.astype(dt).astype(np.float64)
    ref = Xr.T @ (w[:, None] * Yr)
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    assert rel(Gd.download(), ref) <= tol * np.sqrt(n) / 10 + tol
    # symmetric path (X == Y, no weight): exactly symmetric output
    Sd = dev.buffer((m, m), np.float64, order="F")
    _fe.check(L().fe_gram(dev.ctx, 0, prec, n, m, m, Xd.ptr, m, Xd.ptr, m, None, Sd.ptr, m))
    S = Sd.download()
    assert np.array_equal(S, S.T)
    assert rel(S, Xr.T @ Xr) <= tol * np.sqrt(n) / 10 + tol


def test_gram_bf16x3(dev):
    rng = np.random.default_rng(3)
    n, m = 50000, 128
    X = rng.uniform(-1, 1, (n, m)).astype(np.float32)
    Xd = dev.put(X, np.float32)
    Gd = dev.buffer((m, m), np.float64, order="F")
    _fe.check(L().fe_gram(dev.ctx, _fe.GRAM_BF16X3, _fe.F32, n, m, m, Xd.ptr, m, Xd.ptr, m, None, Gd.ptr, m))
    ref = X.astype(np.float64).T @ X.astype(np.float64)
    # x0 y0 + x0 y1 + x1 y0 keeps about 16 significant bits per product
    assert rel(Gd.download(), ref) < 2.0 ** -14


@pytest.mark.parametrize("n,m,p", [(50000, 128, 128), (40000, 512, 512), (3001, 65, 64), (4000, 161, 150),
                                   (777, 5, 3), (200003, 192, 48), (31, 300, 260)])
def test_gram_bf16_modes(dev, n, m, p):
    """The compensated Gram products against the FP64 product of the same FP32
    inputs: BF16x3 to about 16 bits; BF16x6 to FP32 accuracy (within twice the
    error of the native FP32 MFMA product). Weighted, ragged and symmetric
    cases; the symmetric output is exactly symmetric."""
    rng = np.random.default_rng(n + 7 * m)
    X = rng.uniform(-1, 1, (n, m))
    Y = X[:, :p] * 0.5 + rng.uniform(-1, 1, (n, p)) * 0.5
    w = rng.uniform(0.5, 1.5, n)
    Xd, Yd, wd = dev.put(X, np.float32), dev.put(Y, np.float32), dev.put(w)
    Xr = X.astype(np.float32).astype(np.float64)
    Yr = Y.astype(np.float32).astype(np.float64)
    ref = Xr.T @ (w[:, None] * Yr)
    sref = Xr.T @ Xr

    def run(mode, sym):
        Gd = dev.buffer((m, m if sym else p), np.float64, order="F")
        if sym:
            _fe.check(L().fe_gram(dev.ctx, mode, _fe.F32, n, m, m, Xd.ptr, m, Xd.ptr, m, None, Gd.ptr, m))
        else:
            _fe.check(L().fe_gram(dev.ctx, mode, _fe.F32, n, m, p, Xd.ptr, m, Yd.ptr, p, wd.ptr, Gd.ptr, m))
        return Gd.download()

    # BF16x6: FP32-class, within twice the native FP32 error or 2^-21 relative
    # (on sums of positive terms, the diagonal of X^T X over ~600-row chunks,
    # the BF16 MFMA's internal sums lose up to 4x more than the FP32 MFMA's;
    # the FP32 worst-case bound there is 600 u = 3.6e-5)
    e32, s32 = rel(run(_fe.GRAM_NATIVE, False), ref), rel(run(_fe.GRAM_NATIVE, True), sref)
    for mode in (_fe.GRAM_BF16X3, _fe.GRAM_BF16X6):
        G, S = run(mode, False), run(mode, True)
        assert np.array_equal(S, S.T)
        if mode == _fe.GRAM_BF16X3:
            assert rel(G, ref) < 2.0 ** -14 and rel(S, sref) < 2.0 ** -14, (rel(G, ref), rel(S, sref))
        else:
            assert rel(G, ref) <= max(2 * e32, 2.0 ** -21), (rel(G, ref), e32)
            assert rel(S, sref) <= max(2 * s32, 2.0 ** -21), (rel(S, sref), s32)


@pytest.mark.parametrize("n,m,p", [(40000, 512, 512), (8191, 256, 512), (100000, 256, 256), (33, 256, 256),
                                   (65, 512, 256)])
def test_gram_bf16x3_fast_path_bitwise(dev, monkeypatch, n, m, p):
    """The BF16x3 Gram's fast path (whole 256-column tiles, no weights; two
    stages of loads in flight, rows past a chunk zeroed by selects) gives the
    general kernel's partials bitwise, symmetric and not, with a chunk whose
    stage count is odd (a stage of zeros) and chunks shorter than a stage."""
    rng = np.random.default_rng(n + m + p)
    X = rng.uniform(-1, 1, (n, m)).astype(np.float32)
    Y = rng.uniform(-1, 1, (n, p)).astype(np.float32)
    Xd, Yd = dev.put(X, np.float32), dev.put(Y, np.float32)
    outs = {}
    for fast in ("1", "0"):
        monkeypatch.setenv("FE_GRAM_FAST", fast)
        Gd = dev.buffer((m, p), np.float64, order="F")
        Sd = dev.buffer((m, m), np.float64, order="F")
        _fe.check(L().fe_gram(dev.ctx, _fe.GRAM_BF16X3, _fe.F32, n, m, p, Xd.ptr, m, Yd.ptr, p, None, Gd.ptr, m))
        _fe.check(L().fe_gram(dev.ctx, _fe.GRAM_BF16X3, _fe.F32, n, m, m, Xd.ptr, m, Xd.ptr, m, None, Sd.ptr, m))
        outs[fast] = (Gd.download(), Sd.download())
    assert np.array_equal(outs["1"][0], outs["0"][0])
    assert np.array_equal(outs["1"][1], outs["0"][1])
    ref = X.astype(np.float64).T @ Y.astype(np.float64)
    assert rel(outs["1"][0], ref) < 2.0 ** -14


@pytest.mark.parametrize("n,m,p", [(20000, 512, 512), (100000, 256, 256), (3333, 200, 70), (999, 130, 130),
                                   (7, 3, 2), (1001, 48, 64), (300, 33, 1)])
def test_gemm_bf16_modes(dev, n, m, p):
    """Y = alpha X C + beta Y in the compensated modes (C rounded to FP32 and
    split), against FP64 on the same rounded inputs; and the triangular
    product, whose C holds NaN below the diagonal that must not be read."""
    rng = np.random.default_rng(n + p)
    X = rng.uniform(-1, 1, (n, m))
    Cm = rng.uniform(-1, 1, (m, p))
    Y0 = rng.uniform(-1, 1, (n, p))
    Xd = dev.put(X, np.float32)
    Cd = dev.put(Cm, np.float64, order="F")
    Xr = X.astype(np.float32).astype(np.float64)
    Cr = Cm.astype(np.float32).astype(np.float64)
    ref = 0.7 * Xr @ Cr - 0.5 * Y0.astype(np.float32).astype(np.float64)

    def gemm(mode):
        Yd = dev.put(Y0, np.float32)
        _fe.check(L().fe_gemm_nn_mode(dev.ctx, mode, _fe.F32, n, m, p, Xd.ptr, m, Cd.ptr, m, Yd.ptr, p, 0.7, -0.5))
        return Yd.download().astype(np.float64)

    e32 = np.max(np.abs(gemm(_fe.GRAM_NATIVE) - ref))
    scale = np.max(np.abs(Xr) @ np.abs(Cr))
    assert np.max(np.abs(gemm(_fe.GRAM_BF16X3) - ref)) <= 2.0 ** -14 * scale
    assert np.max(np.abs(gemm(_fe.GRAM_BF16X6) - ref)) <= 2 * e32 + 1e-7 * scale
    if m == p:
        Cg = Cm.copy()
        Cg[np.tril_indices(m, -1)] = np.nan
        Ct = dev.put(Cg, np.float64, order="F")
        tref = Xr @ np.triu(Cr)

        def trmm(mode):
            Yd = dev.buffer((n, m), np.float32)
            _fe.check(L().fe_trmm_upper_mode(dev.ctx, mode, _fe.F32, n, m, Xd.ptr, m, Ct.ptr, m, Yd.ptr, m))
            return Yd.download().astype(np.float64)

        t32 = np.max(np.abs(trmm(_fe.GRAM_NATIVE) - tref))
        for mode, bound in ((_fe.GRAM_BF16X3, 2.0 ** -14 * scale), (_fe.GRAM_BF16X6, 2 * t32 + 1e-7 * scale)):
            out = trmm(mode)
            assert np.all(np.isfinite(out)) and np.max(np.abs(out - tref)) <= bound, mode


def test_bf16_modes_reject_fp64(dev):
    Xd = dev.put(np.ones((10, 4)))
    Cd = dev.put(np.eye(4), np.float64, order="F")
    Yd = dev.buffer((10, 4), np.float64)
    Gd = dev.buffer((4, 4), np.float64, order="F")
    assert L().fe_gemm_nn_mode(dev.ctx, _fe.GRAM_BF16X6, _fe.F64, 10, 4, 4, Xd.ptr, 4, Cd.ptr, 4, Yd.ptr, 4,
                               1.0, 0.0) != 0
    assert L().fe_gram(dev.ctx, _fe.GRAM_BF16X6, _fe.F64, 10, 4, 4, Xd.ptr, 4, Xd.ptr, 4, None, Gd.ptr, 4) != 0


def test_gram_reproducible(dev):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((30000, 64))
    Xd = dev.put(X)
    outs = []
    for _ in range(2):
        Gd = dev.buffer((64, 64), np.float64, order="F")
        _fe.check(L().fe_gram(dev.ctx, 0, _fe.F64, 30000, 64, 64, Xd.ptr, 64, Xd.ptr, 64, None, Gd.ptr, 64))
        outs.append(Gd.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("n,m", [(1000, 16), (5000, 48), (3333, 200), (20000, 512), (7, 3), (200001, 64),
                                 (999, 130), (300, 1)])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_trmm_upper(dev, n, m, prec):
    """X C with C upper triangular: below-diagonal entries hold garbage that
    must not be read; same result as the full product with triu(C)."""
    rng = np.random.default_rng(n + 3 * m)
    X = rng.uniform(-1, 1, (n, m))
    Cm = rng.uniform(-1, 1, (m, m))
    Cg = Cm.copy()
    Cg[np.tril_indices(m, -1)] = np.nan  # never read
    dt = np.float64 if prec == _fe.F64 else np.float32
    Xd, Yd = dev.put(X, dt), dev.buffer((n, m), dt)
    Cd = dev.put(Cg, np.float64, order="F")
    _fe.check(L().fe_trmm_upper(dev.ctx, prec, n, m, Xd.ptr, m, Cd.ptr, m, Yd.ptr, m))
    Xr = X.astype(dt).astype(np.float64)
    ref = Xr @ np.triu(Cm).astype(dt).astype(np.float64)
    out = Yd.download().astype(np.float64)
    assert np.all(np.isfinite(out))
    tol = 1e-13 if prec == _fe.F64 else 2e-6
    assert rel(out, ref) <= tol * np.sqrt(m) + tol


@pytest.mark.parametrize("n,m,p", [(1000, 16, 16), (5000, 48, 10), (20000, 96, 96), (3333, 200, 70),
                                   (100000, 256, 256), (7, 3, 2), (200001, 192, 48), (1001, 48, 64),
                                   (999, 20, 65), (300, 33, 1), (100003, 20, 1), (7, 64, 1), (4097, 65, 1)])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_gemm_nn(dev, n, m, p, prec):
    rng = np.random.default_rng(n + p)
    X = rng.uniform(-1, 1, (n, m))
    Cm = rng.uniform(-1, 1, (m, p))
    Y0 = rng.uniform(-1, 1, (n, p))
    dt = np.float64 if prec == _fe.F64 else np.float32
    Xd, Yd = dev.put(X, dt), dev.put(Y0, dt)
    Cd = dev.put(Cm, np.float64, order="F")
    _fe.check(L().fe_gemm_nn(dev.ctx, prec, n, m, p, Xd.ptr, m, Cd.ptr, m, Yd.ptr, p, 0.7, -0.5))
    Xr = X.astype(dt).astype(np.float64)
    ref = 0.7 * Xr @ Cm.astype(dt).astype(np.float64) - 0.5 * Y0.astype(dt).astype(np.float64)
    tol = 1e-13 if prec == _fe.F64 else 3e-6
    assert np.max(np.abs(Yd.download() - ref)) <= tol * m


@pytest.mark.parametrize("k", [1, 7, 31, 32, 33, 64, 100, 130, 512])
def test_potrf_rinv_cholqr(dev, k):
    rng = np.random.default_rng(k)
    n = max(4 * k, 50)
    X = rng.uniform(-1, 1, (n, k))
    Xd = dev.put(X)
    Gd = dev.buffer((k, k), np.float64, order="F")
    _fe.check(L().fe_gram(dev.ctx, 0, _fe.F64, n, k, k, Xd.ptr, k, Xd.ptr, k, None, Gd.ptr, k))
    assert _fe.check(L().fe_potrf(dev.ctx, k, Gd.ptr, k)) == 0
    # the factor against LAPACK's (through numpy) on the host Gram matrix
    Lref = np.linalg.cholesky(X.T @ X)
    assert np.max(np.abs(np.tril(Gd.download()) - Lref)) <= 1e-11 * np.max(np.abs(Lref))
    Rd = dev.buffer((k, k), np.float64, order="F")
    _fe.check(L().fe_chol_rinv(dev.ctx, k, Gd.ptr, k, Rd.ptr, k))
    Rref = np.linalg.inv(Lref).T
    assert np.max(np.abs(Rd.download() - Rref)) <= 1e-10 * np.max(np.abs(Rref))
    Qd = dev.buffer((n, k))
    _fe.check(L().fe_gemm_nn(dev.ctx, _fe.F64, n, k, k, Xd.ptr, k, Rd.ptr, k, Qd.ptr, k, 1.0, 0.0))
    Q = Qd.download()
    Rinv = Rd.download()
    assert np.allclose(np.triu(Rinv), Rinv)
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) < 1e-10
    # Q R = X with R = inv(Rinv); oracle CholQR gives the same factors
    info, Qo, Ro = orc.cholqr(X)
    assert info == 0
    assert np.max(np.abs(Q - Qo)) < 1e-9


@pytest.mark.parametrize("k,bad", [(6, 3), (100, 70), (64, 0)])
def test_potrf_reports_failure(dev, k, bad):
    """The first non-positive pivot (column bad + 1) is reported, also in a
    later block of the blocked factorization."""
    Gm = np.eye(k) * 2.0
    Gm[bad, bad] = -1.0
    Gd = dev.put(Gm, np.float64, order="F")
    assert L().fe_potrf(dev.ctx, k, Gd.ptr, k) == bad + 1
    # the next factorization on the same context starts clean
    Gd2 = dev.put(np.eye(k) * 4.0, np.float64, order="F")
    assert L().fe_potrf(dev.ctx, k, Gd2.ptr, k) == 0
    assert np.allclose(np.tril(Gd2.download()), np.eye(k) * 2.0)


@pytest.mark.parametrize("k", [130, 512])
def test_potrf_rinv_repeated(dev, k):
    """Six factorizations through the same buffers (SPD matrices and, in
    between, ones with a negative pivot whose column must be reported), with
    eager work between them, each checked against LAPACK's factor and the
    inverse."""
    rng = np.random.default_rng(k)
    Gd = dev.buffer((k, k), np.float64, order="F")
    Rd = dev.buffer((k, k), np.float64, order="F")
    Xd = dev.put(rng.uniform(-1, 1, (3000, 16)))
    Sd = dev.buffer((16, 16), np.float64, order="F")
    for rep in range(6):
        B = rng.standard_normal((k, k))
        Gm = B @ B.T + k * np.eye(k)
        bad = 37 + 11 * rep if rep in (2, 4) else None
        if bad is not None:
            Gm[bad, bad] = -1.0
        Gd.upload(Gm)
        info = L().fe_potrf(dev.ctx, k, Gd.ptr, k)
        if bad is not None:
            assert info >= 1 and info <= bad + 1, (rep, info)
            continue
        assert info == 0, rep
        Lc = np.tril(Gd.download())
        assert np.max(np.abs(Lc - np.linalg.cholesky(Gm))) <= 1e-12 * np.sqrt(np.max(np.abs(Gm)))
        _fe.check(L().fe_chol_rinv(dev.ctx, k, Gd.ptr, k, Rd.ptr, k))
        assert np.max(np.abs(Lc.T @ Rd.download() - np.eye(k))) <= 1e-12
        _fe.check(L().fe_gram(dev.ctx, 0, _fe.F64, 3000, 16, 16, Xd.ptr, 16, Xd.ptr, 16, None, Sd.ptr, 16))
        dev.sync()


def _syev(dev, A):
    m = A.shape[0]
    Ad = dev.put(A, np.float64, order="F")
    Wd = dev.buffer((m,), np.float64)
    Vd = dev.buffer((m, m), np.float64, order="F")
    info = _fe.check(L().fe_syev(dev.ctx, m, Ad.ptr, m, Wd.ptr, Vd.ptr, m))
    return info, Wd.download(), Vd.download()


@pytest.mark.parametrize("m", [1, 2, 3, 5, 17, 64, 100, 128, 129, 130, 257, 512, 700, 1500, 2048])
def test_syev_random(dev, m):
    rng = np.random.default_rng(m)
    A = rng.standard_normal((m, m))
    A = A + A.T
    info, w, V = _syev(dev, A)
    assert info == 0
    wr = np.linalg.eigvalsh(A)
    scale = max(np.max(np.abs(wr)), 1.0)
    assert np.max(np.abs(w - wr)) <= 1e-13 * scale * max(1, np.sqrt(m))
    assert np.max(np.abs(V.T @ V - np.eye(m))) <= 1e-11
    assert np.max(np.abs(A @ V - V * w)) <= 1e-12 * scale * max(1, np.sqrt(m))
    if m <= 700:  # the oracle's dense QR iteration is O(m^3) on one core
        wo, _ = orc.syev(A)
        assert np.max(np.abs(w - wo)) <= 1e-13 * scale * max(1, np.sqrt(m))


@pytest.mark.parametrize("m", [64, 300, 512])
def test_syev_graph_replays(dev, m):
    """The tridiagonalization's column steps are captured once per size and
    workspace and replayed: five different matrices through the same buffers
    (the first call captures, the others replay), with eager work between the
    calls (a syev of another size, a Gram, a host download), each checked
    against LAPACK."""
    rng = np.random.default_rng(100 + m)
    Ad = dev.buffer((m, m), np.float64, order="F")
    Wd = dev.buffer((m,), np.float64)
    Vd = dev.buffer((m, m), np.float64, order="F")
    Xd = dev.put(rng.uniform(-1, 1, (2000, 8)))
    Gd = dev.buffer((8, 8), np.float64, order="F")
    for rep in range(5):
        A = rng.standard_normal((m, m))
        A = A + A.T
        Ad.upload(A)
        assert _fe.check(L().fe_syev(dev.ctx, m, Ad.ptr, m, Wd.ptr, Vd.ptr, m)) == 0
        w, V = Wd.download(), Vd.download()
        wr = np.linalg.eigvalsh(A)
        scale = max(np.max(np.abs(wr)), 1.0)
        assert np.max(np.abs(w - wr)) <= 1e-13 * scale * np.sqrt(m), rep
        assert np.max(np.abs(A @ V - V * w)) <= 1e-12 * scale * np.sqrt(m), rep
        # eager work between the replays
        B = rng.standard_normal((40, 40))
        info, wb, _ = _syev(dev, B + B.T)
        assert info == 0 and np.max(np.abs(wb - np.linalg.eigvalsh(B + B.T))) <= 1e-12 * 40
        _fe.check(L().fe_gram(dev.ctx, 0, _fe.F64, 2000, 8, 8, Xd.ptr, 8, Xd.ptr, 8, None, Gd.ptr, 8))
        dev.sync()


def test_syev_rejects_above_2048(dev):
    A = dev.buffer((2049, 2049), np.float64, order="F")
    W = dev.buffer((2049,), np.float64)
    V = dev.buffer((2049, 2049), np.float64, order="F")
    assert L().fe_syev(dev.ctx, 2049, A.ptr, 2049, W.ptr, V.ptr, 2049) == -2


def test_syev_degenerate_and_diagonal(dev):
    # exact pairs (2D Laplacian symmetry), a diagonal matrix and a Wilkinson matrix
    L2 = G.laplacian_2d(6).toarray()
    for A in (L2, np.diag(np.arange(10.0)[::-1]), np.diag(np.abs(np.arange(-10, 11.0))) +
              np.diag(np.ones(20), 1) + np.diag(np.ones(20), -1), np.zeros((4, 4)), np.eye(9) * 3):
        info, w, V = _syev(dev, A)
        assert info == 0
        m = A.shape[0]
        scale = max(np.max(np.abs(A)), 1.0)
        assert np.max(np.abs(w - np.linalg.eigvalsh(A))) <= 1e-13 * scale * 4
        assert np.max(np.abs(V.T @ V - np.eye(m))) <= 1e-11
        assert np.max(np.abs(A @ V - V * w)) <= 1e-12 * scale * 4


@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
@pytest.mark.parametrize("n,k", [(12345, 37), (12345, 1), (884736, 1), (1000, 3), (77, 8), (5, 9)])
def test_colnorms_and_ritz_residuals(dev, prec, n, k):
    """Wide blocks take 64 columns per workgroup, at most 8 columns the
    narrow kernel (every thread on rows): both against numpy."""
    rng = np.random.default_rng(11 + k)
    Y = rng.uniform(-1, 1, (n, k))
    AY = rng.uniform(-1, 1, (n, k))
    th = rng.uniform(-2, 2, k)
    dt = np.float64 if prec == _fe.F64 else np.float32
    Yd, AYd, thd = dev.put(Y, dt), dev.put(AY, dt), dev.put(th)
    Rd = dev.buffer((n, k), dt)
    rn, yn = dev.buffer((k,)), dev.buffer((k,))
    _fe.check(L().fe_ritz_residuals(dev.ctx, prec, n, k, Yd.ptr, k, AYd.ptr, k, thd.ptr, Rd.ptr, k, rn.ptr, yn.ptr))
    Yr, AYr = Y.astype(dt).astype(np.float64), AY.astype(dt).astype(np.float64)
    R = AYr - Yr * th
    tol = 1e-13 if prec == _fe.F64 else 1e-6
    assert rel(rn.download(), np.linalg.norm(R, axis=0)) < tol
    assert rel(yn.download(), np.linalg.norm(Yr, axis=0)) < tol
    assert rel(Rd.download().astype(np.float64), R) < (1e-15 if prec == _fe.F64 else 1e-6)
    cn = dev.buffer((k,))
    _fe.check(L().fe_colnorms(dev.ctx, prec, n, k, Yd.ptr, k, cn.ptr))
    assert rel(cn.download(), np.linalg.norm(Yr, axis=0)) < tol


RESID_MODES = [(_fe.F64, _fe.GRAM_NATIVE), (_fe.F32, _fe.GRAM_NATIVE), (_fe.F32, _fe.GRAM_BF16X3),
               (_fe.F32, _fe.GRAM_BF16X6)]


@pytest.mark.parametrize("n,m,p", [(12345, 37, 37), (20000, 512, 256), (3001, 96, 48), (777, 130, 130), (5, 8, 3),
                                   (100003, 64, 64), (4099, 200, 70)])
@pytest.mark.parametrize("prec,mode", RESID_MODES)
def test_gemm_nn_resid(dev, n, m, p, prec, mode):
    """Rotation with the fused Ritz residual (fe_gemm_nn_resid) against the
    unfused pair it replaces (fe_gemm_nn_mode into AY, then
    fe_ritz_residuals): AY and R bitwise (the same product kernel and element
    formula), the norms within summation-order rounding; the variants without
    AY or R give bitwise the same norms; and the norms against numpy on the
    same rounded data."""
    rng = np.random.default_rng(n + m + p)
    X = rng.uniform(-1, 1, (n, m))
    Cm = rng.uniform(-1, 1, (m, p)) / np.sqrt(m)
    Y = rng.uniform(-1, 1, (n, p))
    th = rng.uniform(-2, 2, p)
    dt = np.float64 if prec == _fe.F64 else np.float32
    Xd, Yd, Cd, thd = dev.put(X, dt), dev.put(Y, dt), dev.put(Cm, np.float64, order="F"), dev.put(th)
    AYp, Rp = dev.buffer((n, p), dt), dev.buffer((n, p), dt)
    rn0, yn0 = dev.buffer((p,)), dev.buffer((p,))
    _fe.check(L().fe_gemm_nn_mode(dev.ctx, mode, prec, n, m, p, Xd.ptr, m, Cd.ptr, m, AYp.ptr, p, 1.0, 0.0))
    _fe.check(L().fe_ritz_residuals(dev.ctx, prec, n, p, Yd.ptr, p, AYp.ptr, p, thd.ptr, Rp.ptr, p, rn0.ptr, yn0.ptr))
    AYf, Rf = dev.buffer((n, p), dt), dev.buffer((n, p), dt)
    rn1, yn1 = dev.buffer((p,)), dev.buffer((p,))
    _fe.check(L().fe_gemm_nn_resid(dev.ctx, mode, prec, n, m, p, Xd.ptr, m, Cd.ptr, m, AYf.ptr, p, Yd.ptr, p,
                                   thd.ptr, Rf.ptr, p, rn1.ptr, yn1.ptr))
    assert np.array_equal(AYf.download(), AYp.download())
    assert np.array_equal(Rf.download(), Rp.download())
    r0, r1 = rn0.download(), rn1.download()
    assert rel(r1, r0) <= 1e-13 and rel(yn1.download(), yn0.download()) <= 1e-13
    rn2, yn2 = dev.buffer((p,)), dev.buffer((p,))
    _fe.check(L().fe_gemm_nn_resid(dev.ctx, mode, prec, n, m, p, Xd.ptr, m, Cd.ptr, m, None, p, Yd.ptr, p,
                                   thd.ptr, None, p, rn2.ptr, yn2.ptr))
    assert np.array_equal(rn2.download(), r1) and np.array_equal(yn2.download(), yn1.download())
    Yr = Y.astype(dt).astype(np.float64)
    R = AYp.download().astype(np.float64) - Yr * th
    tol = 1e-13 if prec == _fe.F64 else 1e-6
    assert rel(r1, np.linalg.norm(R, axis=0)) < tol
    assert rel(yn1.download(), np.linalg.norm(Yr, axis=0)) < tol


def test_gemm_nn_resid_rejects_aliasing(dev):
    n, m, p = 100, 8, 8
    Xd, Yd = dev.buffer((n, m)), dev.buffer((n, p))
    Cd, thd, rn, yn = dev.buffer((m, p), order="F"), dev.buffer((p,)), dev.buffer((p,)), dev.buffer((p,))
    for ay, r in ((Xd, None), (None, Yd), (Yd, None)):
        rc = L().fe_gemm_nn_resid(dev.ctx, 0, _fe.F64, n, m, p, Xd.ptr, m, Cd.ptr, m, ay.ptr if ay else None, p,
                                  Yd.ptr, p, thd.ptr, r.ptr if r else None, p, rn.ptr, yn.ptr)
        assert rc == -2
    # an empty inner dimension is rejected (there is no product to fuse the residual into)
    assert L().fe_gemm_nn_resid(dev.ctx, 0, _fe.F64, n, 0, p, Xd.ptr, m, Cd.ptr, m, None, p, Yd.ptr, p, thd.ptr,
                                None, p, rn.ptr, yn.ptr) == -2


@pytest.mark.parametrize("g,m", [(20, 512), (14, 65), (20, 144), (9, 5), (30, 256), (24, 130)])
@pytest.mark.parametrize("prec,mode", RESID_MODES)
def test_gram_symmetric_product(dev, g, m, prec, mode):
    """FE_GRAM_SYMMETRIC on the Rayleigh-Ritz matrix Q^T (A Q) of a symmetric
    A: exactly symmetric output, within the mode's product accuracy of the
    FP64 product of the same rounded data (which is symmetric up to rounding);
    the flag with m != p is rejected."""
    A = G.hamiltonian_27pt(g, seed=4, wells=5)
    n = A.shape[0]
    rng = np.random.default_rng(g + m)
    dt = np.float64 if prec == _fe.F64 else np.float32
    Q = rng.uniform(-1, 1, (n, m)).astype(dt).astype(np.float64)
    AQ = (A @ Q).astype(dt).astype(np.float64)
    Qd, AQd = dev.put(Q, dt), dev.put(AQ, dt)
    Gd = dev.buffer((m, m), np.float64, order="F")
    _fe.check(L().fe_gram(dev.ctx, mode | _fe.GRAM_SYMMETRIC, prec, n, m, m, Qd.ptr, m, AQd.ptr, m, None, Gd.ptr, m))
    S = Gd.download()
    assert np.array_equal(S, S.T)
    ref = Q.T @ AQ
    scale = np.max(np.abs(Q).T @ np.abs(AQ))
    tol = {(_fe.F64, 0): 1e-14, (_fe.F32, 0): 1e-6, (_fe.F32, 1): 2.0 ** -14, (_fe.F32, 2): 1e-6}[(prec, mode)]
    assert np.max(np.abs(S - ref)) <= tol * scale
    assert L().fe_gram(dev.ctx, mode | _fe.GRAM_SYMMETRIC, prec, n, m, m - 1, Qd.ptr, m, AQd.ptr, m, None,
                       Gd.ptr, m) == -2


@pytest.mark.parametrize("m", [1, 7, 256, 300, 512])
def test_identity_deviation(dev, m):
    rng = np.random.default_rng(m)
    Gm = np.eye(m) + 1e-3 * rng.standard_normal((m, m))
    Gd = dev.put(Gm, np.float64, order="F")
    out = ctypes.c_double()
    _fe.check(L().fe_identity_deviation(dev.ctx, m, Gd.ptr, m, ctypes.byref(out)))
    assert abs(out.value - np.linalg.norm(Gm - np.eye(m))) <= 1e-14 * np.linalg.norm(Gm - np.eye(m))


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
@pytest.mark.parametrize("m,n,k", [(512, 512, 512), (1, 1, 1), (65, 33, 100), (130, 7, 64), (200, 300, 17)])
def test_gemm_small(dev, ta, tb, m, n, k):
    """Small column-major FP64 product against numpy, with leading dimensions
    above the sizes and a nonzero beta."""
    rng = np.random.default_rng(m * 7 + n * 3 + k + 2 * ta + tb)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    pad = 3
    def put_ld(M):
        P = np.zeros((M.shape[0] + pad, M.shape[1]))
        P[:M.shape[0]] = M
        return dev.put(P, np.float64, order="F"), M.shape[0] + pad
    Ad, lda = put_ld(A)
    Bd, ldb = put_ld(B)
    Cd, ldc = put_ld(C0)
    _fe.check(L().fe_gemm_small(dev.ctx, ta, tb, m, n, k, 1.5, Ad.ptr, lda, Bd.ptr, ldb, -0.5, Cd.ptr, ldc))
    ref = 1.5 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.5 * C0
    got = Cd.download()[:m]
    assert rel(got, ref) < 1e-13
    assert np.array_equal(Cd.download()[m:], np.zeros((pad, n)))  # rows past m untouched
    # beta = 0 ignores what C holds (NaN here)
    Cn, _ = put_ld(np.full((m, n), np.nan))
    _fe.check(L().fe_gemm_small(dev.ctx, ta, tb, m, n, k, 1.0, Ad.ptr, lda, Bd.ptr, ldb, 0.0, Cn.ptr, ldc))
    assert rel(Cn.download()[:m], (A.T if ta else A) @ (B.T if tb else B)) < 1e-13
    assert L().fe_gemm_small(dev.ctx, ta, tb, m, n, k, 1.0, Ad.ptr, lda, Bd.ptr, ldb, 0.0, Ad.ptr, ldc) == -2


@pytest.mark.parametrize("m", [1, 37, 512])
def test_gram_equilibrate_and_row_scaling(dev, m):
    """D G D with d_j = 1/sqrt(G_jj) (1 for a zero column) and diag(d) A, on a
    Gram whose column norms span twelve decades."""
    rng = np.random.default_rng(m + 5)
    X = rng.standard_normal((3 * m + 4, m)) * np.logspace(-6, 6, m)[None, :]
    if m > 1:
        X[:, m // 2] = 0.0
    Gm = X.T @ X
    ld = m + 2
    Gp = np.zeros((ld, m))
    Gp[:m] = Gm
    Gd = dev.put(Gp, np.float64, order="F")
    dd = dev.buffer((m,), np.float64)
    _fe.check(L().fe_gram_equilibrate(dev.ctx, m, Gd.ptr, ld, dd.ptr))
    g = np.diag(Gm)
    d = np.where(g > 0, 1.0 / np.sqrt(np.where(g > 0, g, 1.0)), 1.0)
    assert np.allclose(dd.download(), d, rtol=1e-15, atol=0)
    got = Gd.download()
    assert rel(got[:m], d[:, None] * Gm * d[None, :]) < 1e-14
    assert np.array_equal(got[m:], np.zeros((2, m)))
    Ap = rng.standard_normal((ld, m + 3))
    Ad = dev.put(Ap, np.float64, order="F")
    _fe.check(L().fe_scale_rows_small(dev.ctx, m, m + 3, dd.ptr, Ad.ptr, ld))
    back = Ad.download()
    assert rel(back[:m], d[:, None] * Ap[:m]) < 1e-15
    assert np.array_equal(back[m:], Ap[m:])


@pytest.mark.parametrize("grid", [1, 7, 4096])
def test_stream_probe(dev, grid):
    """The streaming probe moves what it says: copy, the three-input sum and
    the write pattern, over sizes that are not a multiple of the grid."""
    rng = np.random.default_rng(grid)
    rows = 3001  # 3001 x 4 FP32 = 3001 16-byte accesses
    a, b, c = (rng.uniform(-1, 1, (rows, 4)).astype(np.float32) for _ in range(3))
    Ad, Bd, Cd = dev.put(a, np.float32), dev.put(b, np.float32), dev.put(c, np.float32)
    D = dev.buffer((rows, 4), np.float32)
    nb = rows * 16
    _fe.check(L().fe_stream_probe(dev.ctx, 2, nb, Ad.ptr, None, None, D.ptr, grid))
    assert np.array_equal(D.download(), a)
    _fe.check(L().fe_stream_probe(dev.ctx, 3, nb, Ad.ptr, Bd.ptr, Cd.ptr, D.ptr, grid))
    assert np.array_equal(D.download(), (a + b) + c)
    _fe.check(L().fe_stream_probe(dev.ctx, 1, nb, None, None, None, D.ptr, grid))
    assert np.array_equal(D.download(), np.tile(np.float32([1, 2, 3, 4]), (rows, 1)))
    Z = dev.buffer((rows, 4), np.float32)
    _fe.check(L().fe_zero2d(dev.ctx, Z.ptr, rows, 16, 16, 1))
    _fe.check(L().fe_stream_probe(dev.ctx, 0, nb, Z.ptr, None, None, D.ptr, grid))
    assert np.array_equal(D.download(), np.tile(np.float32([1, 2, 3, 4]), (rows, 1)))  # untouched
    assert L().fe_stream_probe(dev.ctx, 2, nb + 4, Ad.ptr, None, None, D.ptr, grid) == -2
    assert L().fe_stream_probe(dev.ctx, 3, nb, Ad.ptr, None, Cd.ptr, D.ptr, grid) == -2


def test_axpby_copy_upload(dev):
    rng = np.random.default_rng(12)
    n, k = 1000, 9
    X = rng.uniform(-1, 1, (n, k))
    Y = rng.uniform(-1, 1, (n, k))
    s = rng.uniform(-1, 1, k)
    Xd, Yd, sd = dev.put(X), dev.put(Y), dev.put(s)
    _fe.check(L().fe_axpby_cols(dev.ctx, _fe.F64, n, k, Xd.ptr, k, Yd.ptr, k, 0.5, 2.0, sd.ptr))
    assert np.max(np.abs(Yd.download() - (2.0 * Y + (0.5 + s) * X))) < 1e-14
    # column-major host <-> row-major device in both precisions
    F = np.asfortranarray(X)
    for prec, dt in ((_fe.F64, np.float64), (_fe.F32, np.float32)):
        B = dev.buffer((n, k), dt)
        _fe.check(L().fe_upload_colmajor(dev.ctx, n, k, F.ctypes.data, B.ptr, prec, k))
        assert np.array_equal(B.download(), X.astype(dt))
        back = np.zeros((n, k), order="F")
        _fe.check(L().fe_download_colmajor(dev.ctx, n, k, B.ptr, prec, k, back.ctypes.data))
        assert np.array_equal(back, X.astype(dt).astype(np.float64))
    B32 = dev.buffer((n, k), np.float32)
    _fe.check(L().fe_copy2d(dev.ctx, n, k, Xd.ptr, _fe.F64, k, B32.ptr, _fe.F32, k))
    assert np.array_equal(B32.download(), X.astype(np.float32))


@pytest.mark.parametrize("n,k,idx", [(1000, 48, [3, 0, 47, 5, 5]), (70001, 300, list(range(299, -1, -1))),
                                     (13, 4, [])])
@pytest.mark.parametrize("prec", [_fe.F64, _fe.F32])
def test_block_gather_cols(dev, n, k, idx, prec):
    rng = np.random.default_rng(n + k)
    dt = np.float64 if prec == _fe.F64 else np.float32
    X = rng.standard_normal((n, k)).astype(dt)
    Xd = dev.put(X, dt)
    p = len(idx)
    ld = p + 3
    Dd = dev.put(np.full((n, ld), -7.0, dt), dt)
    ia = (ctypes.c_int * max(p, 1))(*idx)
    _fe.check(L().fe_block_gather_cols(dev.ctx, prec, n, p, Xd.ptr, k, ia, Dd.ptr, ld))
    got = Dd.download()
    assert np.array_equal(got[:, :p], X[:, idx])
    assert np.all(got[:, p:] == -7.0)
    bad = (ctypes.c_int * 1)(k)
    assert L().fe_block_gather_cols(dev.ctx, prec, n, 1, Xd.ptr, k, bad, Dd.ptr, ld) == -2

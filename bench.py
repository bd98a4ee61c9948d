"""Benchmark of the block-eigensolver hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): the
Hamiltonian-like H = -0.5 * L27 + V on a 96^3 grid (n = 884,736, isotropic
27-point Laplacian, 200 Gaussian wells), lowest 256 eigenpairs by
Chebyshev-filtered subspace iteration, filter degree 20, tol 1e-8, FP64
residuals, adaptive filter bounds (cheb_bounds=ritz: the damped interval starts
at the largest Ritz value of the 512-column subspace). A "step" is one complete solve with the matrix and the seeded start
block already resident in HBM. `value` is mixed-precision (FP32 filter and
Gram, FP64 Cholesky / Rayleigh-Ritz / residuals / closing sweeps) solves per
second; the plain-FP64 GPU path is timed beside it.

Also measured: the fused Chebyshev step (the dominant HBM-bound kernel, in
the solve's own sparse format) with HIP events on the library's own stream,
against algorithmic bytes nnz*(s+4) + 4(n+1) + 3 n k s; its HBM traffic from
two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) calibrated on a 1 GiB
copy; per-pair eigenvalue errors against the CPU oracle's FP64 answer
(tests/golden/c3_eigs.json); the CPU oracle on the host cores (whole C1 and
32^3 solves timed, and the C3 solve extrapolated from its measured parts);
the generalized 64^3 problem in plain FP64; the serving stream (512 requests,
continuous batching vs one at a time in FP64); and the power-law C4 matrix
(its fused step in CSR and SELL, and its three solvers).

    python bench.py [--steps K] [--warmup W] [--config c3|c2|small|c5] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_matrix(cfg):
    from fasteig_amd import generators as G
    if cfg == "c3":
        return G.hamiltonian_27pt(96, seed=4), dict(nev=256, degree=20, tol=1e-8, blksize=0)
    if cfg == "c2":
        K, M = G.conductivity_3d(64, seed=2, mass_seed=3)
        return (K, M), dict(nev=64, degree=16, tol=1e-8, blksize=96)
    if cfg == "c4":  # the power-law matrix of c4_bench, for the kernel-only timings
        return G.power_law(2_000_000, seed=5), dict(nev=32, degree=20, tol=1e-8, blksize=48)
    # small smoke-sized variant of c3
    return G.hamiltonian_27pt(32, seed=4, wells=30), dict(nev=32, degree=20, tol=1e-8, blksize=0)


KERNEL_OF_FORMAT = {"csr": "k_spmm_csr", "sell": "k_spmm_sell", "blocked": "k_spmm_rb", "tiled": "k_spmm_tile"}


def spmm_roofline(A, k, prec, reps=20, fmt="csr"):
    """Average time of the fused Chebyshev step Y = 2c A X + 2e X - Z on the
    library stream (HIP events), its algorithmic bytes and the kernel's name
    (the stencil form of the tiled copy reads the diagonal and 27
    coefficients instead of the CSR arrays). fmt
    "blocked" times the row-block copy (4-row blocks sharing a column list),
    "tiled" the LDS-tiled copy (on a grid 8 x 4 or 4 x 4 plane patches,
    pencils through the planes; "tiled16" / "tiled32" force the height)."""
    from fasteig_amd import _fe
    L = _fe.lib()
    dev = _fe.Device(0)
    C = dev.csr(A)
    if fmt == "sell":
        C.build_sell(32, 256)
        C.use(_fe.FMT_SELL)
    if fmt.startswith("blocked"):
        built, reuse = C.build_blocked(8 if fmt == "blocked8" else 4)
        if not built:
            raise RuntimeError(f"row-block copy not built (reuse {reuse:.2f})")
        C.block_schedule()  # strip order for grid stencils, as the solvers use it
        C.use(_fe.FMT_BLOCKED)
    kernel = KERNEL_OF_FORMAT[fmt.rstrip("0123456789")]
    if fmt in ("tiled", "tiled16", "tiled32"):
        # tiled16 / tiled32: entry-list tiles of that height forced
        # (FE_TILE_STENCIL=0, FE_TILE_ROWS), for comparison with the stencil form
        saved = {v: os.environ.get(v) for v in ("FE_TILE_ROWS", "FE_TILE_STENCIL")}
        if fmt != "tiled":
            os.environ["FE_TILE_ROWS"] = fmt[5:]
            os.environ["FE_TILE_STENCIL"] = "0"
        try:
            built, lpr = C.build_tiled()
        finally:
            for v, old in saved.items():
                if old is None:
                    os.environ.pop(v, None)
                else:
                    os.environ[v] = old
        if not built:
            raise RuntimeError(f"tiled copy not built ({lpr:.2f} X rows loaded per row)")
        C.use(_fe.FMT_TILED)
        stencil = C.tiled_kind() == 64
        if stencil:
            kernel = "k_spmm_stencil"
    n = A.shape[0]
    dt = np.float64 if prec == _fe.F64 else np.float32
    s = np.dtype(dt).itemsize
    X = dev.buffer((n, k), dt)
    Y = dev.buffer((n, k), dt)
    rng = np.random.default_rng(0)
    X.upload(rng.uniform(-1, 1, (n, k)).astype(dt))
    Y.upload(rng.uniform(-1, 1, (n, k)).astype(dt))
    for _ in range(3):
        _fe.check(L.fe_spmm(dev.ctx, C.h, prec, k, X.ptr, k, Y.ptr, k, 0.01, 0.02, Y.ptr, k, -1.0))
    dev.sync()
    _fe.check(L.fe_event_record(dev.ctx, 0))
    for _ in range(reps):
        _fe.check(L.fe_spmm(dev.ctx, C.h, prec, k, X.ptr, k, Y.ptr, k, 0.01, 0.02, Y.ptr, k, -1.0))
    _fe.check(L.fe_event_record(dev.ctx, 1))
    import ctypes
    ms = ctypes.c_float()
    _fe.check(L.fe_event_elapsed_ms(dev.ctx, 0, 1, ctypes.byref(ms)))
    per = ms.value / reps
    nnz = A.nnz
    if kernel == "k_spmm_stencil":  # no entry lists: the diagonal and 27 coefficients
        algo = (n + 27) * s + 3 * n * k * s
    else:
        algo = nnz * (s + 4) + 4 * (n + 1) + 3 * n * k * s
    del X, Y, C
    dev.close()
    return per, algo, kernel


MFMA_PEAK_TFS = {"fp32": 157.3, "fp64": 78.6, "bf16": 2500.0}


def dense_roofline(n, m=512, reps=10):
    """MFMA rate of the dense products of a Chebyshev cycle at the solve's
    shape (n rows, m = 2 nev columns): the symmetric Gram Q^T Q (upper
    triangle of tiles: n m^2 flops), the Gram Q^T (AQ) (2 n m^2) and the
    rotation Q C (2 n m^2), each in FP32 and FP64; HIP events on the library
    stream, average of `reps` launches. Returns {name: (ms, TF/s, fraction of
    the MFMA peak)}."""
    import ctypes
    from fasteig_amd import _fe
    L = _fe.lib()
    dev = _fe.Device(0)
    out = {}
    rng = np.random.default_rng(3)
    Cm = dev.put(np.asfortranarray(rng.uniform(-1, 1, (m, m))), np.float64, order="F")
    G = dev.buffer((m, m), np.float64, order="F")
    for pname, prec, dt in (("fp32", _fe.F32, np.float32), ("fp64", _fe.F64, np.float64)):
        X = dev.buffer((n, m), dt)
        Yb = dev.buffer((n, m), dt)
        X.upload(rng.uniform(-1, 1, (n, m)).astype(dt))
        Yb.upload(rng.uniform(-1, 1, (n, m)).astype(dt))
        cases = {
            "gram_sym": (lambda: L.fe_gram(dev.ctx, _fe.GRAM_NATIVE, prec, n, m, m, X.ptr, m, X.ptr, m, None, G.ptr, m),
                         float(n) * m * m),
            "gram": (lambda: L.fe_gram(dev.ctx, _fe.GRAM_NATIVE, prec, n, m, m, X.ptr, m, Yb.ptr, m, None, G.ptr, m),
                     2.0 * n * m * m),
            "gemm": (lambda: L.fe_gemm_nn(dev.ctx, prec, n, m, m, X.ptr, m, Cm.ptr, m, Yb.ptr, m, 1.0, 0.0),
                     2.0 * n * m * m),
            # the CholQR product X R^-1 (upper triangular): n m^2 useful flops
            "trmm": (lambda: L.fe_trmm_upper(dev.ctx, prec, n, m, X.ptr, m, Cm.ptr, m, Yb.ptr, m),
                     float(n) * m * m),
        }
        if prec == _fe.F32:
            # compensated BF16 products: useful (FP32-equivalent) flops; the
            # fraction is of the BF16 MFMA peak over the flops executed
            for mname, mode, nprod in (("bf16x3", _fe.GRAM_BF16X3, 3), ("bf16x6", _fe.GRAM_BF16X6, 6)):
                cases[f"gram_sym_{mname}"] = (
                    lambda mode=mode: L.fe_gram(dev.ctx, mode, prec, n, m, m, X.ptr, m, X.ptr, m, None, G.ptr, m),
                    float(n) * m * m, nprod)
                cases[f"gram_{mname}"] = (
                    lambda mode=mode: L.fe_gram(dev.ctx, mode, prec, n, m, m, X.ptr, m, Yb.ptr, m, None, G.ptr, m),
                    2.0 * n * m * m, nprod)
                cases[f"gemm_{mname}"] = (
                    lambda mode=mode: L.fe_gemm_nn_mode(dev.ctx, mode, prec, n, m, m, X.ptr, m, Cm.ptr, m, Yb.ptr,
                                                        m, 1.0, 0.0),
                    2.0 * n * m * m, nprod)
                cases[f"trmm_{mname}"] = (
                    lambda mode=mode: L.fe_trmm_upper_mode(dev.ctx, mode, prec, n, m, X.ptr, m, Cm.ptr, m, Yb.ptr,
                                                           m),
                    float(n) * m * m, nprod)
        for name, case in cases.items():
            fn, flops = case[0], case[1]
            nprod = case[2] if len(case) > 2 else 0
            for _ in range(2):
                _fe.check(fn())
            dev.sync()
            _fe.check(L.fe_event_record(dev.ctx, 4))
            for _ in range(reps):
                _fe.check(fn())
            _fe.check(L.fe_event_record(dev.ctx, 5))
            ms = ctypes.c_float()
            _fe.check(L.fe_event_elapsed_ms(dev.ctx, 4, 5, ctypes.byref(ms)))
            per = ms.value / reps
            tfs = flops / (per * 1e-3) / 1e12
            peak_name = "bf16" if nprod else pname
            frac = tfs * max(nprod, 1) / MFMA_PEAK_TFS[peak_name]
            out[f"{name}_{pname}"] = {"ms": per, "tflops": tfs, "frac": frac}
            log(f"dense {name} {pname} n={n} m={m}: {per:.3f} ms {tfs:.1f} TF/s "
                f"({100 * frac:.0f}% of the {peak_name} peak {MFMA_PEAK_TFS[peak_name]})")
        del X, Yb
    del Cm, G
    dev.close()
    return out


def copy_reference(nbytes=1 << 30, reps=5):
    """A plain device copy of nbytes (fe_copy2d, FP32) used to calibrate the
    FETCH_SIZE / WRITE_SIZE counters: it reads and writes exactly nbytes."""
    import ctypes
    from fasteig_amd import _fe
    L = _fe.lib()
    dev = _fe.Device(0)
    rows = nbytes // (4 * 256)
    a = dev.buffer((rows, 256), np.float32)
    b = dev.buffer((rows, 256), np.float32)
    _fe.check(L.fe_zero2d(dev.ctx, a.ptr, rows, 256, 256, 4))
    _fe.check(L.fe_event_record(dev.ctx, 2))
    for _ in range(reps):
        _fe.check(L.fe_copy2d(dev.ctx, rows, 256, a.ptr, _fe.F32, 256, b.ptr, _fe.F32, 256))
    _fe.check(L.fe_event_record(dev.ctx, 3))
    ms = ctypes.c_float()
    _fe.check(L.fe_event_elapsed_ms(dev.ctx, 2, 3, ctypes.byref(ms)))
    del a, b
    dev.close()
    return ms.value / reps, rows * 256 * 4


def stream_ceiling(nbytes=2 << 30, reps=10, grids=(16384, 65536)):
    """HBM rates a plain streaming kernel reaches on this GPU (fe_stream_probe):
    read only, write only, copy, and three reads per write (the fused
    Chebyshev step's mix: X, Z and the matrix in, Y out). The best of the
    grid sizes, each timed over `reps` launches after one untimed."""
    import ctypes
    from fasteig_amd import _fe
    L = _fe.lib()
    dev = _fe.Device(0)
    rows = nbytes // (4 * 256)
    bufs = [dev.buffer((rows, 256), np.float32) for _ in range(4)]
    for b in bufs:
        _fe.check(L.fe_zero2d(dev.ctx, b.ptr, rows, 256, 256, 4))
    a, b, c, d = (x.ptr for x in bufs)
    moved = {0: 1, 1: 1, 2: 2, 3: 4}
    out = {}
    for kind, name in ((0, "read"), (1, "write"), (2, "copy"), (3, "read3_write1")):
        best = 0.0
        for grid in grids:
            _fe.check(L.fe_stream_probe(dev.ctx, kind, nbytes, a, b, c, d, grid))
            _fe.check(L.fe_event_record(dev.ctx, 2))
            for _ in range(reps):
                _fe.check(L.fe_stream_probe(dev.ctx, kind, nbytes, a, b, c, d, grid))
            _fe.check(L.fe_event_record(dev.ctx, 3))
            ms = ctypes.c_float()
            _fe.check(L.fe_event_elapsed_ms(dev.ctx, 2, 3, ctypes.byref(ms)))
            best = max(best, moved[kind] * nbytes / (ms.value / reps * 1e-3) / 1e9)
        out[name] = best
    del bufs
    dev.close()
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), r


def cpu_timed_solves(reps=3):
    """The CPU oracle's FP64 solves timed like the reference times a solve
    (src/fasteig_problem.cpp:389-416: the solver call only; the final residual
    recompute is outside), median of `reps` on all the host threads OpenMP
    gets and once on one thread, beside the HIP path's time for the same
    solve (median of `reps` after a warm-up):
      c1: 2D 5-point Laplacian 100 x 100, lowest 10, LOBPCG block 16, tol 1e-8
      c3_32: the C3 Hamiltonian family on a 32^3 grid (n = 32,768, 30 wells),
             lowest 32, Chebyshev degree 20 with adaptive bounds, tol 1e-8."""
    from fasteig_amd import generators as G
    from fasteig_amd.api import Problem
    from oracle import oracle as orc
    cores = orc.max_threads()
    cases = {
        "c1_lobpcg_block16": (G.laplacian_2d(100),
                              lambda A: orc.lobpcg(A, 10, block=16, maxiter=2000, tol=1e-8, which="sr", seed=1,
                                                   want_vectors=False),
                              dict(solver="LOBPCG", nev=10, blksize=16, tol=1e-8, maxiter=2000, which="SR", seed=1)),
        "c3_32cube_chebyshev": (G.hamiltonian_27pt(32, seed=4, wells=30),
                                lambda A: orc.chebyshev(A, 32, degree=20, maxiter=300, tol=1e-8, which="sr",
                                                        seed=1, bounds="ritz", want_vectors=False),
                                dict(solver="Chebyshev", nev=32, degree=20, tol=1e-8, maxiter=300, which="SR",
                                     seed=1, cheb_bounds="ritz")),
    }
    out = {"cpu": cpu_model(), "threads": cores}
    for name, (A, ofn, gopts) in cases.items():
        orc.set_threads(cores)
        t_all, r = _median_time(lambda: ofn(A), reps)
        orc.set_threads(1)
        t_one, r1 = _median_time(lambda: ofn(A), 1)
        orc.set_threads(cores)
        p = Problem(A)
        p.solve(**gopts)
        t_gpu, rg = _median_time(lambda: p.solve(**gopts), reps)
        p.close()
        out[name] = {"n": int(A.shape[0]), "cpu_s": t_all, "cpu_1thread_s": t_one, "cpu_iterations": r["iterations"],
                     "cpu_max_resid": float(r["residuals"].max()), "gpu_s": t_gpu, "gpu_iterations": rg.iterations,
                     "gpu_max_resid": float(rg.residual_norms.max()),
                     "max_rel_eig_diff": float(np.max(np.abs(rg.eigenvalues - r["eigenvalues"]) /
                                                      np.abs(r["eigenvalues"])))}
        log(f"cpu {name}: {t_all:.3f}s on {cores} threads, {t_one:.3f}s on 1 ({r['iterations']} iterations); "
            f"GPU {t_gpu:.4f}s ({rg.iterations})")
    return out


def cpu_c3_extrapolated(A, m, degree, cycles):
    """One FP64 oracle Chebyshev cycle at the C3 shape from measured parts,
    times the GPU FP64 solve's cycle count: per cycle `degree` + 1 sparse
    products on m columns (timed on 16 columns, scaled by m/16), two m x m
    Gram products and three n x m times m x m products (timed on an eighth of
    the rows, scaled by 8), and one m x m symmetric eigensolve (timed whole).
    The parts are the oracle's own routines on the host threads OpenMP gets.
    Returns (seconds per solve, detail)."""
    from oracle import oracle as orc
    n = A.shape[0]
    rng = np.random.default_rng(2)
    X16 = np.asfortranarray(rng.uniform(-1, 1, (n, 16)))
    t_sp, _ = _median_time(lambda: orc.spmm(A, X16), 3)
    nr = n // 8
    Q = np.asfortranarray(rng.uniform(-1, 1, (nr, m)))
    t_gram, _ = _median_time(lambda: orc.gram(Q, Q), 2)
    Cm = np.asfortranarray(rng.uniform(-1, 1, (m, m)))
    t_ac, _ = _median_time(lambda: orc.right_multiply(Q, Cm), 2)
    S = rng.standard_normal((m, m))
    t_ev, _ = _median_time(lambda: orc.syev(S + S.T), 1)
    cycle = (degree + 1) * t_sp * (m / 16) + 2 * 8 * t_gram + 3 * 8 * t_ac + t_ev
    detail = {"spmm_16col_s": t_sp, "gram_eighth_rows_s": t_gram, "rotation_eighth_rows_s": t_ac, "syev_s": t_ev,
              "cycle_s": cycle, "cycles": cycles}
    return cycles * cycle, detail


def measure_traffic(config, k, fmt, timeout=180):
    """HBM bytes per launch of the fused step (FP32, k columns, format fmt):
    two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch) over a child `bench.py --only-kernels` run, each scaled by the
    same counter on the 1 GiB reference copy in that run (a copy reads and
    writes exactly 1 GiB). Returns (bytes, detail) or (None, reason)."""
    import shutil
    import sqlite3
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    vals = {}
    with tempfile.TemporaryDirectory(prefix="fe_pmc_") as tmp:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, ctr)
            cmd = [prof, "--pmc", ctr, "--kernel-include-regex", "k_spmm|k_copy2d", "-d", out, "-o", "pmc", "--",
                   sys.executable, os.path.abspath(__file__), "--config", config, "--only-kernels",
                   "--kernel-k", str(k), "--kernel-precs", "fp32", "--kernel-formats", fmt]
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
            except subprocess.TimeoutExpired:
                return None, f"{ctr} pass timed out"
            dbs = [os.path.join(dp, f) for dp, _, fs in os.walk(out) for f in fs if f.endswith("_results.db")]
            if r.returncode != 0 or not dbs:
                return None, f"{ctr} pass failed (rc {r.returncode})"
            con = sqlite3.connect(dbs[0])
            rows = con.execute("select kernel_name, avg(value), count(*) from counters_collection "
                               "where counter_name = ? group by kernel_name", (ctr,)).fetchall()
            con.close()
            spmm = [v for name, v, _ in rows if "k_spmm" in name and "float" in name]
            copy = [v for name, v, _ in rows if "k_copy2d<float, float>" in name]
            if not spmm or not copy:
                return None, f"{ctr}: kernels not found in the counter database"
            vals[ctr] = (spmm[0], copy[0])
    gib = float(1 << 30)
    rd = vals["FETCH_SIZE"][0] * gib / vals["FETCH_SIZE"][1]
    wr = vals["WRITE_SIZE"][0] * gib / vals["WRITE_SIZE"][1]
    detail = {"fetch_kib": vals["FETCH_SIZE"][0], "write_kib": vals["WRITE_SIZE"][0],
              "copy_fetch_kib_per_gib": vals["FETCH_SIZE"][1], "copy_write_kib_per_gib": vals["WRITE_SIZE"][1],
              "read_bytes": rd, "write_bytes": wr}
    return rd + wr, detail


def c2_bench(steps=2):
    """Generalized K u = lambda M u, 7-point conductivity stencil on 64^3
    (n = 262,144), diagonal M, lowest 64 pairs, Chebyshev block 96 degree 16,
    tol 1e-8, plain FP64 on the GPU: seconds per solve and accuracy."""
    from fasteig_amd import generators as G
    from fasteig_amd.api import Problem
    K, M = G.conductivity_3d(64, seed=2, mass_seed=3)
    p = Problem(K, mass=M)
    opts = dict(solver="Chebyshev", nev=64, blksize=96, degree=16, tol=1e-8, which="SR", maxiter=300, seed=1,
                cheb_bounds="ritz", precision="fp64")
    p.solve(**opts)
    ts = []
    for _ in range(steps):
        t = time.perf_counter()
        r = p.solve(**opts)
        ts.append(time.perf_counter() - t)
    out = {"solves_per_s": steps / sum(ts), "s_per_solve": float(np.median(ts)), "iterations": r.iterations,
           "max_resid": float(r.residual_norms.max()), "status": r.status}
    log(f"c2 generalized fp64: {out['s_per_solve']:.3f} s/solve, {r.iterations} cycles, "
        f"max resid {out['max_resid']:.2e}")
    return out


def c4_bench(maxiter=200):
    """Power-law SPD matrix, n = 2,000,000 (row degrees alpha 2.1 on [4, 2000],
    symmetrized, diagonal = row abs-sum + 1), lowest 32 pairs, tol 1e-8: the
    fused step at k = 48 in CSR and SELL-32-sigma (load balance); Chebyshev
    block 48 with the mixed FP32 filter and FP64 refinement (the solver that
    converges here), and the two solvers the configuration names, LOBPCG
    block 48 and block Krylov-Schur (block 48, maxdim 192), in plain FP64 for
    `maxiter` iterations with their final residuals. Why those two stall is
    read off the spectrum: the wanted 32 eigenvalues fill a sliver of the
    interval [lambda_1, lambda_max] (relative width reported), and neither
    method has a preconditioner in the reference."""
    from fasteig_amd import _fe
    from fasteig_amd import generators as G
    from fasteig_amd.api import Problem
    t = time.time()
    A = G.power_law(2_000_000, seed=5)
    log(f"c4: n={A.shape[0]} nnz={A.nnz} built in {time.time() - t:.1f}s")
    out = {"n": A.shape[0], "nnz": int(A.nnz)}
    for fmt in ("csr", "sell"):
        for pname, prec in (("fp32", _fe.F32), ("fp64", _fe.F64)):
            per, algo, _ = spmm_roofline(A, 48, prec, fmt=fmt)
            out[f"step_{fmt}_{pname}_k48_ms"] = per
            out[f"step_{fmt}_{pname}_k48_gbs"] = algo / (per * 1e-3) / 1e9
            log(f"c4 fused step {fmt} {pname} k=48: {per:.3f} ms {algo / (per * 1e-3) / 1e9:.0f} GB/s")
    p = Problem(A)
    p.solve(solver="LOBPCG", nev=32, blksize=48, tol=1e-8, which="SR", maxiter=1, seed=1)  # upload, start block
    lo, hi = p.lanczos_bounds(1.0, 1)
    runs = (("chebyshev_mixed", dict(solver="Chebyshev", blksize=48, degree=20, precision="mixed",
                                     cheb_bounds="ritz", maxiter=400)),
            ("lobpcg_fp64", dict(solver="LOBPCG", blksize=48, precision="fp64", maxiter=maxiter)),
            ("block_krylov_schur_fp64", dict(solver="BlockKrylovSchur", blksize=48, maxdim=192, maxiter=maxiter)))
    for name, opts in runs:
        t = time.perf_counter()
        r = p.solve(nev=32, tol=1e-8, which="SR", seed=1, **opts)
        dt = time.perf_counter() - t
        out[name] = {"s": dt, "iterations": r.iterations, "status": r.status,
                     "converged": bool(r.residual_norms.max() <= 1e-8),
                     "max_resid": float(r.residual_norms.max()), "lowest": float(r.eigenvalues[0]),
                     "s_per_iteration": dt / max(1, r.iterations)}
        if name == "chebyshev_mixed":
            ev = r.eigenvalues
            out["spectrum"] = {"lambda_1": float(ev[0]), "lambda_32": float(ev[-1]), "lanczos_upper": hi,
                               "wanted_relative_width": float((ev[-1] - ev[0]) / (hi - ev[0]))}
        log(f"c4 {name}: {dt:.2f}s, {r.iterations} iterations, status {r.status}, "
            f"max resid {r.residual_norms.max():.2e}")
    out["stall_reason"] = ("the 32 wanted eigenvalues fill a relative width of "
                           f"{out['spectrum']['wanted_relative_width']:.1e} of [lambda_1, lambda_max]; unpreconditioned "
                           "LOBPCG and block Krylov-Schur converge at a rate set by that gap ratio, the "
                           "Chebyshev filter with adaptive bounds damps the rest of the interval directly")
    p.close()
    return out


def serving_bench(count=512, rate=200.0, baseline=None, degree=16, max_slots=128, seed=11):
    """Serving workload: `count` requests (2D 5-point / 3D 7-point Laplacian +
    N(0,1) diagonal, n log-uniform in [1e3, 2e5], nev in {8,16,32}, tol in
    {1e-6,1e-8}) with Poisson arrivals at `rate`/s, served by the
    continuous-batching server in real time, and again with every request
    waiting at time 0 (the server's capacity); then the same requests (the
    first `baseline` when given) one at a time with the plain-FP64 Chebyshev
    solver (same degree, ritz bounds), their latencies from a FIFO queue over
    the measured solve times at the same arrivals. Eigenvectors are returned
    by both."""
    from fasteig_amd import generators as G
    from fasteig_amd.api import Problem
    from fasteig_amd.serving import Server
    stream = G.serving_stream(count, seed=seed, rate=rate)
    t = time.time()
    mats = [G.stencil_with_potential(r["kind"], r["n_target"], r["seed"]) for r in stream]
    log(f"serving: {count} requests, n {min(A.shape[0] for A in mats)}..{max(A.shape[0] for A in mats)}, "
        f"built in {time.time() - t:.1f}s")
    # warm the kernels and the graph on two small requests
    with Server(max_slots=max_slots, degree=degree) as w:
        for j in range(2):
            w.submit(10**6 + j, mats[j], stream[j]["nev"], stream[j]["tol"])
        while w.pending():
            w.step()
            w.results()
    def run_batched(arrival):
        """Serve every request in real time, request i submitted once
        arrival[i] seconds have passed; latencies from arrival to result."""
        srv = Server(max_slots=max_slots, degree=degree)
        lat, status, cycles = {}, {}, {}
        t0 = time.perf_counter()
        i = 0
        while len(lat) < count:
            now = time.perf_counter() - t0
            while i < count and arrival[i] <= now:
                srv.submit(i, mats[i], stream[i]["nev"], stream[i]["tol"])
                i += 1
            active, queued, _ = srv.counts()
            if active == 0 and queued == 0 and i < count:
                time.sleep(max(0.0, arrival[i] - now))
                continue
            srv.step()
            for r in srv.results():
                lat[r.id] = time.perf_counter() - t0 - arrival[r.id]
                status[r.id] = r.status
                cycles[r.id] = r.cycles
        span = time.perf_counter() - t0 - arrival[0]
        st = srv.stats()
        srv.close()
        lats = np.array([lat[j] for j in range(count)])
        return {"requests": count, "solves_per_s": count / span,
                "p50_s": float(np.percentile(lats, 50)), "p99_s": float(np.percentile(lats, 99)),
                "converged": int(sum(1 for v in status.values() if v == 0)), "cycles": st["cycles"],
                "peak_active": st["peak_active"], "peak_hbm_gb": st["peak_bytes"] / 1e9,
                "mean_problem_cycles": float(np.mean(list(cycles.values())))}

    out = run_batched([r["arrival"] for r in stream])
    out["arrival_rate_per_s"] = rate
    log(f"serving batched: {out['solves_per_s']:.1f} solves/s p50 {out['p50_s']:.3f}s p99 {out['p99_s']:.3f}s "
        f"converged {out['converged']}/{count}, {out['cycles']} cycles, peak {out['peak_active']} active")
    # capacity: all requests waiting at time 0
    sat = run_batched([0.0] * count)
    log(f"serving batched, all at t=0: {sat['solves_per_s']:.1f} solves/s p50 {sat['p50_s']:.3f}s "
        f"p99 {sat['p99_s']:.3f}s converged {sat['converged']}/{count}")
    # one at a time, plain FP64
    durs = []
    p0 = None
    nb = count if baseline is None else min(baseline, count)
    for j in range(nb):
        if j and j % 128 == 0:
            log(f"serving one-at-a-time: {j}/{nb} solved")
        r = stream[j]
        ts = time.perf_counter()
        p = Problem(mats[j], share_with=p0)
        res = p.solve(solver="Chebyshev", nev=r["nev"], tol=r["tol"], degree=degree, which="SR", maxiter=300,
                      seed=1, cheb_bounds="ritz", precision="fp64")
        p.eigenvectors()
        durs.append(time.perf_counter() - ts)
        if p0 is None:
            p0 = p
    durs = np.array(durs)
    fin, lat1 = 0.0, []
    for j, d in enumerate(durs):
        fin = max(fin, stream[j]["arrival"]) + d
        lat1.append(fin - stream[j]["arrival"])
    one = {"requests": len(durs), "solves_per_s": len(durs) / durs.sum(),
           "p50_s": float(np.percentile(lat1, 50)), "p99_s": float(np.percentile(lat1, 99)),
           "mean_solve_s": float(durs.mean())}
    log(f"serving one-at-a-time fp64 ({len(durs)} requests): {one['solves_per_s']:.1f} solves/s "
        f"p50 {one['p50_s']:.3f}s p99 {one['p99_s']:.3f}s")
    # one-at-a-time with every request waiting at time 0: p99 of the FIFO queue
    one["p99_all_at_t0_s"] = float(np.percentile(np.cumsum(durs), 99))
    return {"batched": out, "batched_all_at_t0": sat, "one_at_a_time_fp64": one,
            "throughput_ratio": out["solves_per_s"] / one["solves_per_s"],
            "capacity_ratio": sat["solves_per_s"] / one["solves_per_s"]}


def trace_split(db, marks_path, top=28):
    """Per-kernel totals of a rocprofv3 kernel trace (its SQLite results
    database) inside each window written by --solve-trace: the time of a solve
    by kernel, its kernel count and the idle time between its kernels."""
    import collections
    import sqlite3
    marks = json.load(open(marks_path))
    rows = sqlite3.connect(db).execute("select name, start, end from kernels order by start").fetchall()
    out = []
    for name, (t0, t1) in marks.items():
        agg = collections.defaultdict(lambda: [0, 0.0])
        busy = gaps = 0.0
        prev_end = None
        for kname, s0, s1 in rows:
            if s0 >= t0 and s1 <= t1:
                k = kname.split("(")[0].replace("void ", "")
                agg[k][0] += 1
                agg[k][1] += (s1 - s0) / 1e6
                busy += (s1 - s0) / 1e6
                if prev_end is not None and s0 > prev_end:
                    gaps += (s0 - prev_end) / 1e6
                prev_end = s1 if prev_end is None else max(prev_end, s1)
        out.append(f"== {name}: window {(t1 - t0) / 1e6:.1f} ms, kernels {busy:.1f} ms, idle between kernels "
                   f"{gaps:.1f} ms")
        out.append(f"{'ms':>9} {'share':>6} {'count':>6}  kernel")
        for k, (c, ms) in sorted(agg.items(), key=lambda x: -x[1][1])[:top]:
            out.append(f"{ms:9.2f} {100 * ms / busy:5.1f}% {c:6d}  {k[:90]}")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c3", "c2", "small", "c4", "c5"])
    ap.add_argument("--gram", default="auto", choices=["auto", "native", "bf16x3", "bf16x6"])
    ap.add_argument("--format", default="auto", choices=["auto", "csr", "blocked", "tiled", "sell"],
                    help="sparse format of the solve (auto: LDS-tiled, else row-block, when the matrix allows)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--only-kernels", action="store_true",
                    help="time only the fused Chebyshev step and a reference copy (for counter runs)")
    ap.add_argument("--only-dense", action="store_true",
                    help="time only the Gram and rotation products at the solve's shape (for counter runs)")
    ap.add_argument("--kernel-k", default="", help="block widths timed by --only-kernels (default nev, 2 nev)")
    ap.add_argument("--kernel-precs", default="fp32,fp64", help="precisions timed by --only-kernels")
    ap.add_argument("--no-extra", action="store_true", help="skip the generalized (C2) and serving (C5) runs")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    ap.add_argument("--kernel-formats", default="csr,blocked,tiled",
                    help="sparse formats timed by --only-kernels (csr, blocked = 4-row blocks, blocked8, tiled, "
                         "tiled16, tiled32)")
    ap.add_argument("--solve-trace", default="", metavar="PRECISIONS",
                    help="for a kernel trace: per precision (e.g. mixed,fp64) one warm-up and one timed solve; "
                         "the timed solves' CLOCK_MONOTONIC windows (the trace's clock) go to --marks")
    ap.add_argument("--marks", default="solve_marks.json", help="output of --solve-trace")
    ap.add_argument("--trace-split", nargs=2, metavar=("DB", "MARKS"),
                    help="per-kernel time inside each --solve-trace window of a rocprofv3 results database")
    args = ap.parse_args()
    if args.trace_split:
        print(trace_split(*args.trace_split))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from fasteig_amd import _fe
    from fasteig_amd.api import Problem

    t0 = time.time()
    if args.config == "c5":
        print(json.dumps(serving_bench()), flush=True)
        return
    if args.config == "c4" and not (args.only_kernels or args.only_dense):
        print(json.dumps(c4_bench()), flush=True)
        return
    if args.solve_trace:
        M, cfg = build_matrix(args.config)
        A = M[0] if isinstance(M, tuple) else M
        p = Problem(A)
        base = dict(solver="Chebyshev", nev=cfg["nev"], degree=cfg["degree"], tol=cfg["tol"], which="SR",
                    maxiter=300, seed=1, cheb_bounds="ritz", format=args.format, gram=args.gram)
        marks = {}
        for prec in args.solve_trace.split(","):
            p.solve(precision=prec, **base)
            t_0 = time.monotonic_ns()
            r = p.solve(precision=prec, **base)
            marks[prec] = [t_0, time.monotonic_ns()]
            log(f"trace solve {prec}: {(marks[prec][1] - t_0) / 1e9:.3f} s, {r.iterations} cycles, "
                f"max residual {r.residual_norms.max():.2e}")
        json.dump(marks, open(args.marks, "w"))
        return
    if args.only_dense:
        M, cfg = build_matrix(args.config)
        A = M[0] if isinstance(M, tuple) else M
        print(json.dumps(dense_roofline(A.shape[0], 2 * cfg["nev"] if not cfg["blksize"] else cfg["blksize"])),
              flush=True)
        return
    if args.only_kernels:
        M, cfg = build_matrix(args.config)
        A = M[0] if isinstance(M, tuple) else M
        out = {}
        ks = [int(x) for x in args.kernel_k.split(",")] if args.kernel_k else [cfg["nev"], 2 * cfg["nev"]]
        for kk in ks:
            for fmt in args.kernel_formats.split(","):
                for pname, prec in (("fp32", _fe.F32), ("fp64", _fe.F64)):
                    if pname not in args.kernel_precs.split(","):
                        continue
                    per, algo, _ = spmm_roofline(A, kk, prec, fmt=fmt)
                    out[f"{fmt}_{pname}_k{kk}_ms"] = per
                    out[f"{fmt}_{pname}_k{kk}_gbs"] = algo / (per * 1e-3) / 1e9
                    log(f"fused step {fmt} {pname} k={kk}: {per:.3f} ms {algo / (per * 1e-3) / 1e9:.0f} GB/s")
        copy_ms, copy_bytes = copy_reference()
        out.update(copy_ms=copy_ms, copy_bytes_each_way=copy_bytes)
        print(json.dumps(out), flush=True)
        return
    M, cfg = build_matrix(args.config)
    mass = None
    if isinstance(M, tuple):
        M, mass = M
    A = M
    n = A.shape[0]
    log(f"matrix {args.config}: n={n} nnz={A.nnz} built in {time.time() - t0:.1f}s")
    p = Problem(A, mass=mass)
    base = dict(solver="Chebyshev", nev=cfg["nev"], degree=cfg["degree"], tol=cfg["tol"], which="SR",
                maxiter=300, seed=1, cheb_bounds="ritz", format=args.format)
    if cfg["blksize"]:
        base["blksize"] = cfg["blksize"]
    modes = {"mixed": dict(base, precision="mixed", gram=args.gram)}
    if not args.no_fp64:
        modes["fp64"] = dict(base, precision="fp64")
    lat = {m: [] for m in modes}
    res = {}
    cold_s = None
    for m, opts in modes.items():
        for w in range(args.warmup):
            t = time.perf_counter()
            r = p.solve(**opts)
            if cold_s is None:  # matrix upload, format build, seeded start block on the host
                cold_s = time.perf_counter() - t
            log(f"warmup {m}: {time.perf_counter() - t:.3f}s iters={r.iterations} "
                f"maxres={r.residual_norms.max():.2e}")
        for s in range(args.steps):
            t = time.perf_counter()
            r = p.solve(**opts)
            lat[m].append(time.perf_counter() - t)
            log(f"step {m} {s}: {lat[m][-1]:.3f}s iters={r.iterations} maxres={r.residual_norms.max():.2e}")
        res[m] = r
    mixed_total = sum(lat["mixed"])
    value = args.steps / mixed_total
    out_acc = {"max_resid_mixed": float(res["mixed"].residual_norms.max()),
               "max_scaled_resid_mixed": float(res["mixed"].scaled_residual_norms.max())}
    extra = {}
    if "fp64" in res:
        e64 = res["fp64"].eigenvalues
        rel = np.abs(res["mixed"].eigenvalues - e64) / np.abs(e64)
        out_acc.update(max_rel_eig_err_mixed_vs_fp64=float(rel.max()),
                       max_resid_fp64=float(res["fp64"].residual_norms.max()))
        extra["fp64_solves_per_s"] = args.steps / sum(lat["fp64"])
        extra["mixed_over_fp64"] = extra["fp64_solves_per_s"] and value / extra["fp64_solves_per_s"]
        extra["fp64_iterations"] = res["fp64"].iterations
    extra["mixed_iterations"] = res["mixed"].iterations
    if args.config == "c3":
        # per-pair error against the oracle's FP64 eigenvalues (run to 1e-12)
        ref = np.array(json.load(open(os.path.join(ROOT, "tests", "golden", "c3_eigs.json")))["eigenvalues"])
        for m, r in res.items():
            err = np.abs(r.eigenvalues - ref) / np.abs(ref)
            out_acc[f"max_rel_eig_err_{m}_vs_oracle"] = float(err.max())
            out_acc[f"median_rel_eig_err_{m}_vs_oracle"] = float(np.median(err))
            out_acc[f"rel_eig_err_{m}_vs_oracle_per_pair"] = [float(f"{e:.2e}") for e in err]
    extra["cold_first_solve_s"] = cold_s
    # product modes the automatic choice measured (time and error of each)
    extra["product_tuning"] = p.product_tuning()
    del p

    # dominant kernel: the fused Chebyshev step at the width the solve's FP32
    # cycles filter (the whole subspace: blksize, or 2 nev), in FP32; the
    # step at k = nev is reported beside it
    k = cfg["blksize"] or min(n, 2 * cfg["nev"])
    # the solve's own sparse format: what "auto" picks (tiled, else row-block, else CSR)
    for fmt in {"auto": ["tiled", "blocked", "csr"], "tiled": ["tiled"], "blocked": ["blocked"],
                "csr": ["csr"], "sell": ["sell"]}[args.format]:
        try:
            per32, algo32, kname = spmm_roofline(A, k, _fe.F32, fmt=fmt)
            break
        except RuntimeError:
            continue
    per64, algo64, _ = spmm_roofline(A, k, _fe.F64, fmt=fmt)
    achieved = algo32 / (per32 * 1e-3) / 1e9
    extra["fused_step_k"] = k
    extra["fused_step_fp32_ms"] = per32
    extra["fused_step_fp64_ms"] = per64
    extra["fused_step_fp64_gbs"] = algo64 / (per64 * 1e-3) / 1e9
    log(f"fused step fp32 k={k}: {per32:.3f} ms, {achieved:.0f} GB/s; fp64 {per64:.3f} ms "
        f"{extra['fused_step_fp64_gbs']:.0f} GB/s")
    kn = cfg["nev"]
    if kn != k:
        p32n, a32n, _ = spmm_roofline(A, kn, _fe.F32, fmt=fmt)
        p64n, a64n, _ = spmm_roofline(A, kn, _fe.F64, fmt=fmt)
        extra[f"fused_step_k{kn}"] = {"fp32_ms": p32n, "fp32_gbs": a32n / (p32n * 1e-3) / 1e9,
                                      "fp32_frac": a32n / (p32n * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                      "fp64_ms": p64n, "fp64_gbs": a64n / (p64n * 1e-3) / 1e9}
        log(f"fused step k={kn}: fp32 {p32n:.3f} ms {a32n / (p32n * 1e-3) / 1e9:.0f} GB/s; "
            f"fp64 {p64n:.3f} ms {a64n / (p64n * 1e-3) / 1e9:.0f} GB/s")
    # the cycle's Gram and rotation products at the solve's block width
    extra["dense_mfma"] = dense_roofline(n, cfg["blksize"] or min(n, 2 * cfg["nev"]))
    extra["stream_ceiling_gbs"] = stream_ceiling()
    log("streaming ceilings (GB/s): " + ", ".join(f"{k} {v:.0f}" for k, v in extra["stream_ceiling_gbs"].items()))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        cores = orc.max_threads()
        m_sub = cfg["blksize"] or min(n, 2 * cfg["nev"])
        cycles = res.get("fp64", res["mixed"]).iterations
        t_solve, cdet = cpu_c3_extrapolated(A, m_sub, cfg["degree"], cycles)
        timed = cpu_timed_solves()
        cpu = {"value": 1.0 / t_solve, "unit": "solves/s", "cores": cores, "kind": "port",
               "sample": f"EXTRAPOLATED: one FP64 oracle Chebyshev cycle at this shape from measured parts "
                         f"({cfg['degree'] + 1} sparse products on {m_sub} columns timed on 16, two Gram and "
                         f"three rotation products timed on an eighth of the rows, one {m_sub} x {m_sub} "
                         f"eigensolve) x the plain-FP64 GPU solve's {cycles} cycles; timed whole solves "
                         f"(median of 3, all threads and 1 thread) in cpu_timed",
               "cpu": timed["cpu"], "detail": cdet, "cpu_timed": timed}
        log(f"cpu C3 extrapolated: {t_solve:.1f}s per solve on {cores} threads ({cdet})")

    if not args.no_extra:
        extra["c2_generalized_fp64"] = c2_bench()
        extra["c5_serving"] = serving_bench()
        extra["c4_power_law"] = c4_bench()
    traffic, tdetail = (None, "skipped") if args.no_traffic else measure_traffic(args.config, k, fmt)
    extra["traffic_detail"] = tdetail
    log(f"traffic: {traffic} ({tdetail})")

    line = {
        "metric": "solves/s + p50/p99 latency at resid<=1e-8, mixed vs plain FP64; SpMM HBM GB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": {"c3": "H=-0.5*L27+V 96^3 (n=884736), nev 256, Chebyshev degree 20, tol 1e-8",
                                "c2": "K u = l M u 7-pt 64^3, nev 64, block 96, degree 16",
                                "small": "H=-0.5*L27+V 32^3, nev 32, degree 20"}[args.config],
                   "n": n, "nnz": int(A.nnz), "gram": args.gram, "format": args.format,
                   "timed": "solve only: matrix and the seeded start block resident on the device between "
                            "steps (the first, cold solve is extra.cold_first_solve_s)"},
        "latency_s": {m: {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99))}
                      for m, v in lat.items()},
        "accuracy": out_acc,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"{kname} fused Chebyshev step FP32 k={k}",
                     "algo_bytes": algo32},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

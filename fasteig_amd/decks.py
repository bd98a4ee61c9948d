"""Input decks for the benchmark configurations.

`write_deck(config, directory)` writes the configuration's matrices as Matrix
Market files (coordinate real symmetric, lower triangle, 17 significant
digits) and a runfasteig `.input` deck in the reference's key/value format
(src/parseinput.cpp:171-280), so each configuration runs as

    fasteig_amd/lib/runfasteig_hip <directory>/<config>.input

Configurations (matrices from fasteig_amd.generators, seeded):
  c1  2D 5-point Laplacian 100 x 100, LOBPCG block 16, lowest 10, tol 1e-8
  c2  K u = lambda M u: 7-point conductivity stencil 64^3 with diagonal M,
      Chebyshev block 96 degree 16, lowest 64, tol 1e-8
  c3  H = -0.5 L27 + V on 96^3, Chebyshev degree 20, lowest 256, tol 1e-8,
      mixed FP32/FP64
  c4  power-law SPD matrix n = 2,000,000, block Krylov-Schur block 48
      maxdim 192, lowest 32, tol 1e-8 (a text file of about 2 GB)

Usage: python -m fasteig_amd.decks <c1|c2|c3|c4> <directory> [--small]
[--rsa | --fei] (--small writes the same deck on a reduced grid, for quick
checks; --rsa / --fei write the reference's compressed-column formats).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp

from . import generators as G


def write_mtx(path: str, A, symmetric: bool = True) -> None:
    """Matrix Market coordinate file; for symmetric=True only the lower
    triangle (row >= column) is written, as the reference's readers expect."""
    T = sp.coo_matrix(A)
    n, m = T.shape
    if symmetric:
        keep = T.row >= T.col
        rows, cols, vals = T.row[keep], T.col[keep], T.data[keep]
    else:
        rows, cols, vals = T.row, T.col, T.data
    order = np.lexsort((rows, cols))  # column-major, like mmio writers
    rows, cols, vals = rows[order], cols[order], vals[order]
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real " + ("symmetric" if symmetric else "general") + "\n")
        f.write(f"{n} {m} {len(vals)}\n")
        block = 1 << 20
        for s in range(0, len(vals), block):
            r, c, v = rows[s:s + block] + 1, cols[s:s + block] + 1, vals[s:s + block]
            f.write("\n".join(f"{a} {b} {x:.17g}" for a, b, x in zip(r.tolist(), c.tolist(), v.tolist())))
            f.write("\n")


def _csc_full(A):
    C = sp.csc_matrix(A)
    C.sort_indices()
    return C.shape[0], C.indptr.astype(np.int64), C.indices.astype(np.int64), C.data.astype(np.float64)


def write_rsa(path: str, A) -> None:
    """RSA text file as the reference's readRSA reads it (src/parseinput.cpp:27-67):
    five header lines (the third "RSA nrows ncols nnz 0"), then 1-based column
    pointers, row indices and values. All stored entries form the matrix
    (full storage, both triangles)."""
    n, cp, ri, v = _csc_full(A)
    with open(path, "w") as f:
        f.write("fasteig_amd matrix\n")
        f.write(f"{len(cp) + len(ri) + len(v)} {len(cp)} {len(ri)} {len(v)}\n")
        f.write(f"RSA {n} {n} {len(v)} 0\n")
        f.write("(10I8) (10I8) (4E26.17)\n\n")
        for arr, fmt in ((cp + 1, "{}"), (ri + 1, "{}"), (v, "{:.17g}")):
            vals = arr.tolist()
            for s0 in range(0, len(vals), 10):
                f.write(" ".join(fmt.format(x) for x in vals[s0:s0 + 10]) + "\n")


def write_fei(path: str, A) -> None:
    """FEI binary file (reference readFEI, src/parseinput.cpp:69-111): int32
    nrows, ncols, nnz; int32 colptr (0-based); int32 rowind; float64 values;
    little-endian; full storage."""
    n, cp, ri, v = _csc_full(A)
    with open(path, "wb") as f:
        np.array([n, n, len(v)], dtype="<i4").tofile(f)
        cp.astype("<i4").tofile(f)
        ri.astype("<i4").tofile(f)
        v.astype("<f8").tofile(f)


def _config(name: str, small: bool):
    if name == "c1":
        g = 20 if small else 100
        return ({"a": G.laplacian_2d(g)},
                dict(solver="LOBPCG", nev=10, blksize=16, tol=1e-8, which="SR", maxiter=2000, seed=1))
    if name == "c2":
        K, M = G.conductivity_3d(16 if small else 64, seed=2, mass_seed=3)
        return ({"a": K, "b": sp.diags(M).tocsr()},
                dict(solver="Chebyshev", nev=16 if small else 64, blksize=24 if small else 96, degree=16, tol=1e-8,
                     which="SR", maxiter=300, seed=1, cheb_bounds="ritz"))
    if name == "c3":
        g = 24 if small else 96
        return ({"a": G.hamiltonian_27pt(g, seed=4)},
                dict(solver="Chebyshev", nev=32 if small else 256, degree=20, tol=1e-8, which="SR", maxiter=300,
                     seed=1, cheb_bounds="ritz", precision="mixed"))
    if name == "c4":
        return ({"a": G.power_law(20_000 if small else 2_000_000, seed=5)},
                dict(solver="BlockKrylovSchur", nev=32, blksize=48, maxdim=192, tol=1e-8, which="SR", maxiter=400,
                     seed=1))
    raise ValueError(f"unknown configuration {name!r} (c1, c2, c3, c4)")


def write_deck(name: str, directory: str, small: bool = False, fmt: str = "mtx") -> str:
    """Write the matrices (fmt: mtx, rsa or fei) and the deck of configuration
    `name`; returns the deck's path."""
    mats, opts = _config(name, small)
    os.makedirs(directory, exist_ok=True)
    lines = [f"# configuration {name}" + (" (reduced grid)" if small else "")]
    writers = {"mtx": lambda path, A: write_mtx(path, A, symmetric=True), "rsa": write_rsa, "fei": write_fei}
    if fmt not in writers:
        raise ValueError(f"unknown matrix format {fmt!r} (mtx, rsa, fei)")
    for key, A in mats.items():
        fname = f"{name}_{key}.{fmt}"
        writers[fmt](os.path.join(directory, fname), A)
        lines.append(f"{key}matrix {fname}")
    lines.append(f"logfile {name}")
    lines.append("output_format text")
    lines += [f"{k} {v}" for k, v in opts.items()]
    path = os.path.join(directory, f"{name}.input")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    if len(args) != 2:
        sys.exit("usage: python -m fasteig_amd.decks <c1|c2|c3|c4> <directory> [--small]")
    fmt = "rsa" if "--rsa" in sys.argv else "fei" if "--fei" in sys.argv else "mtx"
    print(write_deck(args[0], args[1], small="--small" in sys.argv, fmt=fmt))

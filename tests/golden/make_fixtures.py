"""Writes the golden fixtures under tests/golden/ from the reference's own test data.

Run in a container where the reference tree is mounted at /root/reference:
    python tests/golden/make_fixtures.py
Outputs (data only, no reference source):
  SiO.npz            0-based triplets of examples/Sparse_Matrices/SiO.mtx (stored
                     triangle, as read by the reference's MTX loader)
  sio_eigs.json      lowest 60 eigenvalues of tests/reference/SiO.out.121mpi
                     (ScaLAPACK PDSYEVX spectrum used by the reference CTest cases)
  deck_defaults.json deck defaults of tests/scripts/run_case.py:225-243 (k=5,
                     tol 1e-8, maxiter 2000, seed 1, blksize 10, maxdim 40) and the
                     acceptance rule of :246-260 (|dl| <= 1e-10 + 1e-10 |l|, res <= 1e-8)
"""
import json
import os
import re

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def sio_matrix():
    path = os.path.join(REF, "examples/Sparse_Matrices/SiO.mtx")
    with open(path) as f:
        header = f.readline()
    assert "symmetric" in header
    d = np.loadtxt(path, comments="%")
    n = int(d[0, 0])
    rows = d[1:, 0].astype(np.int32) - 1
    cols = d[1:, 1].astype(np.int32) - 1
    vals = d[1:, 2]
    np.savez_compressed(os.path.join(HERE, "SiO.npz"), n=n, rows=rows, cols=cols, vals=vals)


def sio_eigs(count=60):
    vals = []
    pat = re.compile(r"ev\(\s*(\d+)\)=\s*([-+0-9.Ee]+)")
    with open(os.path.join(REF, "tests/reference/SiO.out.121mpi")) as f:
        for line in f:
            m = pat.search(line)
            if m:
                vals.append((int(m.group(1)), float(m.group(2))))
    vals.sort()
    with open(os.path.join(HERE, "sio_eigs.json"), "w") as f:
        json.dump({"source": "tests/reference/SiO.out.121mpi", "eigenvalues": [v for _, v in vals[:count]]}, f,
                  indent=1)


def deck_defaults():
    out = {"source": "tests/scripts/run_case.py:225-260",
           "nev": 5, "tol": 1e-8, "maxiter": 2000, "seed": 1, "blksize": 10, "maxdim": 40, "which": "SR",
           "eig_abs_tol": 1e-10, "eig_rel_tol": 1e-10, "residual_tol": 1e-8}
    with open(os.path.join(HERE, "deck_defaults.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    sio_matrix()
    sio_eigs()
    deck_defaults()

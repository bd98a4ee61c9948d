"""Writes tests/golden/c4_powerlaw_20k.npz: the oracle's eigenpairs of the C4
power-law family at n = 20,000 (fasteig_amd.generators.power_law(20000, seed=5):
row degrees from a truncated power law, alpha 2.1 on [4, 2000], N(0,1) values,
symmetrized, diagonal = row abs-sum + 1), lowest 8, seed 1:

  LOBPCG, block 12, tol 1e-12 (the reference's LOBPCGSolver)   eigenvalues, vectors, iterations
  block Krylov-Schur, block 12, maxdim 96, tol 1e-10           eigenvalues, restarts

(block Krylov-Schur's residual estimates level off near 2e-12 here, so it is
run to 1e-10).

The oracle needs about a minute for these (LOBPCG takes ~1,100 iterations on
this spectrum), too long to rerun inside every GPU test, so the GPU parity test
reads this file. Run from the repository root:

    python tests/golden/make_c4_fixture.py [threads]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fasteig_amd import generators as G  # noqa: E402
from oracle import oracle as orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_powerlaw_20k.npz")


def main():
    if len(sys.argv) > 1:
        orc.set_threads(int(sys.argv[1]))
    A = G.power_law(20000, seed=5)
    t = time.time()
    lo = orc.lobpcg(A, 8, block=12, maxiter=3000, tol=1e-12, which="sr", seed=1)
    print(f"LOBPCG: {lo['iterations']} iterations, max relres {lo['residuals'].max():.2e}, {time.time() - t:.0f} s")
    t = time.time()
    bk = orc.block_krylov_schur(A, 8, block=12, maxdim=96, maxiter=3000, tol=1e-10, which="sr", seed=1,
                                want_vectors=False)
    print(f"BKS: {bk['iterations']} restarts, max relres {bk['residuals'].max():.2e}, {time.time() - t:.0f} s")
    assert lo["status"] == 0 and bk["status"] == 0
    np.savez_compressed(OUT, n=A.shape[0], nnz=A.nnz, data_sum=float(np.abs(A.data).sum()),
                        lobpcg_eigenvalues=lo["eigenvalues"], lobpcg_vectors=lo["eigenvectors"],
                        lobpcg_iterations=lo["iterations"], lobpcg_relres=lo["residuals"],
                        bks_eigenvalues=bk["eigenvalues"], bks_iterations=bk["iterations"],
                        bks_relres=bk["residuals"])
    print("wrote", OUT)


if __name__ == "__main__":
    main()

"""The runfasteig_hip input-deck driver on configuration decks written by
fasteig_amd.decks (reduced grids): exit status, the reference's RAW RESULTS
layout, and the eigenvalues against the closed form (C1) and against the
library's own solve of the same matrices (C2, generalized)."""
import os
import subprocess

import numpy as np
import pytest

from fasteig_amd import decks
from fasteig_amd import generators as G
from fasteig_amd.api import Problem

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(__file__)), "fasteig_amd", "lib", "runfasteig_hip")


def run_deck(deck, cwd):
    r = subprocess.run([EXE, deck], cwd=cwd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def raw_results(path):
    lines = open(path).read().splitlines()
    start = next(i for i, l in enumerate(lines) if l.split() == ["Index", "Eigenvalue", "ResidualNorm"])
    rows = []
    for l in lines[start + 1:]:
        parts = l.split()
        if len(parts) != 3 or not parts[0].isdigit():
            break
        rows.append((int(parts[0]), float(parts[1]), float(parts[2])))
    return rows


@pytest.mark.parametrize("fmt", ["mtx", "rsa", "fei"])
def test_runfasteig_c1_deck(tmp_path, fmt):
    deck = decks.write_deck("c1", str(tmp_path), small=True, fmt=fmt)
    run_deck(deck, str(tmp_path))
    rows = raw_results(os.path.join(str(tmp_path), "c1_LOBPCG.log"))
    assert [r[0] for r in rows] == list(range(1, 11))
    ev = np.array([r[1] for r in rows])
    ex = np.sort(G.laplacian_2d_eigenvalues(20))[:10]
    assert np.all(np.abs(ev - ex) <= 1e-10 * np.abs(ex))
    assert max(r[2] for r in rows) <= 1e-8


def test_runfasteig_c2_generalized_deck(tmp_path):
    deck = decks.write_deck("c2", str(tmp_path), small=True)
    run_deck(deck, str(tmp_path))
    rows = raw_results(os.path.join(str(tmp_path), "c2_Chebyshev.log"))
    mats, opts = decks._config("c2", True)
    r = Problem(mats["a"], mass=mats["b"].diagonal()).solve(**opts)
    ev = np.array([x[1] for x in rows])
    assert len(ev) == opts["nev"]
    assert np.all(np.abs(ev - r.eigenvalues) <= 1e-10 * np.abs(r.eigenvalues))
    assert max(x[2] for x in rows) <= 1e-8

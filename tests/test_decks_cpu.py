"""Configuration decks: Matrix Market round trip and deck parsing (no GPU)."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from fasteig_amd import api, decks
from fasteig_amd import generators as G


def test_mtx_round_trip_is_exact(tmp_path):
    A = G.hamiltonian_27pt(6, seed=4)
    path = tmp_path / "h.mtx"
    decks.write_mtx(str(path), A)
    head = path.read_text().splitlines()[:2]
    assert head[0] == "%%MatrixMarket matrix coordinate real symmetric"
    assert head[1].split()[:2] == ["216", "216"]
    B = sp.csr_matrix(scipy.io.mmread(str(path)))
    assert (abs(B - A) > 0).nnz == 0  # 17 significant digits: bitwise values


@pytest.mark.parametrize("name", ["c1", "c2", "c3", "c4"])
def test_deck_parses(tmp_path, name):
    deck = decks.write_deck(name, str(tmp_path), small=True)
    mats, opts = decks._config(name, True)
    text = open(deck).read()
    for k, v in opts.items():
        assert f"\n{k} {v}\n" in text
    p = api.Problem(deck=deck)
    assert p.n == mats["a"].shape[0]
    p.close()
    if name == "c2":
        Bm = sp.csr_matrix(scipy.io.mmread(os.path.join(str(tmp_path), "c2_b.mtx")))
        assert np.array_equal(Bm.diagonal(), mats["b"].diagonal()) and Bm.nnz == mats["b"].shape[0]


def test_unknown_configuration():
    with pytest.raises(ValueError):
        decks.write_deck("c9", "/tmp/never")

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


@pytest.mark.parametrize("fmt", ["rsa", "fei"])
def test_compressed_column_decks_parse(tmp_path, fmt):
    """The reference's RSA text and FEI binary readers (src/parseinput.cpp:27-111):
    a deck naming a .rsa / .fei matrix loads with the right dimension; the
    FEI file's layout is checked byte for byte."""
    deck = decks.write_deck("c2", str(tmp_path), small=True, fmt=fmt)
    mats, _ = decks._config("c2", True)
    p = api.Problem(deck=deck)
    assert p.n == mats["a"].shape[0]
    p.close()
    if fmt == "fei":
        raw = open(os.path.join(str(tmp_path), "c2_a.fei"), "rb").read()
        n, _, nnz = np.frombuffer(raw[:12], dtype="<i4")
        C = sp.csc_matrix(mats["a"])
        assert n == C.shape[0] and nnz == C.nnz
        cp = np.frombuffer(raw[12:12 + 4 * (n + 1)], dtype="<i4")
        assert np.array_equal(cp, C.indptr)


def test_mismatched_matrix_formats_rejected(tmp_path):
    decks.write_deck("c2", str(tmp_path), small=True, fmt="fei")
    decks.write_deck("c2", str(tmp_path), small=True, fmt="mtx")
    deck = tmp_path / "mixed.input"
    deck.write_text("amatrix c2_a.mtx\nbmatrix c2_b.fei\nnev 4\nsolver LOBPCG\n")
    with pytest.raises(Exception):
        api.Problem(deck=str(deck))

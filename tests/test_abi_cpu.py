"""CPU-side checks of the native libraries: they load, export every entry point
their headers declare, and reject bad input before touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fasteig_amd import _fe, api
from fasteig_amd import generators as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fe|feh)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("fe_hip.h", lambda: _fe.lib()), ("fasteig_hip.h", lambda: api.lib())])
def test_every_declared_symbol_is_exported(header, lib):
    L = lib()
    names = declared(header)
    assert len(names) > 10 if header == "fe_hip.h" else len(names) > 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_kernel_library_argument_errors_without_device():
    L = _fe.lib()
    # null context -> -2 and a message, no device work
    assert L.fe_ctx_sync(None) == -2
    assert b"null" in L.fe_last_error()
    assert L.fe_spmm(None, None, 0, 4, None, 4, None, 4, 1.0, 0.0, None, 0, 0.0) == -2
    assert L.fe_gram(None, 0, 0, 10, 2, 2, None, 2, None, 2, None, None, 2) == -2
    assert L.fe_syev(None, 4, None, 4, None, None, 4) == -2


def test_problem_rejects_bad_csr_and_options_on_host():
    A = G.laplacian_2d(6)
    rp = np.ascontiguousarray(A.indptr, dtype=np.int32).copy()
    ci = np.ascontiguousarray(A.indices, dtype=np.int32).copy()
    v = np.ascontiguousarray(A.data)
    h = C.c_void_p()
    bad = ci.copy()
    bad[3] = 1000
    rc = api.lib().feh_problem_create(36, len(ci), rp.ctypes.data, bad.ctypes.data, v.ctypes.data, C.byref(h))
    assert rc == -2 and b"out of range" in api.lib().feh_last_error()
    p = api.Problem(A)
    for opts in ({"nev": "abc"}, {"precision": "fp16"}, {"gram": "bf16x3"}, {"nosuchkey": 1}, {"tol": -1}):
        with pytest.raises(ValueError):
            p.solve(**opts)
    p.close()


def test_deck_parser_reads_keys_and_matrix(tmp_path):
    A = G.laplacian_2d(5)
    T = A.tocoo()
    keep = T.row >= T.col
    mtx = tmp_path / "lap.mtx"
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% lower triangle\n")
        f.write(f"25 25 {keep.sum()}\n")
        for r, c, x in zip(T.row[keep], T.col[keep], T.data[keep]):
            f.write(f"{r + 1} {c + 1} {x}\n")
    deck = tmp_path / "lap.input"
    deck.write_text(f"# test deck\namatrix {mtx}\nnev 4\nwhich SR\nsolver Chebyshev\ndegree 14\ntol 1e-9\n"
                    "maxiter 77\nprecision mixed\n")
    p = api.Problem(deck=str(deck))
    assert p.n == 25
    # an invalid override is rejected after the deck's own options were applied
    with pytest.raises(ValueError):
        p.solve(nev=30)
    p.close()


def test_python_binding_declares_every_kernel_entry_point():
    """ctypes passes undeclared integers as 32-bit C ints: every entry point of
    fe_hip.h must have its argument types declared in the binding."""
    L = _fe.lib()
    missing = [n for n in declared("fe_hip.h") if getattr(L, n).argtypes is None]
    assert not missing, missing

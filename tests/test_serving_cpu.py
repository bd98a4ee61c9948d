"""Host-side logic of the serving loop, without a device: the admission
footprint and the grouped-launch planners (tiling, validation) of
include/fe_hip.h, called through the C ABI on host memory."""
import ctypes as C

import numpy as np
import pytest

from fasteig_amd import _fe
from fasteig_amd import serving


class Header(C.Structure):
    _fields_ = [("nitems", C.c_int32), ("kind", C.c_int32), ("ntiles", C.c_int64)]


class Spmm(C.Structure):
    _fields_ = [("n", C.c_int32), ("k", C.c_int32), ("prec", C.c_int32), ("lpr", C.c_int32),
                ("tile0", C.c_int64), ("rowptr", C.c_void_p), ("colind", C.c_void_p), ("vals", C.c_void_p),
                ("X", C.c_void_p), ("Y", C.c_void_p), ("Z", C.c_void_p), ("ldx", C.c_int64),
                ("ldy", C.c_int64), ("ldz", C.c_int64), ("alpha", C.c_double), ("beta", C.c_double),
                ("gamma", C.c_double)]


class Gram(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("p", C.c_int32), ("prec", C.c_int32),
                ("tile0", C.c_int64), ("X", C.c_void_p), ("Y", C.c_void_p), ("ldx", C.c_int64),
                ("ldy", C.c_int64), ("part", C.c_void_p), ("G", C.c_void_p), ("ldg", C.c_int64)]


class Gemm(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("p", C.c_int32), ("prec", C.c_int32),
                ("tile0", C.c_int64), ("X", C.c_void_p), ("ldx", C.c_int64), ("C", C.c_void_p),
                ("ldc", C.c_int64), ("Y", C.c_void_p), ("ldy", C.c_int64)]


def test_footprint_counts_matrix_blocks_and_partials():
    n, nnz, nev = 10000, 49600, 16
    m = 32
    csr = (n + 1) * 4 + nnz * 16
    blocks = 4 * ((n * m * 8 + 255) // 256 * 256)
    part = ((n + 1023) // 1024) * m * m * 8  # chunks of FE_GRP_CHUNK = 1024 rows
    fp = serving.footprint(n, nnz, nev)
    assert csr + blocks + part <= fp <= csr + blocks + part + 80000
    assert serving.footprint(2 * n, 2 * nnz, nev) > 1.9 * fp
    assert serving.footprint(0, 0, 4) == 0


def test_spmm_plan_tiles_and_lanes():
    L = _fe.lib()
    buf = (C.c_double * 64)()
    base = C.addressof(buf)
    items = (Spmm * 3)()
    for i, (n, k, prec) in enumerate([(1000, 16, _fe.F32), (77, 64, _fe.F64), (5000, 8, _fe.F32)]):
        items[i] = Spmm(n=n, k=k, prec=prec, rowptr=base, X=base, Y=base + 256, ldx=64, ldy=64)
    h = Header()
    assert L.fe_grp_spmm_plan(C.byref(h), items, 3) == 0
    # lanes: 16-byte runs per row rounded up to a power of two
    assert [it.lpr for it in items] == [4, 32, 2]
    rows_per_tile = [256 // it.lpr for it in items]
    tiles = [-(-it.n // r) for it, r in zip(items, rows_per_tile)]
    assert [it.tile0 for it in items] == [0, tiles[0], tiles[0] + tiles[1]]
    assert h.ntiles == sum(tiles) and h.nitems == 3


def test_spmm_plan_rejects_bad_items():
    L = _fe.lib()
    buf = (C.c_double * 64)()
    base = C.addressof(buf)
    h = Header()
    bad = [Spmm(n=10, k=65, prec=_fe.F32, rowptr=base, X=base, Y=base + 256, ldx=68, ldy=68),   # k > 64
           Spmm(n=10, k=8, prec=_fe.F32, rowptr=base, X=base + 4, Y=base + 256, ldx=8, ldy=8),  # misaligned
           Spmm(n=10, k=8, prec=_fe.F32, rowptr=base, X=base, Y=base + 256, ldx=6, ldy=8),      # ld not x4
           Spmm(n=10, k=8, prec=_fe.F32, rowptr=base, X=base, Y=base, ldx=8, ldy=8),            # Y aliases X
           Spmm(n=10, k=8, prec=7, rowptr=base, X=base, Y=base + 256, ldx=8, ldy=8)]            # precision
    for it in bad:
        arr = (Spmm * 1)(it)
        assert L.fe_grp_spmm_plan(C.byref(h), arr, 1) == -2


def test_gram_and_gemm_plans_chunk_by_rows():
    L = _fe.lib()
    buf = (C.c_double * 64)()
    base = C.addressof(buf)
    # chunks of FE_GRP_CHUNK = 1024 rows (include/fe_hip.h)
    g = (Gram * 2)(Gram(n=3000, m=16, p=16, prec=_fe.F32, X=base, Y=base, part=base, G=base, ldg=16),
                   Gram(n=256, m=64, p=8, prec=_fe.F64, X=base, Y=base, part=base, G=base, ldg=64))
    h = Header()
    assert L.fe_grp_gram_plan(C.byref(h), g, 2) == 0
    assert (g[0].tile0, g[1].tile0, h.ntiles) == (0, 3, 4)
    e = (Gemm * 2)(Gemm(n=130, m=8, p=8, prec=_fe.F32, X=base, C=base, ldc=8, Y=base + 512),
                   Gemm(n=64, m=8, p=8, prec=_fe.F32, X=base, C=base, ldc=8, Y=base + 512))
    assert L.fe_grp_gemm_plan(C.byref(h), e, 2) == 0
    assert (e[0].tile0, e[1].tile0, h.ntiles) == (0, 3, 4)
    e[1].ldc = 4  # ldc < m
    assert L.fe_grp_gemm_plan(C.byref(h), e, 2) == -2


def test_struct_sizes_match_header():
    """The ctypes mirrors above (and any FFI binding) must match the C layout."""
    assert C.sizeof(Spmm) == 4 * 4 + 8 + 6 * 8 + 3 * 8 + 3 * 8
    assert C.sizeof(Gram) == 4 * 4 + 8 + 2 * 8 + 2 * 8 + 3 * 8

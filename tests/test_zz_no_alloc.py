"""fe_gram_static, the Gram product that allocates nothing (its split-K
partials live in a fixed per-device array), against fe_gram through the C ABI.

Bar: bitwise equal to fe_gram when fe_gram's row chunks fit the fixed array
(same kernels, same chunks, same reduction order); within 1e-13 (FP64) /
2e-6 (FP32) / 1e-5 (compensated BF16x3) relative to the operand scale when
fewer, taller chunks are used (summation order only). Two calls in a row, the second on a smaller input, so
the array's leftover contents from the first call must not leak into the
second. The device's free memory does not change across a call (fe_gram's
first call of a size, the positive control, does take memory).
"""
import ctypes as C

import numpy as np
import pytest

from fasteig_amd import _fe

pytestmark = pytest.mark.gpu
L = _fe.lib


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def free_bytes():
    # the HIP runtime libfe_hip.so already loaded (same handle, no second copy)
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
    hip = C.CDLL(path)
    f, t = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def gram_pair(dev, mode, prec, X, Y, m, p, w=None):
    """(fe_gram, fe_gram_static) results for X^T diag(w) Y; Y is X when the same array."""
    dt = np.float64 if prec == _fe.F64 else np.float32
    n = X.shape[0]
    Xd = dev.put(X, dt)
    Yd = Xd if Y is X else dev.put(Y, dt)
    wd = dev.put(w) if w is not None else None
    out = []
    for fn in (L().fe_gram, L().fe_gram_static):
        Gd = dev.buffer((m, p), np.float64, order="F")
        _fe.check(fn(dev.ctx, mode, prec, n, m, p, Xd.ptr, X.shape[1], Yd.ptr, Y.shape[1],
                     wd.ptr if wd is not None else None, Gd.ptr, m))
        out.append(Gd.download())
    return out


@pytest.fixture()
def fresh():
    d = _fe.Device(0)
    yield d
    d.close()


def test_static_gram_two_calls_second_smaller(fresh):
    dev = fresh
    rng = np.random.default_rng(21)
    # 1) FP64 symmetric 256-column Gram on 300,000 rows: fe_gram's chunks need
    #    about 130 MB of partials, so the static variant runs taller chunks
    X = rng.uniform(-1, 1, (300000, 256))
    g, s = gram_pair(dev, 0, _fe.F64, X, X, 256, 256)
    assert rel(s, g) <= 1e-13
    assert np.array_equal(s, s.T)
    assert rel(s, X.T @ X) <= 1e-12
    # 2) smaller: FP32 thin product (64 x 64 tiles), chunks fit -> bitwise
    X2 = rng.uniform(-1, 1, (50000, 48))
    Y2 = rng.uniform(-1, 1, (50000, 20))
    w2 = rng.uniform(0.5, 1.5, 50000)
    g2, s2 = gram_pair(dev, 0, _fe.F32, X2, Y2, 48, 20, w2)
    assert np.array_equal(g2, s2)
    ref = X2.astype(np.float32).astype(np.float64).T @ (w2[:, None] * Y2.astype(np.float32).astype(np.float64))
    assert rel(s2, ref) <= 2e-6
    # 3) smaller still, FP64 non-symmetric, chunks fit -> bitwise
    X3 = rng.standard_normal((20000, 32))
    Y3 = rng.standard_normal((20000, 20))
    g3, s3 = gram_pair(dev, 0, _fe.F64, X3, Y3, 32, 20)
    assert np.array_equal(g3, s3)


def test_static_gram_bf16x3_matches(fresh):
    """The compensated BF16 Gram keeps 256 x 256 partial tiles: fe_gram's 512
    row chunks need 256 MB of them at m = 128, so the static variant runs fewer,
    taller chunks (summation order only: within 1e-5 of fe_gram, and both
    within the mode's 2^-14 of the FP64 product); two calls in a row, the second
    smaller."""
    rng = np.random.default_rng(22)
    for n, m in ((100000, 128), (3000, 64)):
        X = rng.uniform(-1, 1, (n, m)).astype(np.float32)
        g, s = gram_pair(fresh, _fe.GRAM_BF16X3, _fe.F32, X, X, m, m)
        ref = X.astype(np.float64).T @ X.astype(np.float64)
        assert np.array_equal(s, s.T)
        assert rel(s, g) <= 1e-5 and rel(s, ref) <= 2.0 ** -14 and rel(g, ref) <= 2.0 ** -14


def test_static_gram_allocates_nothing(fresh):
    dev = fresh
    rng = np.random.default_rng(23)
    X = dev.put(rng.standard_normal((200000, 192)))
    G = dev.buffer((192, 192), np.float64, order="F")
    # warm up the kernels on a small product (code-object loading)
    _fe.check(L().fe_gram_static(dev.ctx, 0, _fe.F64, 1000, 192, 192, X.ptr, 192, X.ptr, 192, None, G.ptr, 192))
    _fe.check(L().fe_gram(dev.ctx, 0, _fe.F64, 1000, 8, 8, X.ptr, 192, X.ptr, 192, None, G.ptr, 192))
    dev.sync()
    f0 = free_bytes()
    _fe.check(L().fe_gram_static(dev.ctx, 0, _fe.F64, 200000, 192, 192, X.ptr, 192, X.ptr, 192, None, G.ptr, 192))
    dev.sync()
    assert free_bytes() == f0
    _fe.check(L().fe_gram(dev.ctx, 0, _fe.F64, 200000, 192, 192, X.ptr, 192, X.ptr, 192, None, G.ptr, 192))
    dev.sync()
    assert free_bytes() < f0  # the growable scratch did allocate (the measure sees it)


def test_static_gram_rejects_oversized_chunk(fresh):
    dev = fresh
    rng = np.random.default_rng(24)
    n, m, p = 64, 2200, 2200   # 18 x 18 tiles of 128 x 128: one chunk needs 42 MB
    X = dev.put(rng.standard_normal((n, m)))
    Y = dev.put(rng.standard_normal((n, p)))
    G = dev.buffer((m, p), np.float64, order="F")
    rc = L().fe_gram_static(dev.ctx, 0, _fe.F64, n, m, p, X.ptr, m, Y.ptr, p, None, G.ptr, m)
    assert rc == -2
    assert b"exceed the fixed scratch" in L().fe_last_error()
    # the same shape as a symmetric product (171 tiles) fits
    _fe.check(L().fe_gram_static(dev.ctx, 0, _fe.F64, n, m, m, X.ptr, m, X.ptr, m, None, G.ptr, m))

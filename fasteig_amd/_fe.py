"""ctypes binding of libfe_hip.so (the C ABI in include/fe_hip.h).

Device blocks are row-major: a numpy array of shape (n, k) in C order maps
one-to-one to a device block with leading dimension k. Small matrices are
column-major FP64 (numpy Fortran order). The binding raises if the library
is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_PKG, "lib")
FE_SO = os.path.join(LIBDIR, "libfe_hip.so")

F64, F32 = 0, 1
GRAM_NATIVE, GRAM_BF16X3, GRAM_BF16X6 = 0, 1, 2
GRAM_SYMMETRIC = 16
FMT_CSR, FMT_SELL, FMT_BLOCKED, FMT_TILED = 0, 1, 2, 3

_lib = None


class FeError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(FE_SO):
            raise FeError(f"{FE_SO} is not built; run `python -m fasteig_amd.build` (no CPU fallback exists)")
        L = C.CDLL(FE_SO, mode=C.RTLD_GLOBAL)
        vp, i64, dbl, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
        L.fe_last_error.restype = C.c_char_p
        L.fe_version.restype = C.c_char_p
        L.fe_ctx_create.argtypes = [i32, C.POINTER(vp)]
        L.fe_ctx_destroy.argtypes = [vp]
        L.fe_ctx_stream.argtypes = [vp]
        L.fe_ctx_stream.restype = vp
        L.fe_ctx_sync.argtypes = [vp]
        L.fe_ctx_info.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_size_t), C.c_char_p, i32]
        L.fe_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
        L.fe_free.argtypes = [vp, vp]
        L.fe_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
        L.fe_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
        L.fe_memcpy_d2d.argtypes = [vp, vp, vp, C.c_size_t]
        L.fe_zero2d.argtypes = [vp, vp, i64, i64, i64, i32]
        L.fe_csr_create.argtypes = [vp, i32, i64, vp, vp, vp, C.POINTER(vp)]
        L.fe_csr_destroy.argtypes = [vp]
        L.fe_csr_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i64)]
        L.fe_csr_scale_sym.argtypes = [vp, vp, vp]
        L.fe_csr_build_sell.argtypes = [vp, vp, i32, i32]
        L.fe_csr_use.argtypes = [vp, i32]
        L.fe_csr_build_blocked.argtypes = [vp, vp, i32, C.POINTER(dbl)]
        L.fe_csr_block_schedule.argtypes = [vp, vp, i64, i64]
        L.fe_csr_build_tiled.argtypes = [vp, vp, C.POINTER(dbl)]
        L.fe_csr_tiled_kind.argtypes = [vp]
        L.fe_fill_random.argtypes = [vp, i32, i64, i32, vp, i64, C.c_uint64]
        L.fe_scale_rows.argtypes = [vp, i32, i64, i32, vp, i64, vp]
        L.fe_last_error.argtypes = []
        L.fe_version.argtypes = []
        L.fe_csr_arrays.argtypes = [vp, vp, vp, vp, vp]
        L.fe_memcpy_h2d_async.argtypes = [vp, vp, vp, C.c_size_t]
        L.fe_memcpy_d2h_async.argtypes = [vp, vp, vp, C.c_size_t]
        L.fe_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
        L.fe_host_free.argtypes = [vp]
        L.fe_graph_begin.argtypes = [vp]
        L.fe_graph_end.argtypes = [vp, C.POINTER(vp)]
        L.fe_graph_launch.argtypes = [vp, vp]
        L.fe_graph_destroy.argtypes = [vp]
        for nm in ("spmm", "gram", "resid", "gemm", "small"):
            getattr(L, f"fe_grp_{nm}_plan").argtypes = [vp, vp, i32]
        L.fe_grp_spmm_run.argtypes = [vp, vp]
        L.fe_grp_gemm_run.argtypes = [vp, vp]
        for nm in ("gram", "resid", "cholinv", "eig"):
            getattr(L, f"fe_grp_{nm}_run").argtypes = [vp, vp, i32]
        L.fe_spmm.argtypes = [vp, vp, i32, i32, vp, i64, vp, i64, dbl, dbl, vp, i64, dbl]
        L.fe_gram.argtypes = [vp, i32, i32, i64, i32, i32, vp, i64, vp, i64, vp, vp, i64]
        L.fe_gram_static.argtypes = [vp, i32, i32, i64, i32, i32, vp, i64, vp, i64, vp, vp, i64]
        L.fe_gemm_nn.argtypes = [vp, i32, i64, i32, i32, vp, i64, vp, i64, vp, i64, dbl, dbl]
        L.fe_trmm_upper.argtypes = [vp, i32, i64, i32, vp, i64, vp, i64, vp, i64]
        L.fe_gemm_nn_mode.argtypes = [vp, i32, i32, i64, i32, i32, vp, i64, vp, i64, vp, i64, dbl, dbl]
        L.fe_trmm_upper_mode.argtypes = [vp, i32, i32, i64, i32, vp, i64, vp, i64, vp, i64]
        L.fe_gemm_nn_resid.argtypes = [vp, i32, i32, i64, i32, i32, vp, i64, vp, i64, vp, i64, vp, i64, vp, vp,
                                       i64, vp, vp]
        L.fe_identity_deviation.argtypes = [vp, i32, vp, i64, C.POINTER(dbl)]
        L.fe_potrf.argtypes = [vp, i32, vp, i64]
        L.fe_chol_rinv.argtypes = [vp, i32, vp, i64, vp, i64]
        L.fe_syev.argtypes = [vp, i32, vp, i64, vp, vp, i64]
        L.fe_axpby_cols.argtypes = [vp, i32, i64, i32, vp, i64, vp, i64, dbl, dbl, vp]
        L.fe_colnorms.argtypes = [vp, i32, i64, i32, vp, i64, vp]
        L.fe_ritz_residuals.argtypes = [vp, i32, i64, i32, vp, i64, vp, i64, vp, vp, i64, vp, vp]
        L.fe_copy2d.argtypes = [vp, i64, i32, vp, i32, i64, vp, i32, i64]
        L.fe_block_gather_cols.argtypes = [vp, i32, i64, i32, vp, i64, vp, vp, i64]
        L.fe_upload_colmajor.argtypes = [vp, i64, i32, vp, vp, i32, i64]
        L.fe_download_colmajor.argtypes = [vp, i64, i32, vp, i32, i64, vp]
        L.fe_sym_small.argtypes = [vp, i32, vp, i64]
        L.fe_gather_cols.argtypes = [vp, i32, i32, vp, i64, vp, vp, i64]
        L.fe_zero_rows_small.argtypes = [vp, i32, i32, vp, i64]
        L.fe_event_record.argtypes = [vp, i32]
        L.fe_event_elapsed_ms.argtypes = [vp, i32, i32, C.POINTER(C.c_float)]
        L.fe_stream_probe.argtypes = [vp, i32, i64, vp, vp, vp, vp, i32]
        L.fe_gemm_small.argtypes = [vp, i32, i32, i32, i32, i32, dbl, vp, i64, vp, i64, dbl, vp, i64]
        L.fe_gram_equilibrate.argtypes = [vp, i32, vp, i64, vp]
        L.fe_scale_rows_small.argtypes = [vp, i32, i32, vp, vp, i64]
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise FeError(lib().fe_last_error().decode())
    return rc


class Buffer:
    """Device allocation holding a numpy-shaped array (row-major for blocks,
    column-major for small matrices when order='F')."""

    def __init__(self, dev: "Device", shape, dtype=np.float64, order="C"):
        self.dev = dev
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.order = order
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().fe_alloc(dev.ctx, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    @property
    def prec(self):
        return F64 if self.dtype == np.float64 else F32

    def upload(self, a):
        a = np.asarray(a, dtype=self.dtype)
        a = np.ascontiguousarray(a) if self.order == "C" else np.asfortranarray(a)
        assert a.shape == self.shape, (a.shape, self.shape)
        check(lib().fe_memcpy_h2d(self.dev.ctx, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype, order=self.order)
        check(lib().fe_memcpy_d2h(self.dev.ctx, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.dev.ctx:
            lib().fe_free(self.dev.ctx, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Csr:
    def __init__(self, dev: "Device", A):
        self.dev = dev
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        self.n = len(rp) - 1
        self.nnz = len(ci)
        h = C.c_void_p()
        check(lib().fe_csr_create(dev.ctx, self.n, self.nnz, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                  C.byref(h)))
        self.h = h.value
        dev._matrices.add(self)

    def build_sell(self, C_=32, sigma=256):
        check(lib().fe_csr_build_sell(self.dev.ctx, self.h, C_, sigma))

    def build_blocked(self, rb=4):
        """Build the row-block gather copy; returns (built, reuse)."""
        r = C.c_double(0.0)
        rc = check(lib().fe_csr_build_blocked(self.dev.ctx, self.h, rb, C.byref(r)))
        return rc == 0, r.value

    def block_schedule(self, plane=-1, strip=0):
        """Row-block processing order for 3D grid stencils (plane -1: detect,
        0: remove); returns True when an order is in place."""
        return check(lib().fe_csr_block_schedule(self.dev.ctx, self.h, plane, strip)) == 0 and plane != 0

    def build_tiled(self):
        """Build the LDS-tiled copy; returns (built, X rows loaded per row)."""
        r = C.c_double(0.0)
        rc = check(lib().fe_csr_build_tiled(self.dev.ctx, self.h, C.byref(r)))
        return rc == 0, r.value

    def tiled_kind(self):
        """Form of the tiled copy: 0 none, 16 / 32 entry-list tiles of that
        many rows, 64 the stencil form (8 x 8 patches)."""
        return lib().fe_csr_tiled_kind(self.h)

    def use(self, fmt):
        check(lib().fe_csr_use(self.h, fmt))

    def destroy(self):
        """Release the device copy (must happen before its context dies)."""
        if self.h:
            lib().fe_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Device:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().fe_ctx_create(device, C.byref(h)))
        self.ctx = h.value
        # matrices are synchronised on their context's stream when destroyed,
        # so close() releases the live ones first
        self._matrices = weakref.WeakSet()

    def sync(self):
        check(lib().fe_ctx_sync(self.ctx))

    def info(self):
        cus = C.c_int()
        hbm = C.c_size_t()
        name = C.create_string_buffer(128)
        check(lib().fe_ctx_info(self.ctx, C.byref(cus), C.byref(hbm), name, 128))
        return {"cus": cus.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}

    def buffer(self, shape, dtype=np.float64, order="C"):
        return Buffer(self, shape, dtype, order)

    def put(self, a, dtype=None, order="C"):
        a = np.asarray(a)
        b = Buffer(self, a.shape, dtype or a.dtype, order)
        return b.upload(a)

    def csr(self, A):
        return Csr(self, A)

    def close(self):
        if self.ctx:
            for m in list(self._matrices):
                m.destroy()
            lib().fe_ctx_destroy(self.ctx)
            self.ctx = None

"""ctypes loader for the oracle (CPU FP64 restatement of the pspaQR ParTree path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_tree_create.restype = C.c_void_p
        L.orc_tree_create.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_tree_create_parts.restype = C.c_void_p
        L.orc_tree_create_parts.argtypes = [C.c_int, _ip, _ip, _dp, _ip, C.c_int, C.c_double, C.c_int]
        for f in ("orc_partition_assemble", "orc_factorize", "orc_num_clusters", "orc_trace_len"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        L.orc_tree_destroy.argtypes = [C.c_void_p]
        L.orc_factor_seconds.argtypes = [C.c_void_p]
        L.orc_factor_seconds.restype = C.c_double
        L.orc_get_cperm.argtypes = [C.c_void_p, _ip]
        L.orc_get_clusters.argtypes = [C.c_void_p, _ip]
        L.orc_get_trace.argtypes = [C.c_void_p, _ip]
        L.orc_get_diag_R.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int]
        L.orc_get_diag_R.restype = C.c_int
        L.orc_solve.argtypes = [C.c_void_p, _dp]
        L.orc_spmv.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_gmres.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_double, _dp, C.c_int,
                                C.POINTER(C.c_int)]
        L.orc_gmres.restype = C.c_int
        L.orc_tree_create_rect.restype = C.c_void_p
        L.orc_tree_create_rect.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _ip, C.c_int, C.c_double, C.c_int]
        L.orc_nrows.argtypes = [C.c_void_p]
        L.orc_nrows.restype = C.c_int
        L.orc_get_rperm.argtypes = [C.c_void_p, _ip]
        L.orc_get_rows.argtypes = [C.c_void_p, _ip]
        L.orc_get_r2c.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int]
        L.orc_get_r2c.restype = C.c_int
        L.orc_solve_ls.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_solve_nrml.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_spmv_t.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_cgls.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.orc_cgls.restype = C.c_int
        # dense kernels
        L.orc_geqrt.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_tpqrt.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_tpmqrt_lt.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_larft.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_larfb_lt.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_larfb_rn.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_geqp3.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp]
        L.orc_geqp3_tie.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, C.c_double]
        L.orc_geqp3_near_ties.restype = C.c_long
        L.orc_geqp3_near_ties.argtypes = [C.c_int]
        L.orc_gemqrt_lt.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_choose_rank.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_choose_rank.restype = C.c_int
        L.orc_trsm_right_upper.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_trsv_upper.argtypes = [C.c_int, _dp, _dp]
        L.orc_ormqr_lt.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_ormqr_ln.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_set_fault.argtypes = [C.c_int]
        # rank path
        L.orc_gauss.argtypes = [C.c_int, C.c_int, C.c_uint64, _dp]
        L.orc_colnorm.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_colnorm.restype = C.c_double
        L.orc_sketch.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp]
        L.orc_rank_features.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                        np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")]
        L.orc_tail_select.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_double, C.c_double]
        L.orc_tail_select.restype = C.c_int
        L.orc_sketch_rank.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_int]
        L.orc_sketch_rank.restype = C.c_int
        _lib = L
    return _lib


def F(a):
    """Column-major float64 copy, flattened (the oracle's storage)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64)).ravel(order="F").copy()


def unF(v, r, c):
    return np.asarray(v).reshape((r, c), order="F")


class OracleTree:
    """The reference's ParTree driver sequence on the CPU: partition, assemble, factorize,
    solve, spmv and dist_gmres (pspaQR.cpp:183-268)."""

    def __init__(self, A, n_side: int, dim: int, nlevels: int, tol: float, skip: int = 0, parts=None):
        """parts: leaf parts (one per column) ordering the tree as the reference's METIS path
        does (MetisPartition with these parts, src/partition.cpp:372-424); n_side and dim are
        then ignored."""
        A = A.tocsc()
        A.sort_indices()
        self.m, self.n = A.shape
        self._colptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._rowind = np.ascontiguousarray(A.indices, dtype=np.int32)
        self._val = np.ascontiguousarray(A.data, dtype=np.float64)
        L = lib()
        if self.m != self.n:
            if parts is None:
                raise ValueError("a rectangular system is ordered by given leaf parts")
            self._parts = np.ascontiguousarray(parts, dtype=np.int32)
            self._h = L.orc_tree_create_rect(self.m, self.n, self._colptr, self._rowind, self._val, self._parts, nlevels,
                                             tol, skip)
        elif parts is None:
            self._h = L.orc_tree_create(self.n, self._colptr, self._rowind, self._val, n_side, dim, nlevels, tol, skip)
        else:
            self._parts = np.ascontiguousarray(parts, dtype=np.int32)
            self._h = L.orc_tree_create_parts(self.n, self._colptr, self._rowind, self._val, self._parts, nlevels, tol,
                                              skip)
        self.nlevels = nlevels
        if L.orc_partition_assemble(self._h) != 0:
            raise RuntimeError("oracle partition/assemble failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_tree_destroy(h)
            self._h = None

    def factorize(self):
        if lib().orc_factorize(self._h) != 0:
            raise RuntimeError("oracle factorize failed")
        return lib().orc_factor_seconds(self._h)

    def cperm(self):
        p = np.zeros(self.n, dtype=np.int32)
        lib().orc_get_cperm(self._h, p)
        return p

    def clusters(self):
        nc = lib().orc_num_clusters(self._h)
        out = np.zeros(13 * nc, dtype=np.int32)
        lib().orc_get_clusters(self._h, out)
        return out.reshape(nc, 13)

    def trace(self):
        n = lib().orc_trace_len(self._h)
        out = np.zeros(5 * max(n, 1), dtype=np.int32)
        lib().orc_get_trace(self._h, out)
        return out[: 5 * n].reshape(n, 5)

    def diag_R(self, order: int, maxn: int = 1 << 22):
        buf = np.zeros(maxn)
        k = lib().orc_get_diag_R(self._h, order, buf, maxn)
        if k < 0:
            return None
        return unF(buf[: k * k], k, k)

    def solve(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc_solve(self._h, x)
        return x

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.m)
        lib().orc_spmv(self._h, x, y)
        return y

    def spmv_t(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        y = np.zeros(self.n)
        lib().orc_spmv_t(self._h, r, y)
        return y

    def rperm(self):
        p = np.zeros(self.m, dtype=np.int32)
        lib().orc_get_rperm(self._h, p)
        return p

    def rows(self):
        """Per cluster: rstart, rows at creation, extra rows received, row_rank, extra rows moved."""
        nc = lib().orc_num_clusters(self._h)
        out = np.zeros(5 * nc, dtype=np.int32)
        lib().orc_get_rows(self._h, out)
        return out.reshape(nc, 5)

    def r2c(self, order: int):
        buf = np.zeros(1 << 20, dtype=np.int32)
        k = lib().orc_get_r2c(self._h, order, buf, buf.size)
        return buf[:k].copy()

    def solve_ls(self, b):
        """ParTree::solve(b, x): b (m, permuted rows) -> x (n, permuted columns)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        lib().orc_solve_ls(self._h, b, x)
        return x

    def solve_nrml(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        x = np.zeros(self.n)
        lib().orc_solve_nrml(self._h, y, x)
        return x

    def cgls(self, b, iters: int = 300, tol: float = 1e-12):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        rel = C.c_double(0.0)
        it = lib().orc_cgls(self._h, b, x, iters, tol, C.byref(rel))
        if it < 0:
            raise RuntimeError("oracle cgls failed")
        return x, it, rel.value

    def gmres(self, b, iters: int = 300, tol: float = 1e-12, restart: int = 0, true_tol: float = 1e-12):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        hist = np.zeros(max(iters, 1) + 16)
        hl = C.c_int(0)
        it = lib().orc_gmres(self._h, b, x, iters, tol, restart, true_tol, hist, hist.size, C.byref(hl))
        return x, it, hist[: hl.value]


class planted_fault:
    """Context manager: run the oracle with one of its planted mistakes (spaqr_oracle.cpp,
    g_fault) so that a test can show a parity check catches it."""

    def __init__(self, code: int):
        self.code = code

    def __enter__(self):
        lib().orc_set_fault(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_set_fault(0)
        return False


def threads() -> int:
    """Threads the oracle's factorization uses (1: it runs sequentially)."""
    L = lib()
    return int(L.orc_threads()) if hasattr(L, "orc_threads") else 1

"""Matrix Market files through the library's reader and writer (spaqr_mm_* in
include/spaqr_gpu.h; the reference's include/mmio.hpp:174-377).  Host only: no GPU is touched.

    A = read_sparse("matrix.mtx")     # scipy.sparse.csc_matrix, duplicates summed
    X = read_dense("coords.mtx")      # numpy array (d x n coordinates for set_Xcoo)
    b = read_vector("rhs.mtx")        # an n x 1 array file as a vector
    write_sparse("out.mtx", A, symmetric=False)
    write_dense("out.mtx", X)
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from ._lib import lib


def _err(rc, what, path):
    msg = lib().spaqr_last_error(None)
    raise OSError("%s %s failed (%d): %s" % (what, path, rc, msg.decode() if msg else ""))


def read_sparse(path: str) -> sp.csc_matrix:
    p = str(path).encode()
    nr, nc, nnz = C.c_int(0), C.c_int(0), C.c_longlong(0)
    rc = lib().spaqr_mm_read_sparse(p, C.byref(nr), C.byref(nc), C.byref(nnz), None, None, None)
    if rc:
        _err(rc, "reading", path)
    colptr = np.zeros(nc.value + 1, dtype=np.int32)
    rowind = np.zeros(max(nnz.value, 1), dtype=np.int32)
    val = np.zeros(max(nnz.value, 1), dtype=np.float64)
    rc = lib().spaqr_mm_read_sparse(p, C.byref(nr), C.byref(nc), C.byref(nnz), colptr.ctypes.data,
                                    rowind.ctypes.data, val.ctypes.data)
    if rc:
        _err(rc, "reading", path)
    return sp.csc_matrix((val[: nnz.value], rowind[: nnz.value], colptr), shape=(nr.value, nc.value))


def read_dense(path: str) -> np.ndarray:
    p = str(path).encode()
    nr, nc = C.c_int(0), C.c_int(0)
    rc = lib().spaqr_mm_read_dense(p, C.byref(nr), C.byref(nc), None)
    if rc:
        _err(rc, "reading", path)
    out = np.zeros(max(nr.value * nc.value, 1), dtype=np.float64)
    rc = lib().spaqr_mm_read_dense(p, C.byref(nr), C.byref(nc), out.ctypes.data)
    if rc:
        _err(rc, "reading", path)
    return out[: nr.value * nc.value].reshape(nc.value, nr.value).T.copy()


def read_vector(path: str) -> np.ndarray:
    X = read_dense(path)
    if X.shape[1] != 1:
        raise ValueError("%s holds a %d x %d array, not a vector" % (path, X.shape[0], X.shape[1]))
    return X[:, 0].copy()


def write_sparse(path: str, A, symmetric: bool = False):
    A = sp.csc_matrix(A)
    A.sort_indices()
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    rowind = np.ascontiguousarray(A.indices, dtype=np.int32)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    rc = lib().spaqr_mm_write_sparse(str(path).encode(), A.shape[0], A.shape[1], colptr.ctypes.data,
                                     rowind.ctypes.data, val.ctypes.data, int(bool(symmetric)))
    if rc:
        _err(rc, "writing", path)


def write_dense(path: str, X):
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    data = np.ascontiguousarray(X.T).ravel()  # column-major
    rc = lib().spaqr_mm_write_dense(str(path).encode(), X.shape[0], X.shape[1], data.ctypes.data)
    if rc:
        _err(rc, "writing", path)

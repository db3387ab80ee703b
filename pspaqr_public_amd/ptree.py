"""Python mirror of the reference's ParTree driver API over the HIP library.

Method names and argument meaning follow include/ptree.h:123-159 and include/tree.h:137-143
(set_tol, set_skip, set_Xcoo, partition, assemble, factorize, solve, spmv) plus
dist_gmres (src/is.cpp:447-527).  Vectors are numpy arrays in the permuted ordering the
reference's solve/spmv/dist_gmres work in; they are copied to HBM for the call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceBuffer, GmresResult, check, lib, register


def linspace_nd(n: int, d: int) -> np.ndarray:
    """Tensor-grid coordinates (d x n^d, column-major), id = sum_i x_i n^(d-1-i)
    (src/util.cpp:901-930)."""
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64)] * d, indexing="ij")
    X = np.stack([g.ravel() for g in grids], axis=0)
    return np.asfortranarray(X)


class ParTree:
    def __init__(self, nlevels: int, skip: int = 0):
        self._h = lib().spaqr_tree_create(nlevels, skip)
        if not self._h:
            raise RuntimeError("spaqr_tree_create failed: %s" % lib().spaqr_last_error(None).decode())
        self.nlevels = nlevels
        self._n = 0
        self._m = 0
        self._keep = []
        register(self)

    def close(self):
        """Releases the tree's device memory and stream (also done at interpreter exit)."""
        for b in ("_bx", "_by"):
            buf = getattr(self, b, None)
            if buf is not None:
                buf.free()
        h = getattr(self, "_h", None)
        if h:
            lib().spaqr_tree_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setters (include/tree.h:137-143, include/ptree.h:123-125)
    def set_tol(self, tol: float):
        check(lib().spaqr_tree_set_tol(self._h, float(tol)), self._h, "set_tol")

    def set_skip(self, s: int):
        check(lib().spaqr_tree_set_skip(self._h, int(s)), self._h, "set_skip")

    def set_scale(self, s: int):
        check(lib().spaqr_tree_set_scale(self._h, int(s)), self._h, "set_scale")

    def set_square(self, s: int):
        check(lib().spaqr_tree_set_square(self._h, int(s)), self._h, "set_square")

    def set_order(self, o: float):
        check(lib().spaqr_tree_set_order(self._h, float(o)), self._h, "set_order")

    def set_verbose(self, v: int):
        check(lib().spaqr_tree_set_verbose(self._h, int(v)), self._h, "set_verbose")

    def set_precision(self, bits: int):
        check(lib().spaqr_tree_set_precision(self._h, int(bits)), self._h, "set_precision")

    def set_device_symbolic(self, on: bool = True):
        """Partition and assemble on the device (square systems with coordinates)."""
        check(lib().spaqr_tree_set_device_symbolic(self._h, 1 if on else 0), self._h, "set_device_symbolic")

    def symbolic_ms(self):
        """(partition ms, assemble ms) of the last partition and assemble."""
        out = np.zeros(2)
        check(lib().spaqr_tree_symbolic_ms(self._h, out.ctypes.data), self._h, "symbolic_ms")
        return float(out[0]), float(out[1])

    def set_Xcoo(self, X: np.ndarray):
        X = np.asfortranarray(X, dtype=np.float64)
        self._X = X
        check(lib().spaqr_tree_set_Xcoo(self._h, X.ctypes.data, X.shape[0], X.shape[1]), self._h, "set_Xcoo")

    def set_parts(self, parts):
        """Leaf parts in place of coordinates (spaqr_tree_set_parts); None clears them."""
        if parts is None:
            check(lib().spaqr_tree_set_parts(self._h, None, 0), self._h, "set_parts")
            return
        p = np.ascontiguousarray(parts, dtype=np.int32)
        self._parts = p
        check(lib().spaqr_tree_set_parts(self._h, p.ctypes.data, int(p.size)), self._h, "set_parts")

    def parts(self) -> np.ndarray:
        """Leaf part of every column (natural order) used by the last partition."""
        n = lib().spaqr_tree_get_parts(self._h, None, 0)
        out = np.zeros(max(n, 0), dtype=np.int32)
        if n > 0:
            lib().spaqr_tree_get_parts(self._h, out.ctypes.data, n)
        return out

    # ---- partition / assemble / factorize
    def partition(self, A):
        A = A.tocsc()
        A.sort_indices()
        self._m, self._n = A.shape
        cp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ri = np.ascontiguousarray(A.indices, dtype=np.int32)
        va = np.ascontiguousarray(A.data, dtype=np.float64)
        self._keep = [cp, ri, va]
        check(lib().spaqr_tree_partition(self._h, A.shape[0], A.shape[1], cp.ctypes.data, ri.ctypes.data,
                                         va.ctypes.data), self._h, "partition")

    TUNABLES = ("big_front", "tpmqrt_tile", "larfb_cols", "trsm_rows", "spars_rows", "spars_cols", "rank_cols",
                "sketch_unpivoted", "tpqrt_coop", "small_tasks", "sketch_bf16", "sketch_tile", "sketch_grow",
                "sketch_pad", "panel_cols")

    def set_tuning(self, params: dict):
        """Planner tunables by name (spaqr_tree_set_tuning); 0 restores the built-in value."""
        for k, v in params.items():
            check(lib().spaqr_tree_set_tuning(self._h, self.TUNABLES.index(k), int(v)), self._h, "set_tuning " + k)

    def set_krylov_mode(self, mode: int):
        """GMRES orthogonalisation: 0 MGS with host Givens (the reference's), 1 fused CGS2 with
        device Givens (default)."""
        check(lib().spaqr_tree_set_krylov_mode(self._h, int(mode)), self._h, "set_krylov_mode")

    def set_pivot_check(self, rel: float):
        check(lib().spaqr_tree_set_pivot_check(self._h, float(rel)), self._h, "set_pivot_check")

    def info(self) -> np.ndarray:
        """Status flags per cluster: 1 non-finite R entry, 2 zero pivot, 4 predictor miss."""
        n = lib().spaqr_tree_get_info(self._h, None, 0)
        out = np.zeros(max(n, 0), dtype=np.int32)
        if n > 0:
            lib().spaqr_tree_get_info(self._h, out.ctypes.data, n)
        return out

    def set_graphs(self, on: bool = True):
        """hipGraph mode: record the factorization's launch segments once, replay them after."""
        check(lib().spaqr_tree_set_graphs(self._h, int(bool(on))), self._h, "set_graphs")

    def graph_info(self) -> dict:
        out = np.zeros(3, dtype=np.int64)
        check(lib().spaqr_tree_graph_info(self._h, out.ctypes.data), self._h, "graph_info")
        return dict(on=bool(out[0]), segments=int(out[1]), replays=int(out[2]))

    def tuning(self) -> dict:
        return {k: lib().spaqr_tree_get_tuning(self._h, i) for i, k in enumerate(self.TUNABLES)}

    def partition_batch(self, A, sizes, levels):
        """Independent systems on the diagonal blocks of A (spaqr_tree_partition_batch)."""
        A = A.tocsc()
        A.sort_indices()
        self._n = A.shape[1]
        self._m = A.shape[0]
        cp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ri = np.ascontiguousarray(A.indices, dtype=np.int32)
        va = np.ascontiguousarray(A.data, dtype=np.float64)
        sz = np.ascontiguousarray(sizes, dtype=np.int32)
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        self._keep = [cp, ri, va, sz, lv]
        check(lib().spaqr_tree_partition_batch(self._h, len(sz), sz.ctypes.data, lv.ctypes.data, A.shape[1],
                                               cp.ctypes.data, ri.ctypes.data, va.ctypes.data), self._h,
              "partition_batch")

    def systems(self) -> np.ndarray:
        """(offset, size) per system."""
        ns = lib().spaqr_tree_num_systems(self._h)
        out = np.zeros(2 * ns, dtype=np.int32)
        check(lib().spaqr_tree_get_systems(self._h, out.ctypes.data), self._h, "get_systems")
        return out.reshape(ns, 2)

    def cluster_systems(self) -> np.ndarray:
        out = np.zeros(lib().spaqr_tree_num_clusters(self._h), dtype=np.int32)
        check(lib().spaqr_tree_get_cluster_systems(self._h, out.ctypes.data), self._h, "get_cluster_systems")
        return out

    def gmres_batch(self, b: np.ndarray, iters: int = 300, tol: float = 1e-12, restart: int = 0,
                    true_tol: float = 1e-12, flexible: bool = False):
        """GMRES on every system of a batch; returns (x, [GmresResult per system])."""
        ns = lib().spaqr_tree_num_systems(self._h)
        self._bx.upload(np.asarray(b, dtype=np.float64))
        res = (GmresResult * ns)()
        check(lib().spaqr_gmres_batch(self._h, self._bx.ptr, self._by.ptr, iters, tol, restart, true_tol,
                                      int(bool(flexible)), res), self._h, "gmres_batch")
        return self._by.download(np.float64, self._n), list(res)

    def gmres_batch_device(self, b_ptr, x_ptr, iters=300, tol=1e-12, restart=0, true_tol=1e-12, flexible=False):
        ns = lib().spaqr_tree_num_systems(self._h)
        res = (GmresResult * ns)()
        check(lib().spaqr_gmres_batch(self._h, b_ptr, x_ptr, iters, tol, restart, true_tol, int(bool(flexible)), res),
              self._h, "gmres_batch")
        return list(res)

    def assemble(self, A=None):
        check(lib().spaqr_tree_assemble(self._h), self._h, "assemble")
        self._bx = DeviceBuffer(8 * max(self._m, self._n))
        self._by = DeviceBuffer(8 * max(self._m, self._n))

    def factorize(self) -> int:
        check(lib().spaqr_tree_factorize(self._h), self._h, "factorize")
        return 0

    # ---- solve / spmv / gmres (numpy in, numpy out)
    def solve(self, x: np.ndarray) -> np.ndarray:
        self._bx.upload(np.asarray(x, dtype=np.float64))
        check(lib().spaqr_tree_solve(self._h, self._bx.ptr), self._h, "solve")
        check(lib().spaqr_tree_sync(self._h), self._h, "sync")
        return self._bx.download(np.float64, self._n)

    def spmv(self, x: np.ndarray) -> np.ndarray:
        self._bx.upload(np.asarray(x, dtype=np.float64))
        check(lib().spaqr_tree_spmv(self._h, self._bx.ptr, self._by.ptr), self._h, "spmv")
        check(lib().spaqr_tree_sync(self._h), self._h, "sync")
        return self._by.download(np.float64, self._m)

    # ---- least squares (rectangular A, m > n)
    def spmv_t(self, r: np.ndarray) -> np.ndarray:
        """A^T r (permuted orderings): r has m entries, the result n."""
        self._bx.upload(np.asarray(r, dtype=np.float64))
        check(lib().spaqr_tree_spmv_t(self._h, self._bx.ptr, self._by.ptr), self._h, "spmv_t")
        check(lib().spaqr_tree_sync(self._h), self._h, "sync")
        return self._by.download(np.float64, self._n)

    def solve_ls(self, b: np.ndarray) -> np.ndarray:
        """ParTree::solve(b, x): b (m, permuted rows) -> x (n, permuted columns)."""
        self._bx.upload(np.asarray(b, dtype=np.float64))
        check(lib().spaqr_tree_solve_ls(self._h, self._bx.ptr, self._by.ptr), self._h, "solve_ls")
        check(lib().spaqr_tree_sync(self._h), self._h, "sync")
        return self._by.download(np.float64, self._n)

    def solve_nrml(self, y: np.ndarray) -> np.ndarray:
        """x = M^{-1} M^{-T} y (n entries, permuted columns): cgls's preconditioner."""
        self._bx.upload(np.asarray(y, dtype=np.float64))
        check(lib().spaqr_tree_solve_nrml(self._h, self._bx.ptr, self._by.ptr), self._h, "solve_nrml")
        check(lib().spaqr_tree_sync(self._h), self._h, "sync")
        return self._by.download(np.float64, self._n)

    def cgls(self, b: np.ndarray, iters: int = 300, tol: float = 1e-12):
        """cgls(A, b, x, M, iters, tol) with x0 = 0 (src/is.cpp:274-345); returns (x, iterations,
        ||A^T r|| / ||A^T b||).  b has m entries in the permuted row ordering, x n entries."""
        self._bx.upload(np.asarray(b, dtype=np.float64))
        it, rel = C.c_int(0), C.c_double(0.0)
        check(lib().spaqr_cgls(self._h, self._bx.ptr, self._by.ptr, iters, tol, C.byref(it), C.byref(rel)), self._h,
              "cgls")
        return self._by.download(np.float64, self._n), it.value, rel.value

    def cgls_device(self, b_ptr, x_ptr, iters=300, tol=1e-12):
        it, rel = C.c_int(0), C.c_double(0.0)
        check(lib().spaqr_cgls(self._h, b_ptr, x_ptr, iters, tol, C.byref(it), C.byref(rel)), self._h, "cgls")
        return it.value, rel.value

    @property
    def m(self) -> int:
        return self._m

    def rperm(self) -> np.ndarray:
        p = np.zeros(self._m, dtype=np.int32)
        check(lib().spaqr_tree_get_rperm(self._h, p.ctypes.data), self._h, "get_rperm")
        return p

    def rows(self) -> np.ndarray:
        """Per cluster: first row, rows at creation, extra rows received, row_rank, extra rows moved."""
        nc = lib().spaqr_tree_num_clusters(self._h)
        out = np.zeros(5 * nc, dtype=np.int32)
        check(lib().spaqr_tree_get_rows(self._h, out.ctypes.data), self._h, "get_rows")
        return out.reshape(nc, 5)

    def gmres(self, b: np.ndarray, iters: int = 300, tol: float = 1e-12, restart: int = 0, true_tol: float = 1e-12,
              flexible: bool = False):
        """dist_gmres(A, b, M, iters, tol, verb=true); returns (x, result, history).

        flexible=True runs the right-preconditioned flexible variant (no reference counterpart)."""
        self._bx.upload(np.asarray(b, dtype=np.float64))
        res = GmresResult()
        hist = np.zeros(iters + 16)
        hl = C.c_int(0)
        fn = lib().spaqr_fgmres if flexible else lib().spaqr_gmres
        check(fn(self._h, self._bx.ptr, self._by.ptr, iters, tol, restart, true_tol, C.byref(res),
                 hist.ctypes.data, hist.size, C.byref(hl)), self._h, "gmres")
        x = self._by.download(np.float64, self._n)
        return x, res, hist[: hl.value]

    def gmres_device(self, b_ptr, x_ptr, iters=300, tol=1e-12, restart=0, true_tol=1e-12, flexible=False):
        res = GmresResult()
        fn = lib().spaqr_fgmres if flexible else lib().spaqr_gmres
        check(fn(self._h, b_ptr, x_ptr, iters, tol, restart, true_tol, C.byref(res), None, 0, None), self._h, "gmres")
        return res

    # ---- rank path (no reference counterpart; see include/spaqr_gpu.h)
    def set_rank_mode(self, mode: int):
        check(lib().spaqr_tree_set_rank_mode(self._h, int(mode)), self._h, "set_rank_mode")

    def set_rank_model(self, weights: np.ndarray):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        check(lib().spaqr_tree_set_rank_model(self._h, w.ctypes.data, int(w.size)), self._h, "set_rank_model")

    def set_rank_log(self, on: bool = True):
        check(lib().spaqr_tree_set_rank_log(self._h, int(bool(on))), self._h, "set_rank_log")

    def rank_log(self):
        """(X [n x 32] float32, meta [n x 5] int32: rank, k, w, level, mode used)."""
        n = lib().spaqr_tree_rank_log(self._h, None, None, 0)
        nf = lib().spaqr_mlp_features()
        X = np.zeros((n, nf), dtype=np.float32)
        meta = np.zeros((n, 5), dtype=np.int32)
        if n:
            lib().spaqr_tree_rank_log(self._h, X.ctypes.data, meta.ctypes.data, n)
        return X, meta

    def rank_log_clusters(self) -> np.ndarray:
        """Cluster order (row of clusters()) of each rank-log entry."""
        n = lib().spaqr_tree_rank_log_clusters(self._h, None, 0)
        out = np.zeros(max(n, 0), dtype=np.int32)
        if n > 0:
            lib().spaqr_tree_rank_log_clusters(self._h, out.ctypes.data, n)
        return out

    def set_rank_retries(self, retries: int):
        check(lib().spaqr_tree_set_rank_retries(self._h, int(retries)), self._h, "set_rank_retries")

    def rank_log_clear(self):
        check(lib().spaqr_tree_rank_log_clear(self._h), self._h, "rank_log_clear")

    def rank_stats(self) -> dict:
        out = np.zeros(7)
        check(lib().spaqr_tree_rank_stats(self._h, out.ctypes.data), self._h, "rank_stats")
        return dict(interfaces=int(out[0]), sketched=int(out[1]), fallback=int(out[2]), direct=int(out[3]),
                    sketch_cols=float(out[4]), full_cols=float(out[5]), retried=int(out[6]))

    def rank_stats_reset(self):
        check(lib().spaqr_tree_rank_stats_reset(self._h), self._h, "rank_stats_reset")

    # ---- introspection
    @property
    def n(self) -> int:
        return self._n

    def cperm(self) -> np.ndarray:
        p = np.zeros(self._n, dtype=np.int32)
        check(lib().spaqr_tree_get_cperm(self._h, p.ctypes.data), self._h, "get_cperm")
        return p

    def clusters(self) -> np.ndarray:
        nc = lib().spaqr_tree_num_clusters(self._h)
        out = np.zeros(13 * nc, dtype=np.int32)
        check(lib().spaqr_tree_get_clusters(self._h, out.ctypes.data), self._h, "get_clusters")
        return out.reshape(nc, 13)

    def level_stats(self) -> np.ndarray:
        out = np.zeros(8 * self.nlevels)
        nl = lib().spaqr_tree_level_stats(self._h, out.ctypes.data, self.nlevels)
        return out[: 8 * nl].reshape(nl, 8)

    def diag_R(self, order: int, maxn: int = 1 << 22):
        buf = np.zeros(maxn)
        k = lib().spaqr_tree_diag_R(self._h, order, buf.ctypes.data, maxn)
        if k < 0:
            return None
        return buf[: k * k].reshape((k, k), order="F")

    def factor_ms(self) -> float:
        return lib().spaqr_tree_factor_ms(self._h)

    @property
    def handle(self):
        return self._h


def build_tree(problem, precision: int = 64, nlevels=None, tol=None, skip=None, device_symbolic: bool = False) -> ParTree:
    """The reference driver's set-up sequence (pspaQR.cpp:183-198) for a Problem."""
    nl = problem.nlevels if nlevels is None else nlevels
    t = ParTree(nl, problem.skip if skip is None else skip)
    if device_symbolic:
        t.set_device_symbolic(True)
    t.set_scale(1)
    t.set_tol(problem.tol if tol is None else tol)
    t.set_order(1)
    t.set_square(1 if problem.A.shape[0] == problem.A.shape[1] else 0)
    t.set_precision(precision)
    rect = problem.A.shape[0] != problem.A.shape[1]
    if rect:  # ordered without coordinates: the problem's leaf parts, else the graph bisection
        if getattr(problem, "parts", None) is not None:
            t.set_parts(problem.parts)
    elif getattr(problem, "X", None) is not None:
        t.set_Xcoo(problem.X)
    elif problem.n_side > 0 and problem.dim > 0:
        t.set_Xcoo(linspace_nd(problem.n_side, problem.dim))
    elif getattr(problem, "parts", None) is not None:
        t.set_parts(problem.parts)
    t.partition(problem.A)
    t.assemble(problem.A)
    return t


def build_batch_tree(batch, precision: int = 64, tol=None) -> ParTree:
    """One tree over every system of a Batch (problems.make_batch)."""
    t = ParTree(max(batch.levels), 0)
    t.set_scale(1)
    t.set_tol(batch.tol if tol is None else tol)
    t.set_order(1)
    t.set_square(1)
    t.set_precision(precision)
    t.set_Xcoo(batch.X)
    t.partition_batch(batch.A, batch.sizes, batch.levels)
    t.assemble()
    return t

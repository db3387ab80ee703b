"""ctypes binding of libspaqr_gpu.so (the C ABI in include/spaqr_gpu.h).

The library is loaded from this package directory only.  There is no CPU fallback:
if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspaqr_gpu.so")
_lib = None

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double


class GmresResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("cycles", C.c_int), ("prec_res", C.c_double), ("true_res", C.c_double)]


class GeqrtDesc(C.Structure):
    _fields_ = [("A", _vp), ("T", _vp), ("m", _i), ("n", _i), ("lda", _i), ("pad", _i)]


class LarfbDesc(C.Structure):
    _fields_ = [("V", _vp), ("T", _vp), ("C", _vp), ("R", _vp), ("m", _i), ("n", _i), ("k", _i), ("ldv", _i),
                ("ldc", _i), ("k2", _i), ("ldr", _i), ("pad", _i)]


class TpqrtLink(C.Structure):
    _fields_ = [("B", _vp), ("T", _vp), ("rows", _i), ("ldb", _i)]


class TpqrtDesc(C.Structure):
    _fields_ = [("R", _vp), ("k", _i), ("ldr", _i), ("link0", _i), ("nlinks", _i)]


class TpmqrtLink(C.Structure):
    _fields_ = [("V", _vp), ("T", _vp), ("B", _vp), ("rows", _i), ("ldv", _i), ("ldb", _i), ("pad", _i)]


class TpmqrtDesc(C.Structure):
    _fields_ = [("Atop", _vp), ("k", _i), ("lda", _i), ("ncols", _i), ("link0", _i), ("nlinks", _i), ("col0", _i)]


class Geqp3Desc(C.Structure):
    _fields_ = [("C", _vp), ("Vout", _vp), ("tau", _vp), ("T", _vp), ("vn1", _vp), ("vn2", _vp), ("jpvt", _vp),
                ("rank_out", _vp), ("k", _i), ("w", _i), ("pad0", _i), ("pad1", _i)]


class SparsDesc(C.Structure):
    _fields_ = [("A", _vp), ("Vn", _vp), ("Tn", _vp), ("Vc", _vp), ("Tc", _vp), ("dst", _vp), ("mc", _i), ("nc", _i),
                ("lda", _i), ("kn", _i), ("rn", _i), ("kc", _i), ("rc", _i), ("ldvn", _i), ("ldvc", _i),
                ("ldd", _i), ("ident", _i), ("pad", _i)]


class ColNormDesc(C.Structure):
    _fields_ = [("C", _vp), ("cn", _vp), ("cmax2", _vp), ("k", _i), ("w", _i)]


class SketchDesc(C.Structure):
    _fields_ = [("C", _vp), ("Y", _vp), ("k", _i), ("w", _i), ("s", _i), ("pad", _i)]


class FeatDesc(C.Structure):
    _fields_ = [("Y16", _vp), ("cn", _vp), ("cmax2", _vp), ("x", _vp), ("k", _i), ("w", _i), ("nsout", _i),
                ("nsin", _i), ("level", _i), ("pad", _i)]


class TailDesc(C.Structure):
    _fields_ = [("B", _vp), ("cmax2", _vp), ("ry", _vp), ("rank", _vp), ("k", _i), ("w", _i)]


class TailLeadDesc(C.Structure):
    _fields_ = [("B", _vp), ("cn", _vp), ("cmax2", _vp), ("ry", _vp), ("rank", _vp), ("r", _i), ("w", _i)]


class TailGemmDesc(C.Structure):
    _fields_ = [("M", _vp), ("C", _vp), ("B", _vp), ("r", _i), ("k", _i), ("w", _i), ("ldm", _i)]


EXPORTS = {
    "spaqr_tree_create": (_vp, [_i, _i]),
    "spaqr_tree_destroy": (None, [_vp]),
    "spaqr_tree_set_tol": (_i, [_vp, _d]),
    "spaqr_tree_set_skip": (_i, [_vp, _i]),
    "spaqr_tree_set_scale": (_i, [_vp, _i]),
    "spaqr_tree_set_square": (_i, [_vp, _i]),
    "spaqr_tree_set_order": (_i, [_vp, C.c_float]),
    "spaqr_tree_set_verbose": (_i, [_vp, _i]),
    "spaqr_tree_set_Xcoo": (_i, [_vp, _vp, _i, _i]),
    "spaqr_tree_set_precision": (_i, [_vp, _i]),
    "spaqr_tree_partition": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "spaqr_tree_assemble": (_i, [_vp]),
    "spaqr_tree_factorize": (_i, [_vp]),
    "spaqr_tree_solve": (_i, [_vp, _vp]),
    "spaqr_tree_spmv": (_i, [_vp, _vp, _vp]),
    "spaqr_gmres": (_i, [_vp, _vp, _vp, _i, _d, _i, _d, C.POINTER(GmresResult), _vp, _i, C.POINTER(_i)]),
    "spaqr_fgmres": (_i, [_vp, _vp, _vp, _i, _d, _i, _d, C.POINTER(GmresResult), _vp, _i, C.POINTER(_i)]),
    "spaqr_tree_n": (_i, [_vp]),
    "spaqr_tree_m": (_i, [_vp]),
    "spaqr_tree_set_device_symbolic": (_i, [_vp, _i]),
    "spaqr_tree_symbolic_ms": (_i, [_vp, _vp]),
    "spaqr_tree_get_rperm": (_i, [_vp, _vp]),
    "spaqr_tree_permute_rows": (_i, [_vp, _i, _vp, _vp]),
    "spaqr_tree_get_rows": (_i, [_vp, _vp]),
    "spaqr_tree_solve_ls": (_i, [_vp, _vp, _vp]),
    "spaqr_tree_solve_nrml": (_i, [_vp, _vp, _vp]),
    "spaqr_tree_spmv_t": (_i, [_vp, _vp, _vp]),
    "spaqr_cgls": (_i, [_vp, _vp, _vp, _i, _d, C.POINTER(_i), C.POINTER(_d)]),
    "spaqr_tree_get_cperm": (_i, [_vp, _vp]),
    "spaqr_tree_num_clusters": (_i, [_vp]),
    "spaqr_tree_get_clusters": (_i, [_vp, _vp]),
    "spaqr_tree_level_stats": (_i, [_vp, _vp, _i]),
    "spaqr_tree_diag_R": (_i, [_vp, _i, _vp, _i]),
    "spaqr_tree_factor_ms": (_d, [_vp]),
    "spaqr_tree_stream": (_vp, [_vp]),
    "spaqr_last_error": (C.c_char_p, [_vp]),
    "spaqr_tree_sync": (_i, [_vp]),
    "spaqr_tree_set_profiling": (_i, [_vp, _i]),
    "spaqr_tree_profile": (_i, [_vp, _vp, _i]),
    "spaqr_tree_profile_reset": (_i, [_vp]),
    "spaqr_tree_set_rank_mode": (_i, [_vp, _i]),
    "spaqr_tree_set_rank_model": (_i, [_vp, _vp, _i]),
    "spaqr_tree_set_rank_log": (_i, [_vp, _i]),
    "spaqr_tree_rank_log": (_i, [_vp, _vp, _vp, _i]),
    "spaqr_tree_rank_log_clear": (_i, [_vp]),
    "spaqr_tree_rank_stats": (_i, [_vp, _vp]),
    "spaqr_tree_rank_stats_reset": (_i, [_vp]),
    "spaqr_mlp_params": (_i, []),
    "spaqr_mlp_features": (_i, []),
    "spaqr_mlp_predict": (_i, [_vp, _i, _vp, _vp]),
    "spaqr_mlp_grad": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "spaqr_mlp_train": (_i, [_vp, _vp, _i, _vp, _i, C.c_float, _vp]),
    "spaqr_device_alloc": (_vp, [C.c_size_t]),
    "spaqr_device_free": (_i, [_vp]),
    "spaqr_memcpy_h2d": (_i, [_vp, _vp, C.c_size_t]),
    "spaqr_memcpy_d2h": (_i, [_vp, _vp, C.c_size_t]),
    "spaqr_device_synchronize": (_i, []),
    "spaqr_geqrt_batched": (_i, [_i, _i, _vp, _vp]),
    "spaqr_larfb_batched": (_i, [_i, _i, _vp, _vp]),
    "spaqr_tpqrt_batched": (_i, [_i, _i, _vp, _i, _vp, _vp]),
    "spaqr_tpmqrt_batched": (_i, [_i, _i, _vp, _i, _vp, _vp]),
    "spaqr_geqp3_batched": (_i, [_i, _i, _vp, _d, _d, _vp]),
    "spaqr_spars_apply_batched": (_i, [_i, _i, _vp, _vp]),
    "spaqr_tree_set_rank_retries": (_i, [_vp, _i]),
    "spaqr_tree_rank_log_clusters": (_i, [_vp, _vp, _i]),
    "spaqr_gauss_fill": (_i, [_i, _vp, _i, _i, _i, C.c_uint64, _vp]),
    "spaqr_colnorm_batched": (_i, [_i, _i, _vp, _vp]),
    "spaqr_sketch_batched": (_i, [_i, _i, _vp, _vp, _i, _vp]),
    "spaqr_sketch_bf16_batched": (_i, [_i, _vp, _vp, _i, _i, _vp]),
    "spaqr_tree_set_krylov_mode": (_i, [_vp, _i]),
    "spaqr_tree_set_parts": (_i, [_vp, _vp, _i]),
    "spaqr_tree_get_parts": (_i, [_vp, _vp, _i]),
    "spaqr_mm_read_sparse": (_i, [C.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spaqr_mm_read_dense": (_i, [C.c_char_p, _vp, _vp, _vp]),
    "spaqr_mm_write_sparse": (_i, [C.c_char_p, _i, _i, _vp, _vp, _vp, _i]),
    "spaqr_mm_write_dense": (_i, [C.c_char_p, _i, _i, _vp]),
    "spaqr_rank_features_batched": (_i, [_i, _i, _vp, _d, _vp]),
    "spaqr_tail_select_batched": (_i, [_i, _i, _vp, _d, _vp]),
    "spaqr_tail_lead_batched": (_i, [_i, _i, _vp, _d, _vp]),
    "spaqr_tail_gemm_batched": (_i, [_i, _vp, _vp]),
    "spaqr_shutdown": (_i, []),
    "spaqr_geqp3_batched_noalloc": (_i, [_i, _i, _vp, _d, _d, _vp]),
    "spaqr_geqp3_noalloc_scratch_bytes": (_i, []),
    "spaqr_tree_partition_batch": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "spaqr_tree_num_systems": (_i, [_vp]),
    "spaqr_tree_set_tuning": (_i, [_vp, _i, _i]),
    "spaqr_tree_set_graphs": (_i, [_vp, _i]),
    "spaqr_tree_set_pivot_check": (_i, [_vp, _d]),
    "spaqr_tree_get_info": (_i, [_vp, _vp, _i]),
    "spaqr_tree_solve_load": (_i, [_vp, _vp]),
    "spaqr_tree_solve_phase": (_i, [_vp, _i, _i, _i]),
    "spaqr_tree_solve_store": (_i, [_vp, _vp]),
    "spaqr_tree_permute": (_i, [_vp, _i, _vp, _vp]),
    "spaqr_tree_graph_info": (_i, [_vp, _vp]),
    "spaqr_tree_get_tuning": (_i, [_vp, _i]),
    "spaqr_tree_get_systems": (_i, [_vp, _vp]),
    "spaqr_tree_get_cluster_systems": (_i, [_vp, _vp]),
    "spaqr_gmres_batch": (_i, [_vp, _vp, _vp, _i, _d, _i, _d, _i, _vp]),
}

# Objects holding device memory or a tree (ParTree, DeviceBuffer).  At interpreter exit they are
# released explicitly, while the HIP runtime is certainly alive, and the library's host threads
# are joined: no HIP call is left to finalizers or static destructors.
_live = weakref.WeakSet()


def register(obj):
    _live.add(obj)


@atexit.register
def _shutdown():
    if _lib is None:
        return
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass
    _lib.spaqr_shutdown()


def lib():
    """Load the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libspaqr_gpu.so is not built (run python -m pspaqr_public_amd.build)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class DeviceBuffer:
    """A plain device allocation owned by Python (spaqr_device_alloc / free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().spaqr_device_alloc(max(self.nbytes, 8))
        if not self.ptr:
            raise MemoryError("spaqr_device_alloc(%d) failed" % nbytes)
        register(self)

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if lib().spaqr_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes) != 0:
            raise RuntimeError("H2D copy failed")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if lib().spaqr_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes) != 0:
            raise RuntimeError("D2H copy failed")
        return out

    def free(self):
        if self.ptr:
            lib().spaqr_device_free(self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def check(rc: int, handle=None, what: str = ""):
    if rc != 0:
        msg = lib().spaqr_last_error(handle)
        raise RuntimeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

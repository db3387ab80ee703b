"""The C-ABI library loads (no GPU needed) and exports every symbol include/spaqr_gpu.h
declares; argument checks that return before touching the GPU behave like LAPACK info."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spaqr_gpu.h")
LIB = os.path.join(ROOT, "pspaqr_public_amd", "libspaqr_gpu.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spaqr_[a-zA-Z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def so():
    if not os.path.exists(LIB):
        pytest.fail("libspaqr_gpu.so not built (python -m pspaqr_public_amd.build)")
    return C.CDLL(LIB)


def test_every_declared_symbol_is_exported(so):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from pspaqr_public_amd import _lib

    assert set(declared_functions()) == set(_lib.EXPORTS)


def test_argument_errors_without_gpu(so):
    from pspaqr_public_amd import _lib

    L = _lib.lib()
    assert L.spaqr_geqrt_batched(16, 1, None, None) == -1       # bad precision
    assert L.spaqr_geqrt_batched(64, 1, None, None) == -2       # null descriptors
    d = (_lib.GeqrtDesc * 1)(_lib.GeqrtDesc(None, None, 10, 5, 3, 0))  # lda < m
    assert L.spaqr_geqrt_batched(64, 1, d, None) == -3
    assert L.spaqr_geqp3_batched(64, 0, None, -1.0, 0.0, None) == -4  # negative tolerance
    assert L.spaqr_geqp3_batched(64, 0, None, 0.1, 2.0, None) == -5  # tie band out of range
    assert L.spaqr_geqrt_batched(64, 0, None, None) == 0        # empty batch is a no-op


def test_rank_path_argument_errors_without_gpu(so):
    from pspaqr_public_amd import _lib

    L = _lib.lib()
    assert L.spaqr_colnorm_batched(16, 1, None, None) == -1
    assert L.spaqr_colnorm_batched(64, 1, None, None) == -2
    d = (_lib.ColNormDesc * 1)(_lib.ColNormDesc(None, None, None, 0, 5))  # k < 1
    assert L.spaqr_colnorm_batched(64, 1, d, None) == -3
    s = (_lib.SketchDesc * 1)(_lib.SketchDesc(1, 1, 4, 4, 4, 0))
    assert L.spaqr_sketch_batched(64, 1, s, None, 4, None) == -4   # no Omega
    assert L.spaqr_sketch_batched(64, 1, s, 1, 3, None) == -5      # ldo < w
    # the bf16 sketch: Omega null or misaligned, ldo not a multiple of 8 or below w rounded up to
    # 32, tile width not 64 / 128 / 256
    assert L.spaqr_sketch_bf16_batched(1, s, None, 32, 128, None) == -4
    assert L.spaqr_sketch_bf16_batched(1, s, 4104, 32, 128, None) == -4
    assert L.spaqr_sketch_bf16_batched(1, s, 4096, 36, 128, None) == -5
    assert L.spaqr_sketch_bf16_batched(1, s, 4096, 24, 128, None) == -5
    w40 = (_lib.SketchDesc * 1)(_lib.SketchDesc(1, 1, 4, 40, 4, 0))
    assert L.spaqr_sketch_bf16_batched(1, w40, 4096, 56, 128, None) == -5   # 40 rounds up to 64
    assert L.spaqr_sketch_bf16_batched(1, s, 4096, 32, 96, None) == -6
    assert L.spaqr_sketch_bf16_batched(0, None, 4096, 32, 128, None) == 0
    assert L.spaqr_gauss_fill(8, 1, 10, 10, 10, 1, None) == -1     # precision 16, 32 or 64
    assert L.spaqr_tail_select_batched(64, 0, None, -1.0, None) == -4
    assert L.spaqr_rank_features_batched(32, 0, None, 0.1, None) == 0
    assert L.spaqr_gauss_fill(64, None, 10, 10, 10, 1, None) == -2
    assert L.spaqr_gauss_fill(64, 1, 10, 10, 5, 1, None) == -5    # ld < rows
    assert L.spaqr_tpqrt_batched(64, 0, None, -1, None, None) == -4  # negative link count
    assert L.spaqr_tree_set_rank_retries(None, 1) == -1           # null tree

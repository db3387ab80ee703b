"""The BF16 sketch (spaqr_sketch_bf16_batched and the planner's rank path on FP32 fronts).

* Omega in bf16 (spaqr_gauss_fill prec 16) is the FP32 Omega rounded to nearest even, bit for
  bit (the rounding done here in numpy on the FP32 matrix's bits).
* Y = C Omega equals the FP64 product of the bf16-rounded C and Omega within the FP32
  accumulation bound (w + 32) 2^-24 (|C_b| |Omega_b|) elementwise -- so every product uses the
  right operands, rounded to nearest even -- and the unrounded product within
  (2 u + u^2 + (w + 32) 2^-24) (|C| |Omega|), u = 2^-8 (the bound the header states).  NaN in
  the rows of Omega past w that the kernel reads must not reach Y.  All three tile widths.
* In a tree: the three tile widths give bit-identical factorizations (each entry of Y is
  accumulated in the same order), and bf16 sketches against FP32 sketches reach the same true
  residual 1e-12 with GMRES iteration counts within one and rank sums within 3%.
"""
import numpy as np
import pytest

from oracle.oracle import F, unF
from pspaqr_public_amd import _lib as G
from pspaqr_public_amd import rank_model
from pspaqr_public_amd.problems import make_problem
from pspaqr_public_amd.ptree import build_tree

pytestmark = pytest.mark.gpu

SEED = 0x5EED5CA1E
U = 2.0 ** -8


def bf16_round(x32):
    """float32 -> bf16 bits (uint16), round to nearest even (finite inputs)."""
    b = np.ascontiguousarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return b.astype(np.uint16)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("tile_cols", [64, 128, 256])
def test_bf16_sketch_rounding_bounds(tile_cols):
    rng = np.random.default_rng(11)
    shapes = [(37, 150, 21), (130, 400, 70), (8, 9, 8), (300, 1000, 130), (150, 333, 290), (64, 32, 64)]
    wmax, smax = max(w for _, w, _ in shapes), max(s for _, _, s in shapes)
    ldo = (wmax + 31) // 32 * 32 + 32
    # Omega: FP32 and bf16 from the device generator, ld ldo
    om32 = G.DeviceBuffer(4 * ldo * smax)
    omb = G.DeviceBuffer(2 * ldo * smax)
    assert G.lib().spaqr_gauss_fill(32, om32.ptr, ldo, smax, ldo, SEED, None) == 0
    assert G.lib().spaqr_gauss_fill(16, omb.ptr, ldo, smax, ldo, SEED, None) == 0
    O32 = om32.download(np.float32, ldo * smax).reshape(smax, ldo).T
    Ob_bits = omb.download(np.uint16, ldo * smax).reshape(smax, ldo).T
    ref = bf16_round(O32)
    bad = Ob_bits != ref
    low = O32.view(np.uint32)[bad] & 0xFFFF
    assert not bad.any(), "%d of %d differ; float32 low halves %s; device %s host %s" % (
        bad.sum(), bad.size, [hex(v) for v in low[:8]], Ob_bits[bad][:8], ref[bad][:8])
    # rows past wmax: NaN (read by the w = 1000 task, whose last 32-row step ends at 1024)
    Ob_nan = Ob_bits.copy()
    Ob_nan[wmax:, :] = 0x7FC0
    omn = G.DeviceBuffer.from_array(np.ascontiguousarray(Ob_nan.T).ravel())
    Ob = bf16_value(Ob_bits).astype(np.float64)
    keep, sds, refs = [], [], []
    for (k, w, s) in shapes:
        C32 = rng.standard_normal((k, w)).astype(np.float32)
        cb = G.DeviceBuffer.from_array(F(C32).astype(np.float32))
        yb = G.DeviceBuffer(4 * k * s)
        keep += [cb, yb]
        sds.append(G.SketchDesc(cb.ptr, yb.ptr, k, w, s, 0))
        refs.append(C32)
    assert G.lib().spaqr_sketch_bf16_batched(len(sds), (G.SketchDesc * len(sds))(*sds), omn.ptr, ldo, tile_cols,
                                             None) == 0
    for i, ((k, w, s), C32) in enumerate(zip(shapes, refs)):
        Y = unF(keep[2 * i + 1].download(np.float32, k * s), k, s).astype(np.float64)
        assert np.all(np.isfinite(Y))
        Cb = bf16_value(bf16_round(C32)).astype(np.float64)
        Ys = Cb @ Ob[:w, :s]
        acc = (w + 32) * 2.0 ** -24
        assert np.all(np.abs(Y - Ys) <= acc * (np.abs(Cb) @ np.abs(Ob[:w, :s])) + 1e-30), (k, w, s)
        C, Om = C32.astype(np.float64), O32[:w, :s].astype(np.float64)
        bound = (2 * U + U * U + acc) * (np.abs(C) @ np.abs(Om))
        err = np.abs(Y - C @ Om)
        assert np.all(err <= bound), (k, w, s)
        if w >= 32:  # and the products really are bf16: FP32 inputs would be ~2^-16 closer
            assert np.max(err / np.maximum(np.abs(C) @ np.abs(Om), 1e-30)) > 1e-4


def _tree(P, bf16, tile=0):
    T = build_tree(P, 32)
    T.set_rank_mode(1)
    T.set_rank_model(rank_model.load_weights())
    T.set_tuning({"sketch_bf16": 1 if bf16 else 2, "sketch_tile": tile})
    T.rank_stats_reset()
    T.factorize()
    return T


def test_bf16_tile_widths_are_bit_identical_and_match_fp32_sketches():
    P = make_problem(3, 24)
    b = None
    xs, its = [], []
    for tile in (64, 128, 256):
        T = _tree(P, True, tile)
        b = P.b[T.cperm()]
        xs.append(T.solve(b))
        x, res, _ = T.gmres(b, iters=200, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
        assert res.true_res <= 1e-12
        its.append(res.iters)
        ranks = T.clusters()[:, 12].copy()
        if tile == 64:
            r64 = ranks
        else:
            assert np.array_equal(ranks, r64)
    assert all(np.array_equal(x, xs[0]) for x in xs[1:]) and len(set(its)) == 1
    T32 = _tree(P, False)
    _, res32, _ = T32.gmres(b, iters=200, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
    assert res32.true_res <= 1e-12
    assert abs(res32.iters - its[0]) <= 1
    r32 = T32.clusters()[:, 12]
    assert abs(int(r64.sum()) - int(r32.sum())) <= 0.03 * r32.sum()
    print("[bf16] iters bf16 %d fp32 %d; rank sums %d vs %d" % (its[0], res32.iters, r64.sum(), r32.sum()))

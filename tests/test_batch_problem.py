"""The 48-system batch workload (problems.make_batch): draws, levels and structure."""
import numpy as np

from pspaqr_public_amd.problems import BATCH_SIDES, default_levels, make_batch, mt_canonical


def test_batch_draws_and_levels():
    B = make_batch(count=6, sides=(6, 8, 10))
    assert len(B.systems) == 6 and B.A.shape[0] == sum(B.sizes)
    for i, P in enumerate(B.systems):
        seed = 3 + i
        u = mt_canonical(seed, 2)
        assert P.n_side == (6, 8, 10)[int(u[0] * 3)]
        assert abs(P.meta["sigma"] - (0.5 + 1.5 * u[1])) < 1e-15
        assert P.nlevels == default_levels(P.n_side ** 3) == B.levels[i]
    # block diagonal, each block the system's own matrix
    A = B.A.tocsr()
    for off, P in zip(B.offsets, B.systems):
        n = P.A.shape[0]
        assert (abs(A[off:off + n, off:off + n] - P.A)).max() == 0
        assert A[off:off + n].nnz == P.A.nnz
    assert B.X.shape == (3, B.A.shape[0])
    # deterministic
    B2 = make_batch(count=6, sides=(6, 8, 10))
    assert np.array_equal(B.b, B2.b) and (abs(B.A - B2.A)).max() == 0


def test_batch_level_rule_matches_the_survey():
    # sides 16 / 24 / 32 / 40 / 48 -> 6 / 8 / 9 / 10 / 11 levels by the driver's default rule
    assert [default_levels(s ** 3) for s in BATCH_SIDES] == [6, 8, 9, 10, 11]

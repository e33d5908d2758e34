"""Pins the oracle's restatement of the deterministic FP64 order (kor_det_sum)
on values whose sum depends on the order, so that a different lane fold, tree
pairing, pairing distance order or chunk size gives a different answer; on
exact sums; on the -0.0 identity; and on math.fsum error bounds."""
import math

import numpy as np

from oracle import pyoracle as po

BIG = 2.0 ** 54  # ulp 4: BIG + 1 rounds back to BIG


def test_exact_sums_any_order():
    rng = np.random.default_rng(1)
    for n in (1, 3, 255, 256, 257, 4096, 70_000):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert po.det_sum(v) == float(v.sum())


def test_lane_fold_is_left_to_right():
    # ((BIG + 1) + -BIG) + 1 = 1; a pairwise fold (BIG + 1) + (-BIG + 1) gives 0.
    assert po.det_sum([BIG, 1.0, -BIG, 1.0]) == 1.0


def test_tree_pairs_distance_32_first():
    # slot 0 = BIG, slot 1 = 1, slot 32 = -BIG: distance 32 cancels BIG first
    # (result 1); a sequential fold, or distances 1..32, absorbs the 1 (result 0).
    v = np.zeros(256)
    v[0], v[4], v[128] = BIG, 1.0, -BIG
    assert po.det_sum(v) == 1.0
    assert math.fsum(v) == 1.0 and float(np.cumsum(v)[-1]) == 0.0


def test_chunks_of_256():
    # Chunk 0 = BIG (its 1 at item 255 is absorbed by the tree), chunk 1 = -BIG:
    # 0. Chunks of 512 would cancel BIG at distance 32 first and give 1.
    v = np.zeros(257)
    v[0], v[255], v[256] = BIG, 1.0, -BIG
    assert po.det_sum(v) == 0.0


def test_negative_zero_identity():
    neg = lambda x: math.copysign(1.0, x) < 0
    assert neg(po.det_sum([]))
    assert neg(po.det_sum([-0.0]))
    assert neg(po.det_sum([-0.0] * 300))
    assert not neg(po.det_sum([0.0, -0.0]))
    assert po.det_sum([5.0] + [-0.0] * 1000) == 5.0


def test_error_bound_vs_fsum():
    rng = np.random.default_rng(2)
    for n in (10, 1000, 100_000):
        v = rng.normal(0, 1e6, n)
        bound = 2 * math.ceil(math.log(n, 256) + 1) * 10 * np.finfo(float).eps * np.abs(v).sum()
        assert abs(po.det_sum(v) - math.fsum(v)) <= bound

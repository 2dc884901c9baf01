"""The per-antenna mask algebra of the antenna-factorized Viss form (CPU).

A half row (a, b) takes a pixel with mask 1 when both endpoint cones hold the
tile, 0 when either misses it, else the product of the two per-pixel
predicates (vw_pair_class, vissfast .cuh:521-527).  The factorized form sums
with a product of per-antenna masks: m = hidden ? 0 : 1 in a tile where no
antenna is partly visible, else m = hidden ? 0 : predicate; where an
all-visible antenna's predicate fails on a pixel, a correction pass adds
o_a o_b - o_a p_a o_b p_b (o: all-visible).  These tests check that the sum
equals the row mask for every pair and pixel, including bit patterns the cone
bound never produces (all-visible antennas with failing predicates).
"""
import itertools

import numpy as np

HIDDEN, ALLVIS, PARTLY = 2, 1, 0


def row_mask(ba, bb, pa, pb):
    if (ba | bb) & 2:
        return 0
    if ba & bb & 1:
        return 1
    return pa & pb


def factorized_mask(bits, pred):
    """[n_ant, n_ant, n_pix] pair weights of the factorized sum, correction included."""
    vis = (bits != HIDDEN)[:, None]
    o = (bits == ALLVIS)[:, None].astype(int)
    if np.any(bits == PARTLY):
        m = (vis & (pred == 1)).astype(int)
        failed = np.any((o == 1) & (pred == 0))
    else:
        m = np.broadcast_to(vis, pred.shape).astype(int)
        failed = False
    w = m[:, None, :] * m[None, :, :]
    if failed:
        op = o * pred
        w = w + o[:, None, :] * o[None, :, :] - op[:, None, :] * op[None, :, :]
    return w


def _check(bits, pred):
    w = factorized_mask(bits, pred)
    for a, b in itertools.combinations(range(bits.size), 2):
        want = [row_mask(int(bits[a]), int(bits[b]), int(pred[a, p]), int(pred[b, p])) for p in range(pred.shape[1])]
        assert np.array_equal(w[a, b], want), (bits, a, b)


def test_factorized_mask_with_failing_predicates():
    rng = np.random.default_rng(5)
    for _ in range(60):
        bits = rng.choice([HIDDEN, ALLVIS, PARTLY], size=9)
        _check(bits, rng.integers(0, 2, size=(9, 300)))


def test_factorized_mask_consistent_cones():
    # all-visible antennas whose predicates hold everywhere (what the cone
    # bound guarantees): no correction pass is needed
    rng = np.random.default_rng(6)
    for _ in range(60):
        bits = rng.choice([HIDDEN, ALLVIS, PARTLY], size=9)
        pred = rng.integers(0, 2, size=(9, 300))
        pred[bits == ALLVIS] = 1
        _check(bits, pred)


def test_no_partly_visible_antenna_needs_no_predicate():
    rng = np.random.default_rng(7)
    for _ in range(20):
        bits = rng.choice([HIDDEN, ALLVIS], size=9)
        _check(bits, rng.integers(0, 2, size=(9, 300)))

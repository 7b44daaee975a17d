"""The score bound that picks FP16 R0 operands (bpmax_score_bound; host only, no GPU).

A pair runs R0 on FP16 when the bound times 2^p is at most 2^11, so the bound must never be
below a score: checked against brute-force maximum matchings of small base multisets under
random tables, and against the oracle's largest F cell.  That it is the fractional matching
optimum is checked against the matching itself where the two coincide: under the default
table the pairing graph (G-C, G-U, A-U) is bipartite, so the optimum is integral and a
greedy matching reaches it."""
import numpy as np
import pytest

import bpmaxcpu_amd as B
from oracle import oracle as O


def _max_matching(codes, W):
    """Largest total weight of disjoint pairs over the multiset `codes` (brute force)."""
    codes = list(codes)
    if len(codes) < 2:
        return 0.0
    x, rest = codes[0], codes[1:]
    best = _max_matching(rest, W)  # x unpaired
    for k in range(len(rest)):
        best = max(best, W[x][rest[k]] + _max_matching(rest[:k] + rest[k + 1:], W))
    return best


def _pair_weights(intra, inter):
    t, e = np.asarray(intra, float).reshape(5, 5), np.asarray(inter, float).reshape(5, 5)
    return np.maximum(np.maximum(t, t.T), np.maximum(e, e.T))


def test_bound_is_never_below_a_matching():
    rng = np.random.default_rng(3)
    for trial in range(60):
        intra = rng.integers(0, 9, 25).astype(np.float32) / 2  # multiples of 0.5 in [0, 4]
        inter = rng.integers(0, 9, 25).astype(np.float32) / 2
        a, b = rng.integers(0, 5, rng.integers(1, 5)), rng.integers(0, 5, rng.integers(1, 6))
        W = _pair_weights(intra, inter)
        m = _max_matching(np.concatenate([a, b]), W)
        assert B.score_bound(a, b, intra, inter) >= m, trial


def test_bound_covers_every_oracle_cell():
    rng = np.random.default_rng(4)
    for trial in range(8):
        intra = rng.integers(0, 7, 25).astype(np.float32) / 2
        inter = rng.integers(0, 7, 25).astype(np.float32) / 2
        a, b = rng.integers(0, 5, 6).astype(np.int32), rng.integers(0, 5, 24).astype(np.int32)
        F, s = O.bpmax(a, b, intra, inter)
        assert np.max(F[np.isfinite(F)]) <= B.score_bound(a, b, intra, inter), trial


def test_bound_is_the_matching_optimum_under_the_default_table():
    t = B.reference_table()
    rng = np.random.default_rng(6)
    for trial in range(40):
        gc = rng.uniform(0.1, 0.9)
        p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
        a = rng.choice(4, size=int(rng.integers(1, 65)), p=p)
        b = rng.choice(4, size=int(rng.integers(1, 1025)), p=p)
        nA, nC, nG, nU = np.bincount(np.concatenate([a, b]), minlength=4)
        gcp = min(nG, nC)
        gup = min(nG - gcp, nU)
        aup = min(nA, nU - gup)
        assert B.score_bound(a, b, t, t) == 3 * gcp + 1.5 * gup + aup, trial


def test_bound_is_reached_at_the_edge_of_fp16():
    """GGGGCCCC against G^336 C^335 U^3: the oracle's score equals the bound, 1018.5 (2037
    halves, below 2^11: FP16); two more bases of G and C give 1026 (2052: FP32)."""
    t = B.reference_table()
    a = np.array([2, 2, 2, 2, 1, 1, 1, 1], np.int32)
    b = np.array([2] * 336 + [1] * 335 + [3] * 3, np.int32)
    assert B.score_bound(a, b, t, t) == 1018.5
    assert O.bpmax(a, b, t, t)[1] == 1018.5
    assert B.score_bound(a, np.array([2] * 338 + [1] * 338), t, t) == 1026.0


def test_bound_rejects_bad_input():
    with pytest.raises(B.BPMaxError):
        B.score_bound([0, 5], [1])
    with pytest.raises(B.BPMaxError):
        B.score_bound([], [1])
    neg = B.reference_table().copy()
    neg[0, 3] = -1
    with pytest.raises(B.BPMaxError):
        B.score_bound([0], [3], neg, neg)

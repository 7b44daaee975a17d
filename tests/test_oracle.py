"""Pins the CPU oracle (oracle/bpmax_oracle.c) before it is trusted as the checker.

* every golden vector written by the reference's own bpmax() (tests/golden/) -- whole
  F tables bit-for-bit for small shapes, SHA-256 of the whole table for the reference's
  sweep shapes (and 64x512 with BPMAX_SLOW=1);
* planted mistakes (dropped terms, a lost k2+1 shift, transposed operands, an unreversed
  strand, a wrong hairpin threshold) must each break at least one fixture;
* S1/S2 against brute-force enumeration of non-crossing pairings;
* the isolated R0 product against a numpy brute force;
* invariants the recurrence implies (monotonicity, superadditivity, exact halves).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_io as G

FULL = G.full_fixtures()
HASHED = G.hash_fixtures()


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                                                 np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("path", FULL, ids=[G.fid(p) for p in FULL])
def test_oracle_full_fixture(path):
    fx = G.load_full(path)
    F, score = O.bpmax(fx["seq1"], fx["seq2"], fx["intra"], fx["inter"])
    assert score == fx["score"]
    assert _bits_equal(F, fx["F"])


@pytest.mark.parametrize("path", HASHED, ids=[G.fid(p) for p in HASHED])
def test_oracle_hash_fixture(path):
    fx = G.load_hash(path)
    if fx["M"] * fx["N"] > 20000:
        pytest.skip("64x512 takes ~45 s on 8 cores: test_oracle_64x512_fixture, BPMAX_SLOW=1")
    F, score = O.bpmax(fx["seq1"], fx["seq2"], fx["intra"], fx["inter"])
    assert score == fx["score"]
    for i1, j1, i2, j2, v in fx["samples"]:
        assert F[O.packed_index(i1, j1, i2, j2, fx["M"], fx["N"])] == v
    assert G.sha256_f32(F) == fx["sha256"]


@pytest.mark.slow
def test_oracle_64x512_fixture():
    for t in ("ref", "321"):
        fx = G.load_hash(f"{G.GOLDEN}/hash_64x512_{t}.json")
        F, score = O.bpmax(fx["seq1"], fx["seq2"], fx["intra"], fx["inter"])
        assert score == fx["score"] and G.sha256_f32(F) == fx["sha256"]


@pytest.mark.parametrize("mut", sorted(O.MUT), ids=sorted(O.MUT))
def test_planted_mistake_is_caught(mut):
    caught = False
    for name in ("full_16x100_asym", "full_4x64_ref", "full_5x17_asym", "full_3x65_321"):
        fx = G.load_full(f"{G.GOLDEN}/{name}.npz")
        F, _ = O.bpmax(fx["seq1"], fx["seq2"], fx["intra"], fx["inter"], mutation=O.MUT[mut])
        if not _bits_equal(F, fx["F"]):
            caught = True
            break
    assert caught, f"planted mistake {mut} not detected by the fixtures"


# ---------------------------------------------------------------- S1/S2 brute force
def _matchings(i, j):
    """All non-crossing sets of pairs (p,q), q-p >= 4, inside [i, j] (explicit enumeration)."""
    if j - i < 4:
        return [()]
    out = [m for m in _matchings(i + 1, j)]             # i unpaired
    for k in range(i + 4, j + 1):                        # i paired with k
        for inner in _matchings(i + 1, k - 1):
            for rest in _matchings(k + 1, j):
                out.append(((i, k),) + inner + rest)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_single_strand_vs_enumeration(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 11))
    x = rng.integers(0, 5, n).astype(np.int32)
    intra = O.table_asym()[0] if seed % 2 else O.reference_table()
    S = O.single_strand(x, intra)
    for i in range(n):
        for j in range(i, n):
            best = max(sum(float(intra[x[p], x[q]]) for p, q in m) for m in _matchings(i, j))
            assert S[i, j] == best, (i, j)
        for j in range(i):
            assert S[i, j] == -np.inf


def test_s2_is_single_strand_of_reversed_seq2():
    """S1+S2 is the no-interaction floor: F(i1,j1,i2,j2) >= S1(i1,j1)+S2(i2,j2), with
    equality when no inter pair can help (inter table all zero)."""
    rng = np.random.default_rng(3)
    s1, s2 = rng.integers(0, 4, 6).astype(np.int32), rng.integers(0, 4, 20).astype(np.int32)
    intra = O.reference_table()
    zero = np.zeros((5, 5), np.float32)
    F, score = O.bpmax(s1, s2, intra, zero)
    S1 = O.single_strand(s1, intra)
    S2 = O.single_strand(s2[::-1].copy(), intra)
    M, N = 6, 20
    for i1 in range(M):
        for j1 in range(i1, M):
            for i2 in range(N):
                for j2 in range(i2, N):
                    v = F[O.packed_index(i1, j1, i2, j2, M, N)]
                    base = 0.0 if (i1 == j1 and i2 == j2) else S1[i1, j1] + S2[i2, j2]
                    assert v == base


# ------------------------------------------------------------------ R0 brute force
@pytest.mark.parametrize("n", [1, 2, 7, 33])
def test_maxplus_tri_vs_numpy(n):
    rng = np.random.default_rng(n)
    up = np.triu(np.ones((n, n), bool))
    A = np.where(up, rng.integers(0, 201, (n, n)), -np.inf).astype(np.float32)
    B = np.where(up, rng.integers(0, 201, (n, n)), -np.inf).astype(np.float32)
    C = np.where(up, rng.integers(0, 201, (n, n)), -np.inf).astype(np.float32)
    got = O.maxplus_tri(A, B, C)
    S = A[:, :, None] + B[None, :, :]                  # S[i,k,j]
    mask = up[:, :, None] & up[None, :, :]             # i<=k and k<=j
    want = np.maximum(C, np.where(mask, S, -np.inf).max(axis=1))
    assert _bits_equal(np.where(up, got, 0), np.where(up, want, 0))


# ------------------------------------------------------------------ invariants
def test_invariants_random_pair():
    rng = np.random.default_rng(11)
    M, N = 6, 40
    s1, s2 = rng.integers(0, 4, M).astype(np.int32), rng.integers(0, 4, N).astype(np.int32)
    tab = O.reference_table()
    F, score = O.bpmax(s1, s2, tab, tab)
    b = s2[::-1]
    idx = lambda i1, j1, i2, j2: O.packed_index(i1, j1, i2, j2, M, N)  # noqa: E731
    assert np.all(F * 2 == np.round(F * 2))                            # exact halves
    for i1 in range(M):
        for i2 in range(N):
            assert F[idx(i1, i1, i2, i2)] == tab[s1[i1], b[i2]]         # base case
    for _ in range(400):
        i1 = int(rng.integers(0, M)); j1 = int(rng.integers(i1, M))
        i2 = int(rng.integers(0, N)); j2 = int(rng.integers(i2, N))
        v = F[idx(i1, j1, i2, j2)]
        if j2 > i2:
            assert v >= F[idx(i1, j1, i2 + 1, j2)] and v >= F[idx(i1, j1, i2, j2 - 1)]
        if j1 > i1:
            assert v >= F[idx(i1 + 1, j1, i2, j2)] and v >= F[idx(i1, j1 - 1, i2, j2)]
            k1 = int(rng.integers(i1, j1))
            if j2 > i2:
                k2 = int(rng.integers(i2, j2))
                assert v >= F[idx(i1, k1, i2, k2)] + F[idx(k1 + 1, j1, k2 + 1, j2)]
    assert score == F[idx(0, M - 1, 0, N - 1)]


def _small_set_pairs(workload, max_cells):
    out = []
    for name in G.set_fixtures(workload):
        fx = G.load_set(name)
        for p in fx["pairs"]:
            if O.tri(p["M"]) * O.tri(p["N"]) <= max_cells:
                out.append((name, fx["intra"], fx["inter"], p))
    return out


@pytest.mark.parametrize("workload,max_cells", [("batch4", 4.0e6), ("serve5", 1.0e7), ("serve5-1024", 0)])
def test_oracle_on_workload_fixture_sets(workload, max_cells):
    """The oracle against the reference's tables for the pairs of the batched and serving
    fixture sets that it finishes in a few seconds, and for the 8x1024 pair (reference
    table): N = 1024, the largest inner length of the serving workload."""
    if workload == "serve5-1024":
        cases = [c for c in _small_set_pairs("serve5", 2.0e7) if (c[3]["M"], c[3]["N"]) == (8, 1024)
                 and c[0].endswith("_ref.json")]
        assert len(cases) == 1
    else:
        cases = _small_set_pairs(workload, max_cells)
        assert len(cases) >= 2
    for name, intra, inter, p in cases:
        F, score = O.bpmax(p["seq1"], p["seq2"], intra, inter)
        assert score == p["score"], (name, p["M"], p["N"])
        assert G.sha256_f32(F) == p["sha256"], (name, p["M"], p["N"])

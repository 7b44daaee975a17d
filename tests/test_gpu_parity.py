"""GPU parity: the HIP engine against the reference's golden vectors and the CPU oracle,
bit-for-bit.  Every test calls through the C ABI (libbpmax_hip.so via ctypes)."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import bpmaxcpu_amd as B
from oracle import oracle as O
from tests import golden_io as G

pytestmark = pytest.mark.gpu

FULL = G.full_fixtures()
HASHED = G.hash_fixtures()


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = B.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module",
                params=["pipe", "pipe-f32", "pipe-eager", "pipe-ahead", "pipe-push", "pipe-t32", "pipe-r0whole",
                        "launch-narrow", "launch-wide", "launch-ahead", "tasks"])
def eng2(request):
    """An engine under each schedule:
    * pipe: the two-stream pipeline with the default settings (inner closure in one launch
      per diagonal -- cross-tile steps carried by the previous diagonal's launch and by the
      in-tile workgroups -- and R0 items of one K each), graph-replayed;
    * pipe-eager: the pipeline with a cross-tile and an in-tile launch per diagonal, in-tile
      launches 1024-lane up to 256 tiles, blocked above; pipe-ahead: the default closure,
      eager; pipe-push: every in-tile launch in the 512-lane push form;
    * launch-*: per-phase launches on one stream with the in-tile closure always blocked
      (narrow; R0 items over every K of a tile), always 1024-lane (wide; R0 items of two K
      each, the last one ragged) or in one launch per diagonal (ahead);
    * pipe-t32: the default pipeline with 32 x 32 storage tiles for every N (the tuner's
      tile candidate);
    * pipe-r0whole: the pipeline with each outer diagonal's R0 run whole after the closure
      of the diagonal below (instead of its bulk one diagonal ahead), in 2 ranges;
    * pipe-f32: the default pipeline with FP32 R0 operands for every pair (the others use the
      FP16 mirrors wherever the score bound allows, which is every fixture here);
    * tasks: the dependency-tile scheduler (FP32 R0 operands)."""
    push = request.param.endswith("push")
    env = {"BPMAX_INTRA_BLOCKED": "0"} if push else ({"BPMAX_R0_AHEAD": "0"} if request.param == "pipe-r0whole" else {})
    os.environ.update(env)  # read when the context is created
    try:
        e = B.Engine(0)
    finally:
        for k in env:
            os.environ.pop(k, None)
    sched = request.param.split("-")[0]
    ahead = request.param.endswith("ahead")
    if sched == "launch":
        narrow = request.param.endswith("narrow")
        e.set_default_tuning({"schedule": "launch", "outer_split": 1, "outer_items": 2048, "graphs": False,
                              "intra_wide": 0 if narrow else 1 << 30, "outer_kgroup": 0 if narrow else 2,
                              "closure_ahead": ahead})
    elif request.param == "pipe-t32":
        e.set_default_tuning({"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "graphs": False,
                              "intra_wide": 256, "outer_kgroup": 1, "closure_ahead": True, "tile": 32})
    elif sched == "pipe" and request.param != "pipe":
        e.set_default_tuning({"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "graphs": False,
                              "intra_wide": 0 if push else 256, "outer_kgroup": 1, "closure_ahead": ahead})
    e.set_schedule(sched)
    e.set_graphs(request.param in ("pipe", "pipe-f32"))
    e.set_r0_half(request.param != "pipe-f32")
    yield e
    e.close()


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("path", FULL, ids=[G.fid(p) for p in FULL])
def test_full_fixture(eng2, path):
    fx = G.load_full(path)
    _say(f"full fixture {G.fid(path)}")
    F, score = eng2.solve(fx["seq1"], fx["seq2"], fx["intra"], fx["inter"])
    assert score == fx["score"]
    assert _bits_equal(F, fx["F"])


@pytest.mark.parametrize("path", HASHED, ids=[G.fid(p) for p in HASHED])
def test_hash_fixture(eng2, path):
    fx = G.load_hash(path)
    _say(f"hash fixture {G.fid(path)} M={fx['M']} N={fx['N']}")
    F, score = eng2.solve(fx["seq1"], fx["seq2"], fx["intra"], fx["inter"])
    assert score == fx["score"]
    for i1, j1, i2, j2, v in fx["samples"]:
        assert F[O.packed_index(i1, j1, i2, j2, fx["M"], fx["N"])] == v
    assert G.sha256_f32(F) == fx["sha256"]


# shapes around the tile edges (T = 32 for N <= 64, 64 above), the P2/P1 hairpin
# thresholds (d <= 3), single-base strands and N below the reference's own minimum of 8
EDGE = [(1, 1), (1, 2), (2, 1), (1, 5), (5, 1), (3, 4), (4, 5), (6, 8), (2, 31), (3, 32), (2, 33),
        (1, 64), (2, 65), (3, 63), (7, 64), (2, 96), (2, 127), (2, 128), (3, 129), (9, 70)]


@pytest.mark.parametrize("M,N", EDGE)
def test_edge_shapes_vs_oracle(eng2, M, N):
    rng = np.random.default_rng(1000 * M + N)
    s1 = rng.integers(0, 5, M).astype(np.int32)   # includes code 4 (N, never pairs)
    s2 = rng.integers(0, 5, N).astype(np.int32)
    intra, inter = O.table_asym()
    _say(f"edge M={M} N={N}")
    F, score = eng2.solve(s1, s2, intra, inter)
    Fo, so = O.bpmax(s1, s2, intra, inter)
    assert score == so
    assert _bits_equal(F, Fo)


@pytest.mark.parametrize("seed", range(4))
def test_random_shapes_vs_oracle(eng, seed):
    rng = np.random.default_rng(seed)
    M = int(rng.integers(2, 14))
    N = int(rng.integers(20, 200))
    s1 = rng.integers(0, 4, M).astype(np.int32)
    s2 = rng.integers(0, 4, N).astype(np.int32)
    tab = O.table_321() if seed % 2 else O.reference_table()
    _say(f"random seed={seed} M={M} N={N}")
    F, score = eng.solve(s1, s2, tab, tab)
    Fo, so = O.bpmax(s1, s2, tab, tab)
    assert score == so and _bits_equal(F, Fo)


def test_dense_layout_matches_packed(eng):
    rng = np.random.default_rng(5)
    M, N = 5, 70
    s1, s2 = rng.integers(0, 4, M).astype(np.int32), rng.integers(0, 4, N).astype(np.int32)
    _say("dense layout 5x70")
    Fp, sp = eng.solve(s1, s2, layout=B.LAYOUT_PACKED)
    Fd, sd = eng.solve(s1, s2, layout=B.LAYOUT_DENSE_4D)
    _, sn = eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    assert sp == sd == sn
    k = 0
    for i1 in range(M):
        for j1 in range(M):
            if j1 < i1:
                assert np.all(Fd[i1, j1] == 0)
                continue
            for i2 in range(N):
                row = Fp[k:k + N - i2]
                k += N - i2
                assert _bits_equal(Fd[i1, j1, i2, i2:], row)
                assert np.all(Fd[i1, j1, i2, :i2] == 0)


def test_batch_matches_single_and_oracle(eng2):
    eng = eng2
    rng = np.random.default_rng(77)
    pairs = [(int(rng.integers(1, 12)), int(rng.integers(1, 150))) for _ in range(12)]
    s1s = [rng.integers(0, 4, m).astype(np.int32) for m, _ in pairs]
    s2s = [rng.integers(0, 4, n).astype(np.int32) for _, n in pairs]
    _say(f"batch of {len(pairs)} mixed pairs")
    scores = eng.solve_batch(s1s, s2s)
    for p in range(len(pairs)):
        _, so = O.bpmax(s1s[p], s2s[p])
        _, ss = eng.solve(s1s[p], s2s[p], layout=B.LAYOUT_NONE)
        assert scores[p] == so == ss, (p, pairs[p])


def test_invalid_inputs_are_rejected(eng):
    _say("invalid inputs")
    with pytest.raises(B.BPMaxError) as e:
        eng.solve(np.array([0, 1, 7], np.int32), np.array([0, 1, 2], np.int32))
    assert e.value.code == B.EINVAL
    with pytest.raises(B.BPMaxError) as e:
        eng.solve(np.zeros(0, np.int32), np.array([0, 1, 2], np.int32))
    assert e.value.code == B.EINVAL
    bad = O.reference_table().copy()
    bad[0, 3] = -1.0
    with pytest.raises(B.BPMaxError) as e:
        eng.solve(np.array([0, 1, 2], np.int32), np.array([3, 2, 1], np.int32), bad)
    assert e.value.code == B.EWEIGHTS
    bad[0, 3] = 0.1
    with pytest.raises(B.BPMaxError) as e:
        eng.solve(np.array([0, 1, 2], np.int32), np.array([3, 2, 1], np.int32), bad)
    assert e.value.code == B.EWEIGHTS
    # the context is still usable afterwards
    _, s = eng.solve(np.array([0], np.int32), np.array([3], np.int32))
    assert s == 1.0


def test_r0_operand_precision_follows_the_score_bound(eng):
    """R0 runs on FP16 operands exactly when every score is an FP16 integer multiple of 2^-p
    and the pair has 64 x 64 tiles: by length, wmax * ((M+N)/2 + 1) * 2^p <= 2048, or by
    sequence, 2^p times the pair's matching bound (B.score_bound) <= 2048.  F and the score
    are bit-identical to the oracle either way."""
    rng = np.random.default_rng(5)
    a, b = rng.integers(0, 4, 16).astype(np.int32), rng.integers(0, 4, 100).astype(np.int32)
    intra = B.reference_table()  # GC 3, AU 1, GU 1.5: p = 1
    inter = intra
    # bounds at 16 x 100 ((16+100)/2 + 1 = 59 pairs): 3 * 59 * 2^1 = 354.  x16 (GU 24, p = 0):
    # by length 48 * 59 = 2832, by sequence 16 * 101 = 1616 (this draw's matching bound is
    # 101: 23 G, 29 C, 32 A, 32 U); x32: 3232 by sequence; x0.25: 0.75 * 59 * 2^3 = 354 (p = 3)
    assert B.score_bound(a, b, intra, inter) == 101.0
    cases = [(intra, inter, True), (intra * 16, inter * 16, True), (intra * 32, inter * 32, False),
             (intra * 0.25, inter * 0.25, True)]
    for wi, we, expect in cases:
        _say(f"precision 16x100 wmax {float(max(wi.max(), we.max()))}")
        F, sc = eng.solve(a, b, wi, we)
        assert eng.r0_half_used() == expect
        Fo, so = O.bpmax(a, b, wi, we)
        assert sc == so and _bits_equal(F, Fo)
    # N <= 64 pairs use 32 x 32 tiles: FP32 operands
    eng.solve(a, b[:60], intra, inter)
    assert not eng.r0_half_used()
    eng.set_r0_half(False)
    try:
        F, sc = eng.solve(a, b, intra, inter)
        assert not eng.r0_half_used()
        Fo, so = O.bpmax(a, b, intra, inter)
        assert sc == so and _bits_equal(F, Fo)
    finally:
        eng.set_r0_half(True)


def test_fp16_at_the_edge_of_its_exact_range(eng):
    """8 x 674 with scores up to 1018.5 (2037 halves, FP16 holds integers to 2048): GGGGCCCC
    against G^336 C^335 U^3.  The length bound (3 * 342 * 2 = 2052) rules FP16 out; the
    matching bound (339 G-C + 1 G-U = 1018.5, the score itself: 2037 halves) lets it run, and F
    must equal the oracle's bit for bit, 127 of its cells above 900 being half-integers.  Two
    more bases (8 x 676, all G or C: 2052 halves) is past both bounds: FP32 operands."""
    a = np.array([2, 2, 2, 2, 1, 1, 1, 1], np.int32)
    b = np.array([2] * 336 + [1] * 335 + [3] * 3, np.int32)
    t = B.reference_table()
    _say("fp16 edge 8x674")
    F, sc = eng.solve(a, b, t, t)
    assert eng.r0_half_used()
    Fo, so = O.bpmax(a, b, t, t)
    assert so == 1018.5 and sc == so and _bits_equal(F, Fo)
    del F, Fo
    b2 = np.array([2] * 338 + [1] * 338, np.int32)
    _say("fp16 edge 8x676")
    _, sc2 = eng.solve(a, b2, t, t, layout=B.LAYOUT_NONE)
    assert not eng.r0_half_used()
    _, so2 = O.bpmax(a, b2, t, t)
    assert sc2 == so2


@pytest.mark.parametrize("half", [False, True], ids=["fp32", "fp16"])
@pytest.mark.parametrize("n,batch", [(1, 1), (5, 3), (64, 2), (65, 2), (130, 3), (512, 2)])
def test_r0_isolated_vs_oracle(eng, n, batch, half):
    """C max= A (max,+) B on tiled triangles against the oracle, on FP32 operands and on their
    FP16 copies (integers of magnitude <= 1024: every sum below 2^11 is exact in FP16)."""
    rng = np.random.default_rng(n + batch)
    up = np.triu(np.ones((n, n), bool))
    mk = lambda: np.where(up, rng.integers(0, 201, (batch, n, n)), -np.inf).astype(np.float32)  # noqa: E731
    A, Bm, C = mk(), mk(), mk()
    if half:  # the whole exact range, negative values included
        A = np.where(up, rng.integers(-1024, 1025, (batch, n, n)), -np.inf).astype(np.float32)
    _say(f"r0 isolated n={n} batch={batch} half={half}")
    bufs = [eng.tri_alloc(n, batch) for _ in range(5 if half else 3)]
    dA, dB, dC = bufs[:3]
    try:
        eng.tri_upload(A, dA)
        eng.tri_upload(Bm, dB)
        eng.tri_upload(C, dC)
        if half:
            assert eng.tri_to_half(dA, bufs[3], n, batch) and eng.tri_to_half(dB, bufs[4], n, batch)
            eng.r0_tri_batch_half(n, batch, bufs[3], bufs[4], dC)
        else:
            eng.r0_tri_batch(n, batch, dA, dB, dC)
        got = eng.tri_download(dC, n, batch)
    finally:
        for p in bufs:
            eng.tri_free(p)
    for m in range(batch):
        want = O.maxplus_tri(A[m], Bm[m], C[m])
        assert _bits_equal(np.where(up, got[m], 0), np.where(up, want, 0))
        assert np.all(got[m][~up] == -np.inf)


@pytest.mark.parametrize("value,exact", [(0.5, False), (1025.0, False), (-1025.0, False), (np.inf, False),
                                         (1024.0, True), (-1024.0, True), (-np.inf, True)])
def test_tri_to_half_exactness(eng, value, exact):
    """bpmax_tri_to_half reports exact copies only when every value is -inf or an integer of
    magnitude <= 1024: one planted value decides, wherever it sits in the batch; the copy
    itself holds the values (checked through the FP16 R0 against the oracle when exact)."""
    n, batch = 130, 2
    rng = np.random.default_rng(23)
    up = np.triu(np.ones((n, n), bool))
    A = np.where(up, rng.integers(-1000, 1001, (batch, n, n)), -np.inf).astype(np.float32)
    Bm = np.where(up, rng.integers(-1000, 1001, (batch, n, n)), -np.inf).astype(np.float32)
    C = np.full((batch, n, n), -np.inf, np.float32)
    A[1, 70, 129] = value
    _say(f"tri_to_half planted {value}")
    bufs = [eng.tri_alloc(n, batch) for _ in range(5)]
    try:
        eng.tri_upload(A, bufs[0])
        eng.tri_upload(Bm, bufs[1])
        eng.tri_upload(C, bufs[2])
        assert eng.tri_to_half(bufs[0], bufs[3], n, batch) == exact
        assert eng.tri_to_half(bufs[1], bufs[4], n, batch)
        if exact:
            eng.r0_tri_batch_half(n, batch, bufs[3], bufs[4], bufs[2])
            got = eng.tri_download(bufs[2], n, batch)
            for m in range(batch):
                want = O.maxplus_tri(A[m], Bm[m], C[m])
                assert _bits_equal(np.where(up, got[m], 0), np.where(up, want, 0))
    finally:
        for p in bufs:
            eng.tri_free(p)


def _maxplus_tri_np(A, Bm, C):
    """C max= A (max,+) B on one upper-triangular pair, row by row in numpy (k in [i, j]:
    B's -inf lower half masks k > j, A's masks k < i)."""
    n = A.shape[0]
    out = C.copy()
    for i in range(n):
        prod = (A[i, i:, None] + Bm[i:, i:]).max(axis=0)
        out[i, i:] = np.maximum(out[i, i:], prod)
    return out


def test_r0_isolated_at_full_batch_size(eng):
    """The isolated R0 kernel at its benchmark size, 4096 tiled 512x512 triangles of
    synthetic integers (device-generated hash values, checked against their host formula),
    with four whole output matrices checked against a numpy max-plus."""
    n, batch = 512, 4096
    _say(f"r0 isolated n={n} batch={batch}")
    tri = eng.tri_tiled_floats(n)
    dA, dB, dC = eng.tri_alloc(n, batch), eng.tri_alloc(n, batch), eng.tri_alloc(n, batch)
    try:
        for d, seed in ((dA, 1), (dB, 2), (dC, 3)):
            eng.tri_fill_random(d, n, batch, seed, 0, 200)
        picks = [0, 1, 2047, 4095]
        before = {m: [eng.tri_download(d + 4 * m * tri, n, 1)[0] for d in (dA, dB, dC)] for m in picks}
        eng.r0_tri_batch(n, batch, dA, dB, dC)
        for m in picks:
            A, Bm, C = before[m]
            want = _maxplus_tri_np(A, Bm, C)
            got = eng.tri_download(dC + 4 * m * tri, n, 1)[0]
            assert _bits_equal(got, want), m
    finally:
        for p in (dA, dB, dC):
            eng.tri_free(p)


def test_r0_isolated_fp16_at_full_batch_size(eng):
    """The FP16-operand R0 at the benchmark size: 4096 tiled 512x512 triangles of synthetic
    integers in [0, 200], copied to FP16 (exact), the product of every matrix equal bit for
    bit to the FP32 kernel's on the same operands, four of them also to a numpy max-plus."""
    n, batch = 512, 4096
    _say(f"r0 isolated fp16 n={n} batch={batch}")
    tri = eng.tri_tiled_floats(n)
    bufs = [eng.tri_alloc(n, batch) for _ in range(4)]
    dA, dB, dC, dC2 = bufs
    half = [eng.tri_alloc(n, (batch + 1) // 2) for _ in range(2)]
    try:
        for d, seed in ((dA, 1), (dB, 2), (dC, 3), (dC2, 3)):
            eng.tri_fill_random(d, n, batch, seed, 0, 200)
        picks = [0, 4095]
        before = {m: [eng.tri_download(d + 4 * m * tri, n, 1)[0] for d in (dA, dB, dC)] for m in picks}
        assert eng.tri_to_half(dA, half[0], n, batch) and eng.tri_to_half(dB, half[1], n, batch)
        eng.r0_tri_batch_half(n, batch, half[0], half[1], dC)
        eng.r0_tri_batch(n, batch, dA, dB, dC2)
        for m0 in range(0, batch, 512):  # every matrix, 512 at a time
            a = eng.tri_download(dC + 4 * m0 * tri, n, 512)
            b = eng.tri_download(dC2 + 4 * m0 * tri, n, 512)
            assert _bits_equal(a, b), m0
            del a, b
        for m in picks:
            A, Bm, C = before[m]
            got = eng.tri_download(dC + 4 * m * tri, n, 1)[0]
            assert _bits_equal(got, _maxplus_tri_np(A, Bm, C)), m
    finally:
        for p in bufs + half:
            eng.tri_free(p)


def _tree6(M, N, N_tile):
    """The reference driver's 6-level pointer tree over one contiguous buffer
    (src/bpmax-wrapper.c:175-197), built with ctypes."""
    N_sec = (N + N_tile - 1) // N_tile
    R = N_sec * N_tile - N
    sec = (N_tile + 1) * N_tile
    lin = np.full(M * M * N_sec * N_sec * sec, np.nan, np.float32)
    base = lin.ctypes.data
    FP = ctypes.POINTER(ctypes.c_float)
    keep = []

    def arr(t, n):
        a = (t * n)()
        keep.append(a)
        return a

    L5 = ctypes.POINTER(FP)
    L4 = ctypes.POINTER(L5)
    L3 = ctypes.POINTER(L4)
    L2 = ctypes.POINTER(L3)
    top = arr(L2, M)
    for a in range(M):
        t1 = arr(L3, M)
        for b in range(M):
            t2 = arr(L4, N_sec)
            for c in range(N_sec):
                t3 = arr(L5, N_sec)
                for d in range(N_sec):
                    t4 = arr(FP, N_tile + 1)
                    off = (((a * M + b) * N_sec + c) * N_sec + d) * sec
                    for e in range(N_tile + 1):
                        t4[e] = ctypes.cast(base + 4 * (off + e * N_tile), FP)
                    t3[d] = ctypes.cast(t4, L5)
                t2[c] = ctypes.cast(t3, L4)
            t1[b] = ctypes.cast(t2, L3)
        top[a] = ctypes.cast(t1, L2)
    return lin, top, keep, N_sec, R


def test_compat_bpmax_symbol_writes_6d_tree(eng):
    """The V3-signature entry point fills the reference's 6-D layout."""
    fx = G.load_full(os.path.join(G.GOLDEN, "full_16x100_ref.npz"))
    M, N, N_tile = 16, 100, 48
    lin, top, keep, N_sec, R = _tree6(M, N, N_tile)
    L = B.load_library()
    fn = L.bpmax
    fn.restype = None
    s1 = np.ascontiguousarray(fx["seq1"], np.int32)
    s2 = np.ascontiguousarray(fx["seq2"], np.int32)
    _say("compat bpmax() 16x100 6-D tree")
    fn(ctypes.c_long(M), ctypes.c_long(N), ctypes.c_long(N_sec), ctypes.c_long(N_tile), ctypes.c_long(R),
       ctypes.c_long(3), ctypes.c_long(24), ctypes.c_long(8), B._ip(s1), B._ip(s2), top)
    sec = (N_tile + 1) * N_tile
    F = fx["F"]
    k = 0
    for i1 in range(M):
        for j1 in range(i1, M):
            for i2 in range(N):
                for j2 in range(i2, N):
                    ci, cj = i2 + R, j2 + R
                    off = (((i1 * M + j1) * N_sec + ci // N_tile) * N_sec + cj // N_tile) * sec
                    v = lin[off + (ci % N_tile) * N_tile + cj % N_tile]
                    assert v == F[k], (i1, j1, i2, j2)
                    k += 1


def test_graph_replays_match_oracle():
    """The pipelined schedule's graph is captured once per shape and replayed: every replay,
    with eager work and synchronisations between replays, and a shape change in the middle
    (graph rebuilt, then rebuilt back), gives the oracle's table bit for bit."""
    eng = B.Engine(0)
    eng.set_schedule(B.SCHED_PIPE)
    eng.set_graphs(True)
    try:
        _graph_replays(eng)
    finally:
        eng.close()


def _graph_replays(eng):
    rng = np.random.default_rng(11)
    shapes = [(5, 70), (5, 70), (5, 70), (3, 33), (5, 70), (5, 70)]
    tables = [O.table_asym(), (O.table_321(), None), (O.reference_table(), None)]
    for k, (M, N) in enumerate(shapes):
        s1 = rng.integers(0, 4, M).astype(np.int32)
        s2 = rng.integers(0, 4, N).astype(np.int32)
        # the weight table changes between replays of one shape
        intra, inter = tables[k % 3]
        _say(f"graph replay {k} M={M} N={N}")
        F, score = eng.solve(s1, s2, intra, inter)
        Fo, so = O.bpmax(s1, s2, intra, inter)
        assert score == so
        assert _bits_equal(F, Fo)
        # eager work on the same stream between replays
        eng.set_timing(True)
        eng.solve(s1[:2], s2[:9], intra, inter)
        eng.set_timing(False)
        eng.synchronize()


def _splitmix64(x):
    with np.errstate(over="ignore"):
        x = (x + np.uint64(0x9E3779B97F4A7C15))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def test_tri_fill_random_matches_host_hash(eng):
    """The synthetic R0 operands are the documented hash of (seed, matrix, row, column)."""
    n, batch, seed, lo, hi = 70, 3, 5, 0, 200
    _say(f"tri_fill_random n={n} batch={batch}")
    dev = eng.tri_alloc(n, batch)
    try:
        eng.tri_fill_random(dev, n, batch, seed, lo, hi)
        got = eng.tri_download(dev, n, batch)
    finally:
        eng.tri_free(dev)
    m, r, c = np.meshgrid(np.arange(batch, dtype=np.uint64), np.arange(n, dtype=np.uint64),
                          np.arange(n, dtype=np.uint64), indexing="ij")
    key = (m << np.uint64(42)) | (r << np.uint64(21)) | c
    want = (lo + (_splitmix64(np.uint64(seed) ^ key) % np.uint64(hi - lo + 1))).astype(np.float32)
    want[r > c] = -np.inf
    assert _bits_equal(got, want)


@pytest.mark.parametrize("name,ntile", [("full_5x17_ref.npz", 16), ("full_16x100_ref.npz", 48)])
def test_reference_driver_linked_to_gpu_library(name, ntile, eng):
    """The reference's own driver (bpmax-wrapper.c + external_functions.c, CHECKING build,
    unchanged) linked against libbpmax_hip.so instead of the reference engine: its dump of
    every F cell equals the dump of the reference-written fixture, line for line."""
    if not os.path.exists(O.REF_DROPIN_CHECK):
        pytest.skip("oracle/_ref/bpmax_dropin.check not built (make -C oracle dropin)")
    fx = G.load_full(os.path.join(G.GOLDEN, name))
    M, N = len(fx["seq1"]), len(fx["seq2"])
    _say(f"reference driver via the GPU library, {name}")
    stdin = " ".join(str(int(x)) for x in list(fx["seq1"]) + list(fx["seq2"])) + "\n"
    r = subprocess.run([O.REF_DROPIN_CHECK, str(M), str(N), str(ntile), "3", "24"], input=stdin,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-500:]
    got = [ln[ln.index("FTable("):] for ln in r.stdout.splitlines() if "FTable(" in ln]
    want = [f"FTable({i1},{j1},{i2},{j2})={v:0.2f}" for (i1, j1, i2, j2), v in
            zip(((i1, j1, i2, j2) for i1 in range(M) for j1 in range(i1, M)
                 for i2 in range(N) for j2 in range(i2, N)), fx["F"])]
    assert got == want


def test_autotuned_schedule_is_used_and_exact():
    """bpmax_tune_shape picks a schedule for the class; solves of that class then run with it
    and still match the oracle bit for bit; set/get round-trip the table."""
    eng = B.Engine(0)
    try:
        _say("autotune 6x70")
        best = eng.tune(6, 70, 1, reps=1)
        assert best["schedule"] in ("pipe", "launch") and best["outer_items"] >= 64 and best["ms"] > 0
        got = eng.get_tuning(6, 70, 1)
        assert got == {k: best[k] for k in got}
        intra, inter = O.table_asym()
        rng = np.random.default_rng(4)
        for M, N in [(5, 66), (8, 127)]:  # same class: M <= 8, 65 <= N <= 128
            s1 = rng.integers(0, 4, M).astype(np.int32)
            s2 = rng.integers(0, 4, N).astype(np.int32)
            F, score = eng.solve(s1, s2, intra, inter)
            Fo, so = O.bpmax(s1, s2, intra, inter)
            assert score == so and _bits_equal(F, Fo)
        for setting in [{"schedule": "launch", "outer_split": 1, "outer_items": 64, "graphs": False,
                         "intra_wide": 0, "outer_kgroup": 0, "closure_ahead": True, "tile": 0},
                        {"schedule": "pipe", "outer_split": 0, "outer_items": 4096, "graphs": True,
                         "intra_wide": 1 << 30, "outer_kgroup": 3, "closure_ahead": False, "tile": 32},
                        {"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "graphs": False,
                         "intra_wide": 256, "outer_kgroup": 1, "closure_ahead": True, "tile": 32}]:
            eng.set_tuning(6, 70, 1, setting)
            assert eng.get_tuning(6, 70, 1) == setting
            F, score = eng.solve(s1, s2, intra, inter)
            assert score == so and _bits_equal(F, Fo)
        eng.set_tuning(6, 70, 1, None)
        assert eng.get_tuning(6, 70, 1) is None
    finally:
        eng.close()

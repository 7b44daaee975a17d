"""GPU tests of the continuous-batching server: every result is bit-identical to the same
pair solved alone and to the reference's tables for the serving workload's fixture set,
pairs run side by side, admission respects the memory budget, stop() never blocks on
kept tables, and the error paths (too big, queue full, unknown id) report the documented
codes."""
import sys
import threading
import time

import numpy as np
import pytest

import bpmaxcpu_amd as B
from oracle import oracle as O
from tests import golden_io as G

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _requests(n, seed, mlo=3, mhi=24, nlo=20, nhi=200):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        M = int(rng.integers(mlo, mhi + 1))
        N = int(rng.integers(nlo, nhi + 1))
        out.append((rng.integers(0, 4, M).astype(np.int32), rng.integers(0, 4, N).astype(np.int32)))
    return out


def test_server_results_match_solo_solves():
    intra, inter = O.table_asym()
    reqs = _requests(30, seed=21)
    srv = B.Engine(0)
    solo = B.Engine(0)
    try:
        _say("server: 30 mixed requests, 3 slots, batches of <= 6, tables kept")
        srv.server_start(memory_budget=8 << 30, max_batch=6, slots=3, keep_tables=True, intra=intra, inter=inter)
        ids = []
        for k, (s1, s2) in enumerate(reqs):
            ids.append(srv.server_submit(s1, s2))
            if k % 5 == 4:
                time.sleep(0.002)  # arrivals while earlier batches run
        res = [srv.server_wait(i, 60000.0) for i in ids]
        st = srv.server_stats()
        assert st["completed"] == 30 and st["failed"] == 0
        assert st["batches"] >= 5 and st["max_batch_seen"] <= 6
        for k, ((s1, s2), r) in enumerate(zip(reqs, res)):
            assert r is not None
            F, score = solo.solve(s1, s2, intra, inter)
            assert r[0] == score, k
            assert r[1] >= 0.0
            if k % 7 == 0:
                Fs = srv.server_fetch(ids[k], len(s1), len(s2))
                assert _bits_equal(Fs, F), k
        s1, s2 = reqs[0]
        Fo, so = O.bpmax(s1, s2, intra, inter)
        assert res[0][0] == so
        assert _bits_equal(srv.server_fetch(ids[0], len(s1), len(s2)), Fo)
        for i in ids:
            srv.server_release(i)
        assert srv.server_stats()["bytes_in_use"] == 0
        with pytest.raises(B.BPMaxError) as e:
            srv.server_poll(ids[0])
        assert e.value.code == B.ENOTFOUND
    finally:
        srv.server_stop()
        srv.close()
        solo.close()


def test_server_overlaps_batches_and_respects_budget():
    """A budget of about three mid-size pairs: batches must run side by side, the arena
    in use never exceeds the budget, and every score still matches a solo solve."""
    reqs = _requests(16, seed=5, mlo=12, mhi=20, nlo=150, nhi=260)
    one = max(B.load_library().bpmax_engine_pair_device_bytes(len(a), len(b)) for a, b in reqs)
    budget = 3 * one + (1 << 20)
    srv = B.Engine(0)
    solo = B.Engine(0)
    try:
        _say(f"server: 16 requests, budget {budget >> 20} MiB, at most 2 resident, 3 ticks ahead")
        srv.server_start(memory_budget=budget, max_batch=2, slots=3)
        ids = [srv.server_submit(a, b) for a, b in reqs]
        res = [srv.server_wait(i, 60000.0) for i in ids]
        st = srv.server_stats()
        assert st["completed"] == 16
        assert st["peak_bytes_in_use"] <= st["budget"] == budget // 4096 * 4096
        assert st["max_inflight_pairs"] >= 2
        for (a, b), r in zip(reqs, res):
            assert r[0] == solo.solve(a, b, layout=B.LAYOUT_NONE)[1]
        assert st["bytes_in_use"] == 0  # tables not kept: freed with the scores
    finally:
        srv.server_stop()
        srv.close()
        solo.close()


def test_server_error_codes():
    srv = B.Engine(0)
    try:
        _say("server: error codes")
        srv.server_start(memory_budget=64 << 20, max_batch=1, slots=1, queue_capacity=4)
        big = np.zeros(64, np.int32), np.zeros(1024, np.int32)
        with pytest.raises(B.BPMaxError) as e:
            srv.server_submit(*big)
        assert e.value.code == B.ETOOBIG
        with pytest.raises(B.BPMaxError) as e:
            srv.server_submit(np.array([0, 5], np.int32), np.zeros(8, np.int32))
        assert e.value.code == B.EINVAL
        # more submissions than the queue holds while one slot works through them
        s1, s2 = np.zeros(16, np.int32), np.zeros(200, np.int32)
        codes = []
        ids = []
        for _ in range(40):
            try:
                ids.append(srv.server_submit(s1, s2))
            except B.BPMaxError as err:
                codes.append(err.code)
        assert codes and set(codes) == {B.EBUSY}
        for i in ids:
            assert srv.server_wait(i, 60000.0) is not None
        with pytest.raises(B.BPMaxError) as e:
            srv.server_poll(123456789)
        assert e.value.code == B.ENOTFOUND
    finally:
        srv.server_stop()
        srv.close()


def test_server_stop_does_not_block_on_kept_tables():
    """With keep_tables, an unreleased table can hold the arena so that a queued pair never
    fits: stop() must still return (the queued request fails with BPMAX_ENOMEM instead of
    waiting for a release that cannot come while stop blocks the caller)."""
    s1, s2 = np.zeros(12, np.int32), np.zeros(200, np.int32)
    one = B.load_library().bpmax_engine_pair_device_bytes(12, 200)
    srv = B.Engine(0)
    try:
        _say("server: a kept table blocks a second request; stop")
        srv.server_start(memory_budget=one + (one >> 1), max_batch=4, slots=2, keep_tables=True)
        a = srv.server_submit(s1, s2)
        assert srv.server_wait(a, 60000.0) is not None
        b = srv.server_submit(s1, s2)  # fits only once `a` is released
        time.sleep(0.05)
        assert srv.server_poll(b) is None
        st = srv.server_stats()
        assert st["completed"] == 1 and st["submitted"] == 2
        stopper = threading.Thread(target=srv.server_stop)
        stopper.start()
        stopper.join(30.0)
        assert not stopper.is_alive(), "server_stop blocked on a request that can never be admitted"
    finally:
        srv.close()


@pytest.mark.parametrize("name", G.set_fixtures("serve5"))
def test_server_serving_fixture_set_poisson_arrivals(name):
    """The serving workload's fixture set (M log-uniform [8,64], N log-uniform [64,1024],
    including 64x1024) submitted with Poisson-timed arrivals, tables kept: every score, the
    strided cells and the SHA-256 of every fetched F table equal the reference's."""
    fx = G.load_set(name)
    pairs = fx["pairs"]
    rng = np.random.default_rng(99)
    gaps = rng.exponential(0.02, len(pairs))
    srv = B.Engine(0)
    try:
        _say(f"server: {name}, {len(pairs)} requests, Poisson arrivals (mean gap 20 ms), tables kept")
        srv.server_start(keep_tables=True, intra=fx["intra"], inter=fx["inter"])
        ids = []
        for p, g in zip(pairs, gaps):
            time.sleep(float(g))
            ids.append(srv.server_submit(p["seq1"], p["seq2"]))
        for k, (p, i) in enumerate(zip(pairs, ids)):
            r = srv.server_wait(i, 120000.0)
            assert r is not None, k
            assert r[0] == p["score"], (k, p["M"], p["N"])
        st = srv.server_stats()
        assert st["completed"] == len(pairs) and st["failed"] == 0
        assert st["max_inflight_pairs"] >= 2
        for k, (p, i) in enumerate(zip(pairs, ids)):
            _say(f"server: fetch {k} {p['M']}x{p['N']}")
            F = srv.server_fetch(i, p["M"], p["N"])
            for i1, j1, i2, j2, v in p["samples"]:
                assert F[O.packed_index(i1, j1, i2, j2, p["M"], p["N"])] == v
            assert G.sha256_f32(F) == p["sha256"], (k, p["M"], p["N"])
            del F
            srv.server_release(i)
        assert srv.server_stats()["bytes_in_use"] == 0
    finally:
        srv.server_stop()
        srv.close()

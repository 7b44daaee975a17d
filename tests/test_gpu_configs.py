"""GPU parity on the batched and the serving workloads, pair by pair against fixture sets
written by the reference's own bpmax() (tests/golden/set_*.json, make_golden.py --sets):

* batch4: 16 pairs drawn like the batched workload (M~U[16,48], N~U[100,400], per-sequence
  GC~U[0.3,0.7]), solved in ONE bpmax_engine_solve_batch call; every pair's score, strided
  cells and the SHA-256 of its whole packed F table (fetched with bpmax_engine_fetch) must
  equal the reference's;
* serve5: 10 pairs drawn like the serving workload (M log-uniform [8,64], N log-uniform
  [64,1024]) including its corners 8x1024, 12x800, 48x700 and 64x1024, solved one per call,
  each checked the same way.  (The served form of this set is in test_gpu_server.py.)"""
import sys

import numpy as np
import pytest

import bpmaxcpu_amd as B
from oracle import oracle as O
from tests import golden_io as G

pytestmark = pytest.mark.gpu

BATCH4 = G.set_fixtures("batch4")
SERVE5 = G.set_fixtures("serve5")


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=["tuned-pipe", "launch", "tile32"])
def eng(request):
    e = B.Engine(0)
    if request.param == "tile32":
        e.set_default_tuning({"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "graphs": False,
                              "intra_wide": 256, "outer_kgroup": 1, "closure_ahead": True, "tile": 32})
    elif request.param == "launch":
        e.set_default_tuning({"schedule": "launch", "outer_split": 1, "outer_items": 2048, "graphs": False,
                              "intra_wide": 256, "outer_kgroup": 1, "closure_ahead": True})
        e.set_schedule("launch")
    else:
        e.set_graphs(True)
    yield e
    e.close()


def _check_pair(eng, k, p, F):
    M, N = p["M"], p["N"]
    for i1, j1, i2, j2, v in p["samples"]:
        assert F[O.packed_index(i1, j1, i2, j2, M, N)] == v, (k, i1, j1, i2, j2)
    assert G.sha256_f32(F) == p["sha256"], (k, M, N)


@pytest.mark.parametrize("name", BATCH4)
def test_batch4_set_in_one_batched_call(eng, name):
    fx = G.load_set(name)
    pairs = fx["pairs"]
    _say(f"{name}: {len(pairs)} pairs in one batch, shapes {[(p['M'], p['N']) for p in pairs]}")
    scores = eng.solve_batch([p["seq1"] for p in pairs], [p["seq2"] for p in pairs], fx["intra"], fx["inter"])
    assert [float(s) for s in scores] == [p["score"] for p in pairs]
    for k, p in enumerate(pairs):
        _check_pair(eng, k, p, eng.fetch(k, p["M"], p["N"]))


@pytest.mark.parametrize("name", SERVE5)
def test_serve5_set_one_call_per_pair(name):
    fx = G.load_set(name)
    e = B.Engine(0)
    try:
        for k, p in enumerate(fx["pairs"]):
            _say(f"{name}: pair {k} {p['M']}x{p['N']}")
            F, score = e.solve(p["seq1"], p["seq2"], fx["intra"], fx["inter"])
            assert score == p["score"], k
            _check_pair(e, k, p, F)
            del F
    finally:
        e.close()


def test_fetch_rejects_a_missing_result():
    e = B.Engine(0)
    try:
        with pytest.raises(B.BPMaxError) as err:
            e.fetch(0, 2, 9)
        assert err.value.code == B.EINVAL
        s = np.array([0, 1, 2], np.int32)
        e.solve_batch([s, s[:2]], [s[::-1], s])
        assert e.fetch(1, 2, 3).shape == (3 * 6,)
        with pytest.raises(B.BPMaxError):
            e.fetch(2, 2, 3)
    finally:
        e.close()

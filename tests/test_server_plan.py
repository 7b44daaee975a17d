"""Host-only checks of the serving loop's plan (bpmax_server_check, no GPU): requests of
mixed shapes arrive while others run; the server's own admission and tick planner produce
the stream operations, and a two-stream event model verifies that every work item starts
after everything it reads, every wait names an event recorded before it and not re-recorded
since, resident extents never overlap and every operation on a pair finishes before its
extent is released."""
import numpy as np
import pytest

import bpmaxcpu_amd as B

BASE = {"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "graphs": False, "intra_wide": 256,
        "outer_kgroup": 1, "closure_ahead": True}


def _requests(n, seed, mlo=1, mhi=20, nlo=1, nhi=300, span=30):
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.integers(mlo, mhi + 1)), int(rng.integers(nlo, nhi + 1))) for _ in range(n)]
    arrive = sorted(int(x) for x in rng.integers(0, span, n))
    return shapes, arrive


@pytest.mark.parametrize("tuning", [
    BASE,
    dict(BASE, closure_ahead=False, outer_kgroup=0),
    dict(BASE, outer_split=0, outer_items=64, outer_kgroup=2),
    dict(BASE, outer_split=1, outer_items=4096),
])
@pytest.mark.parametrize("depth", [1, 2, 3])
def test_server_plan_orders_every_dependency(tuning, depth):
    shapes, arrive = _requests(14, seed=depth)
    rc, n, msg = B.server_check(shapes, arrive, tuning, 1 << 30, max_active=6, depth=depth)
    assert rc == B.OK, msg
    assert n > 10000


def test_server_plan_under_memory_pressure_reuses_extents_safely():
    """A budget of about two mid-size pairs and a queue of many: extents are released and
    reused many times; the resident-pair limit and FIFO admission hold."""
    shapes, arrive = _requests(24, seed=7, mlo=6, mhi=14, nlo=100, nhi=200, span=5)
    one = max(B.load_library().bpmax_engine_pair_device_bytes(m, n) for m, n in shapes)
    for depth in (1, 3):
        rc, n, msg = B.server_check(shapes, arrive, BASE, 2 * one + (2 << 20), max_active=4, depth=depth)
        assert rc == B.OK, msg


def test_server_plan_serving_mix():
    """The serving workload's shapes (M log-uniform [8,64], N log-uniform [64,1024]; small M
    keeps the host check short), both tile sizes in flight at once."""
    rng = np.random.default_rng(13)
    shapes = [(int(round(np.exp(rng.uniform(np.log(4), np.log(12))))),
               int(round(np.exp(rng.uniform(np.log(20), np.log(1024)))))) for _ in range(16)]
    arrive = sorted(int(x) for x in rng.integers(0, 20, 16))
    assert any(n <= 64 for _, n in shapes) and any(n > 64 for _, n in shapes)
    rc, n, msg = B.server_check(shapes, arrive, BASE, 64 << 30, max_active=256, depth=3)
    assert rc == B.OK, msg


def test_server_plan_check_catches_every_dropped_wait():
    """Planted error: each single event wait of the plan, dropped, must break a dependency the
    check reports (every wait the planner issues is needed)."""
    shapes, arrive = [(7, 53), (1, 60), (4, 163), (4, 19), (3, 120), (7, 145)], [0, 1, 2, 5, 8, 9]
    rc, _, msg = B.server_check(shapes, arrive, BASE, 64 << 20, max_active=3, depth=2)
    assert rc == B.OK, msg
    k = 0
    while True:
        rc, _, msg = B.server_check(shapes, arrive, BASE, 64 << 20, max_active=3, depth=2, skip=k)
        if rc == B.OK:
            break  # k is past the last wait
        assert msg
        k += 1
    assert k == 55  # the number of waits this plan issues: every one of them was caught


def test_server_plan_rejects_pairs_that_never_fit():
    rc, _, msg = B.server_check([(64, 1024)], [0], BASE, 64 << 20)
    assert rc != B.OK and "budget" in msg

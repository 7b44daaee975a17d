"""The no-allocation solve (bpmax_engine_solve_noalloc, the library's fixed per-GPU scratch)
against the allocating bpmax_engine_solve, bit for bit: two calls in a row, the second on a
smaller pair (the scratch keeps the first call's contents, which must not leak into the
second), both layouts; and a pair too large for the scratch is refused before any work."""
import sys

import numpy as np
import pytest

import bpmaxcpu_amd as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_solve_matches_solve_on_two_calls_in_a_row():
    rng = np.random.default_rng(31)
    big = rng.integers(0, 4, 24).astype(np.int32), rng.integers(0, 4, 160).astype(np.int32)
    small = rng.integers(0, 4, 9).astype(np.int32), rng.integers(0, 4, 70).astype(np.int32)
    intra, inter = O.table_asym()
    eng = B.Engine(0)
    try:
        _say("noalloc: 24x160 then 9x70, packed")
        Fb, sb = eng.solve_noalloc(*big, intra, inter)
        Fs, ss = eng.solve_noalloc(*small, intra, inter)
        _say("noalloc: the same pairs through bpmax_engine_solve")
        Fb_ref, sb_ref = eng.solve(*big, intra, inter)
        Fs_ref, ss_ref = eng.solve(*small, intra, inter)
        assert sb == sb_ref and _bits_equal(Fb, Fb_ref)
        assert ss == ss_ref and _bits_equal(Fs, Fs_ref)
        _say("noalloc: score-only, 24x160 then 9x70")
        assert eng.solve_noalloc(*big, intra, inter, layout=B.LAYOUT_NONE)[1] == sb_ref
        assert eng.solve_noalloc(*small, intra, inter, layout=B.LAYOUT_NONE)[1] == ss_ref
        # the allocating path is unaffected by a no-allocation call in between
        assert _bits_equal(eng.solve(*small, intra, inter)[0], Fs_ref)
        Fo, so = O.bpmax(*small, intra, inter)
        assert ss == so and _bits_equal(Fs, Fo)
    finally:
        eng.close()


def test_noalloc_refuses_a_pair_larger_than_its_scratch():
    eng = B.Engine(0)
    try:
        _say("noalloc: 64x512 is refused")
        s1, s2 = np.zeros(64, np.int32), np.zeros(512, np.int32)
        for layout in (B.LAYOUT_PACKED, B.LAYOUT_NONE):
            with pytest.raises(B.BPMaxError) as e:
                eng.solve_noalloc(s1, s2, layout=layout)
            assert e.value.code == B.ENOMEM
        # still usable
        assert eng.solve_noalloc(np.array([0], np.int32), np.array([3], np.int32), layout=B.LAYOUT_NONE)[1] == 1.0
    finally:
        eng.close()

"""C-ABI checks that need no GPU: the library loads, exports every function include/bpmax.h
declares, and its host-only helpers (closed forms, sequence reader, CHECKING dump) match
the reference's conventions."""
import os
import subprocess

import numpy as np
import pytest

import bpmaxcpu_amd as B
from oracle import oracle as O
from tests import golden_io as G


def test_library_exports_every_header_function():
    L = B.load_library()
    names = B.header_functions()
    assert "bpmax" in names and "bpmax_engine_solve" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_cell_update_closed_form():
    # values quoted with the workload descriptions: 64x512 -> 1.0816e12, 16x100 -> 1.6552e8
    assert abs(B.cell_updates(64, 512) / 1.0816e12 - 1) < 1e-4
    assert abs(B.cell_updates(16, 100) / 1.6552e8 - 1) < 1e-4
    # brute-force count of the recurrence's terms on a tiny shape
    M, N = 5, 7
    n = 0
    for i1 in range(M):
        for j1 in range(i1, M):
            for i2 in range(N):
                for j2 in range(i2, N):
                    n += (j1 - i1) * (j2 - i2) + 2 * (j2 - i2) + 2 * (j1 - i1)
    assert B.cell_updates(M, N) == n


def test_layout_bytes():
    L = B.load_library()
    assert L.bpmax_layout_bytes(64, 512, B.LAYOUT_PACKED) == 2080 * 131328 * 4
    assert L.bpmax_layout_bytes(3, 5, B.LAYOUT_DENSE_4D) == 3 * 3 * 5 * 5 * 4
    assert L.bpmax_layout_bytes(0, 5, B.LAYOUT_PACKED) == 0


def test_base_codes_and_default_table():
    L = B.load_library()
    assert [L.bpmax_base_code(c.encode()) for c in "ACGUTNacgut0123X"] == [0, 1, 2, 3, 3, 4, 0, 1, 2, 3, 3, 0, 1, 2, 3, 4]
    assert np.array_equal(B.reference_table(), O.reference_table())


def test_read_sequence_file(tmp_path):
    L = B.load_library()
    p = tmp_path / "s.fa"
    p.write_text(">name one\nACGU\nacgt nX\n>second\nGGGG\n")
    out = np.zeros(16, np.int32)
    n = L.bpmax_read_sequence_file(str(p).encode(), B._ip(out), 16)
    assert n == 10
    assert out[:10].tolist() == [0, 1, 2, 3, 0, 1, 2, 3, 4, 4]
    p.write_text("AC;GU\n")
    assert L.bpmax_read_sequence_file(str(p).encode(), B._ip(out), 16) == 2
    assert L.bpmax_read_sequence_file(b"/nonexistent/file", B._ip(out), 16) == B.EINVAL


def test_checking_dump_matches_reference_format(tmp_path):
    """The dump writer reproduces the reference's CHECKING output byte-for-byte
    (src/bpmax-wrapper.c:318-342), checked against oracle/_ref/bpmax.check when built."""
    L = B.load_library()
    fx = G.load_full(os.path.join(G.GOLDEN, "full_5x17_ref.npz"))
    M, N = len(fx["seq1"]), len(fx["seq2"])
    out = tmp_path / "dump.txt"
    F = np.ascontiguousarray(fx["F"], np.float32)
    assert L.bpmax_write_checking_dump(str(out).encode(), B._fp(F), M, N, 1) == 0
    lines = out.read_text().splitlines()
    assert len(lines) == O.tri(M) * O.tri(N)
    assert lines[0].startswith("FTable(0,0,0,0)=")
    if not os.path.exists(O.REF_CHECK):
        pytest.skip("oracle/_ref/bpmax.check not built")
    stdin = " ".join(str(int(x)) for x in list(fx["seq1"]) + list(fx["seq2"])) + "\n"
    r = subprocess.run([O.REF_CHECK, str(M), str(N), "16", "2", "8"], input=stdin, capture_output=True,
                       text=True, check=True)
    ref_lines = [ln[ln.index("FTable("):] for ln in r.stdout.splitlines() if "FTable(" in ln]
    assert ref_lines == lines


def test_context_without_gpu_fails_cleanly():
    import ctypes
    L = B.load_library()
    h = ctypes.c_void_p()
    rc = L.bpmax_ctx_create(ctypes.byref(h), 0)
    if rc == B.OK:
        L.bpmax_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == B.EDEVICE
    assert L.bpmax_ctx_create(None, 0) == B.EINVAL


def test_server_entry_points_reject_null_context():
    import ctypes
    L = B.load_library()
    rid = ctypes.c_longlong()
    s = np.zeros(4, np.int32)
    assert L.bpmax_server_start(None, None) == B.EINVAL
    assert L.bpmax_server_submit(None, B._ip(s), 4, B._ip(s), 4, ctypes.byref(rid)) == B.EINVAL
    assert L.bpmax_server_poll(None, 1, None, None) == B.EINVAL
    assert L.bpmax_server_release(None, 1) == B.EINVAL
    assert L.bpmax_server_stop(None) == B.EINVAL


def test_tuning_entry_points_reject_bad_arguments():
    L = B.load_library()
    t = B.Tuning(B.SCHED_PIPE, 2, 2048, 0)
    assert L.bpmax_ctx_set_tuning(None, 8, 8, 1, ctypes_ref(t)) == B.EINVAL
    assert L.bpmax_ctx_get_tuning(None, 8, 8, 1, ctypes_ref(t)) == B.EINVAL
    assert L.bpmax_tune_shape(None, 8, 8, 1, 1, None, None) == B.EINVAL


def test_autotune_shape_parser():
    from bpmaxcpu_amd.autotune import parse_shape
    assert parse_shape("64x512") == (64, 512, 1)
    assert parse_shape("48x400x16") == (48, 400, 16)
    with pytest.raises(ValueError):
        parse_shape("64")


def ctypes_ref(x):
    import ctypes
    return ctypes.byref(x)



@pytest.mark.parametrize("tuning", [
    {"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "outer_kgroup": 1},
    {"schedule": "pipe", "outer_split": 1, "outer_items": 512, "outer_kgroup": 0},
    {"schedule": "pipe", "outer_split": 0, "outer_items": 4096, "outer_kgroup": 2},
    {"schedule": "pipe", "outer_split": 4, "outer_items": 64, "outer_kgroup": 3},
    {"schedule": "launch", "outer_split": 1, "outer_items": 2048, "outer_kgroup": 1},
    {"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "outer_kgroup": 1, "closure_ahead": True},
    {"schedule": "pipe", "outer_split": 0, "outer_items": 64, "outer_kgroup": 0, "closure_ahead": True},
    {"schedule": "launch", "outer_split": 1, "outer_items": 2048, "outer_kgroup": 2, "closure_ahead": True},
    {"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "outer_kgroup": 1, "closure_ahead": True, "tile": 32},
    {"schedule": "pipe", "outer_split": 0, "outer_items": 512, "outer_kgroup": 2, "closure_ahead": False, "tile": 32},
])
def test_schedule_runs_every_item_after_what_it_reads(tuning):
    """The launch list of a 40x300 pair and of a mixed batch (two tile sizes, ragged M and N)
    is checked on the host: every R0 item after the closures it reads, every in-tile closure
    after its R0 tile, the tiles left of / below it and its cross-tile part, every cross-tile
    step after the tile it reads and each of them run once (also when the in-tile launches
    carry them, closure_ahead), R0 of the first diagonal after S2 -- under the pipelined
    runner's stream order and event waits."""
    for shapes in ([(40, 300)], [(12, 200), (7, 70), (20, 260), (3, 40), (1, 9), (5, 64)]):
        rc, n, msg = B.schedule_check(shapes, dict(tuning, graphs=False, intra_wide=256))
        assert rc == B.OK, msg
        assert n > 1000


@pytest.mark.parametrize("ahead", [False, True])
def test_schedule_check_catches_a_missing_wait(ahead):
    """Planted error: with the wait of one launch dropped, a dependency is no longer ordered
    and the check must say so -- for every launch whose wait carries one."""
    tuning = {"schedule": "pipe", "outer_split": 2, "outer_items": 2048, "outer_kgroup": 1, "graphs": False,
              "intra_wide": 256, "closure_ahead": ahead}
    caught = 0
    for k in range(0, 700):
        rc, _, msg = B.schedule_check([(10, 200)], tuning, skip=k)
        if rc != B.OK:
            assert msg
            caught += 1
    assert caught >= 20


def test_bench_tree6_addresses_every_level():
    """The reference's 6-level F pointer tree that the bench builds with numpy address
    arithmetic (src/bpmax-wrapper.c:175-197) points every leaf at the reference's offset."""
    import ctypes
    import bench
    M, N, Nt = 3, 20, 8
    lin, top, keep = bench.tree6(M, N, Nt)
    Ns, sec = (N + Nt - 1) // Nt, (Nt + 1) * Nt
    U = ctypes.POINTER(ctypes.c_uint64)
    l1 = ctypes.cast(top.ctypes.data, U)
    for a in range(M):
        l2 = ctypes.cast(int(l1[a]), U)
        for b in range(M):
            l3 = ctypes.cast(int(l2[b]), U)
            for c in range(Ns):
                l4 = ctypes.cast(int(l3[c]), U)
                for d in range(Ns):
                    l5 = ctypes.cast(int(l4[d]), U)
                    for e in range(Nt + 1):
                        assert int(l5[e]) == lin.ctypes.data + 4 * ((((a * M + b) * Ns + c) * Ns + d) * sec + e * Nt)


def test_m0_is_used_only_by_the_lds_stage_loads(tmp_path):
    """The stage loads set M0 in inline assembly, which the compiler cannot be told about (M0
    is reserved): compiled to gfx950 assembly, the kernels must touch M0 nowhere else, and
    every write of it must be followed by the LDS load that reads it."""
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    csrc = os.path.join(os.path.dirname(B.__file__), "csrc")
    out = tmp_path / "kernels.s"
    subprocess.run([hipcc, "-O3", "-std=c++17", "-I", os.path.join(csrc, "..", "..", "include"), "-I", csrc,
                    "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-fno-honor-nans", "--cuda-device-only", "-S", os.path.join(csrc, "bpmax_kernels.hip"),
                    "-o", str(out)], check=True, capture_output=True)
    code = [ln.split(";")[0].strip() for ln in out.read_text().splitlines()]
    code = [ln for ln in code if ln and not ln.startswith(".") and not ln.endswith(":")]
    writes = 0
    for i, ln in enumerate(code):
        if "m0" not in ln.replace(",", " ").split():
            continue
        assert ln.startswith("s_mov_b32 m0,"), ln
        assert code[i + 1].startswith("global_load_lds_dword"), code[i + 1]
        writes += 1
    assert writes >= 4

"""bpmaxcpu_amd -- MI355X-native BPMax engine (Python binding of libbpmax_hip.so).

The engine is C/HIP behind the C ABI declared in ``include/bpmax.h``; this module is a
thin ctypes layer over it for tests, benchmarks and scripting.  It mirrors the
reference's entry point ``bpmax(M, N, ..., seq1, seq2, FTable)``
(TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/src/bpmax.c:129) as
``Engine.solve(seq1, seq2, ...)`` returning the F table and the optimal score.

There is no CPU fallback: if the HIP library is not built, or no GPU is present,
constructing an :class:`Engine` raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BPMAX_LIB selects another build of the library (A/B comparisons of kernel variants)
LIB_PATH = os.environ.get("BPMAX_LIB") or os.path.join(HERE, "libbpmax_hip.so")
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "bpmax.h")

OK, EINVAL, ENOMEM, EDEVICE, EWEIGHTS, EBUSY, ENOTREADY, ENOTFOUND, ETOOBIG = 0, -1, -2, -3, -4, -5, -6, -7, -8
LAYOUT_NONE, LAYOUT_PACKED, LAYOUT_DENSE_4D = 0, 1, 2
K_FILL, K_S1, K_OUTER, K_INNER_EXT, K_INNER_INTRA, K_CONVERT, K_R0TRI, K_TASKS = range(8)
KERNEL_CLASSES = ["fill", "s1", "outer", "inner_ext", "inner_intra", "convert", "r0tri", "tasks"]
SCHED_TASKS, SCHED_LAUNCH, SCHED_PIPE = 0, 1, 2
SCHEDULES = {"tasks": SCHED_TASKS, "launch": SCHED_LAUNCH, "pipe": SCHED_PIPE}

_ERRNAMES = {EINVAL: "EINVAL", ENOMEM: "ENOMEM", EDEVICE: "EDEVICE", EWEIGHTS: "EWEIGHTS",
             EBUSY: "EBUSY", ENOTREADY: "ENOTREADY", ENOTFOUND: "ENOTFOUND", ETOOBIG: "ETOOBIG"}


class BPMaxError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libbpmax_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `make -C bpmaxcpu_amd/csrc`")
    L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    c_int, c_float, c_void_p, c_size_t = ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t
    ip, fp = ctypes.POINTER(c_int), ctypes.POINTER(c_float)
    ipp = ctypes.POINTER(ip)
    sig = {
        "bpmax_ctx_create": (c_int, [ctypes.POINTER(c_void_p), c_int]),
        "bpmax_ctx_destroy": (c_int, [c_void_p]),
        "bpmax_last_error": (ctypes.c_char_p, [c_void_p]),
        "bpmax_engine_score_table": (None, [fp, fp]),
        "bpmax_engine_solve": (c_int, [c_void_p, ip, c_int, ip, c_int, fp, fp, fp, c_int, fp]),
        "bpmax_engine_solve_batch": (c_int, [c_void_p, c_int, ipp, ip, ipp, ip, fp, fp, fp]),
        "bpmax_engine_fetch": (c_int, [c_void_p, c_int, c_int, fp]),
        "bpmax_engine_solve_noalloc": (c_int, [c_void_p, ip, c_int, ip, c_int, fp, fp, fp, c_int, fp]),
        "bpmax_layout_bytes": (c_size_t, [c_int, c_int, c_int]),
        "bpmax_engine_pair_device_bytes": (c_size_t, [c_int, c_int]),
        "bpmax_tri_tiled_floats": (c_size_t, [c_int]),
        "bpmax_r0_tri_batch_device": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
        "bpmax_r0_tri_batch_device_half": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
        "bpmax_tri_to_half": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, ctypes.POINTER(c_int)]),
        "bpmax_tri_upload": (c_int, [c_void_p, c_int, c_int, fp, c_void_p]),
        "bpmax_tri_download": (c_int, [c_void_p, c_int, c_int, c_void_p, fp]),
        "bpmax_tri_fill_random": (c_int, [c_void_p, c_int, c_int, ctypes.c_ulonglong, c_int, c_int, c_void_p]),
        "bpmax_device_alloc": (c_void_p, [c_void_p, c_size_t]),
        "bpmax_device_free": (c_int, [c_void_p, c_void_p]),
        "bpmax_ctx_set_r0_half": (c_int, [c_void_p, c_int]),
        "bpmax_ctx_r0_half_used": (c_int, [c_void_p]),
        "bpmax_ctx_set_timing": (c_int, [c_void_p, c_int]),
        "bpmax_ctx_set_timing_mask": (c_int, [c_void_p, ctypes.c_uint]),
        "bpmax_ctx_kernel_stats": (c_int, [c_void_p, c_int, ctypes.POINTER(ctypes.c_longlong),
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "bpmax_ctx_reset_stats": (c_int, [c_void_p]),
        "bpmax_ctx_task_stats": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)]),
        "bpmax_ctx_set_schedule": (c_int, [c_void_p, c_int, c_int]),
        "bpmax_ctx_set_graphs": (c_int, [c_void_p, c_int]),
        "bpmax_ctx_stream": (c_void_p, [c_void_p]),
        "bpmax_ctx_synchronize": (c_int, [c_void_p]),
        "bpmax_cell_updates": (ctypes.c_double, [c_int, c_int]),
        "bpmax_score_bound": (ctypes.c_double, [c_void_p, c_void_p, ip, c_int, ip, c_int]),
        "bpmax_base_code": (ctypes.c_ubyte, [ctypes.c_char]),
        "bpmax_read_sequence_file": (c_int, [ctypes.c_char_p, ip, c_int]),
        "bpmax_write_checking_dump": (c_int, [ctypes.c_char_p, fp, c_int, c_int, c_int]),
        "bpmax_tune_shape": (c_int, [c_void_p, c_int, c_int, c_int, c_int, ctypes.POINTER(Tuning),
                                     ctypes.POINTER(ctypes.c_double)]),
        "bpmax_tune_batch": (c_int, [c_void_p, c_int, ipp, ip, ipp, ip, c_int, ctypes.POINTER(Tuning),
                                     ctypes.POINTER(ctypes.c_double)]),
        "bpmax_ctx_set_tuning": (c_int, [c_void_p, c_int, c_int, c_int, ctypes.POINTER(Tuning)]),
        "bpmax_ctx_get_tuning": (c_int, [c_void_p, c_int, c_int, c_int, ctypes.POINTER(Tuning)]),
        "bpmax_ctx_set_default_tuning": (c_int, [c_void_p, ctypes.POINTER(Tuning)]),
        "bpmax_server_start": (c_int, [c_void_p, ctypes.POINTER(ServerConfig)]),
        "bpmax_server_submit": (c_int, [c_void_p, ip, c_int, ip, c_int, ctypes.POINTER(ctypes.c_longlong)]),
        "bpmax_server_poll": (c_int, [c_void_p, ctypes.c_longlong, fp, ctypes.POINTER(ctypes.c_double)]),
        "bpmax_server_wait": (c_int, [c_void_p, ctypes.c_longlong, ctypes.c_double, fp,
                                      ctypes.POINTER(ctypes.c_double)]),
        "bpmax_server_fetch": (c_int, [c_void_p, ctypes.c_longlong, c_int, fp]),
        "bpmax_server_release": (c_int, [c_void_p, ctypes.c_longlong]),
        "bpmax_server_get_stats": (c_int, [c_void_p, ctypes.POINTER(ServerStats)]),
        "bpmax_server_stop": (c_int, [c_void_p]),
        "bpmax_schedule_check": (c_int, [c_int, ip, ip, ctypes.POINTER(Tuning), c_int,
                                         ctypes.POINTER(ctypes.c_longlong), ctypes.c_char_p, c_int]),
        "bpmax_server_check": (c_int, [c_int, ip, ip, ip, c_size_t, c_int, c_int, ctypes.POINTER(Tuning), c_int,
                                       ctypes.POINTER(ctypes.c_longlong), ctypes.c_char_p, c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class Tuning(ctypes.Structure):
    """bpmax_tuning: launch-schedule settings of one problem class."""
    _fields_ = [("schedule", ctypes.c_int), ("outer_split", ctypes.c_int), ("outer_items", ctypes.c_int),
                ("graphs", ctypes.c_int), ("intra_wide", ctypes.c_int), ("outer_kgroup", ctypes.c_int),
                ("closure_ahead", ctypes.c_int), ("tile", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {"schedule": {v: k for k, v in SCHEDULES.items()}[self.schedule], "outer_split": self.outer_split,
                "outer_items": self.outer_items, "graphs": bool(self.graphs), "intra_wide": self.intra_wide,
                "outer_kgroup": self.outer_kgroup, "closure_ahead": bool(self.closure_ahead), "tile": self.tile}

    @staticmethod
    def from_dict(tuning: dict) -> "Tuning":
        return Tuning(SCHEDULES[tuning["schedule"]], tuning["outer_split"], tuning["outer_items"],
                      1 if tuning.get("graphs") else 0, tuning.get("intra_wide", 256), tuning.get("outer_kgroup", 1),
                      1 if tuning.get("closure_ahead", True) else 0, tuning.get("tile", 0))


def schedule_check(shapes, tuning: dict, skip: int = -1):
    """Host-only check of the launch schedule pairs of these (M, N) shapes get under
    `tuning` (bpmax_schedule_check): returns (rc, dependencies checked, message)."""
    L = load_library()
    Ms = np.array([m for m, _ in shapes], np.int32)
    Ns = np.array([n for _, n in shapes], np.int32)
    t = Tuning.from_dict(tuning)
    nchk = ctypes.c_longlong(0)
    msg = ctypes.create_string_buffer(256)
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.bpmax_schedule_check(len(shapes), Ms.ctypes.data_as(P), Ns.ctypes.data_as(P), ctypes.byref(t), skip,
                                ctypes.byref(nchk), msg, len(msg))
    return rc, nchk.value, msg.value.decode()


def server_check(shapes, arrive, tuning: dict, memory_budget: int, max_active: int = 256, depth: int = 3,
                 skip: int = -1):
    """Host-only check of the serving loop's plan for requests of these (M, N) shapes arriving
    at these tick counts (bpmax_server_check): returns (rc, dependencies checked, message)."""
    L = load_library()
    Ms = np.array([m for m, _ in shapes], np.int32)
    Ns = np.array([n for _, n in shapes], np.int32)
    At = np.array(arrive, np.int32)
    t = Tuning.from_dict(tuning)
    nchk = ctypes.c_longlong(0)
    msg = ctypes.create_string_buffer(256)
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.bpmax_server_check(len(shapes), Ms.ctypes.data_as(P), Ns.ctypes.data_as(P), At.ctypes.data_as(P),
                              memory_budget, max_active, depth, ctypes.byref(t), skip, ctypes.byref(nchk), msg,
                              len(msg))
    return rc, nchk.value, msg.value.decode()


class ServerConfig(ctypes.Structure):
    """bpmax_server_config."""
    _fields_ = [("memory_budget", ctypes.c_size_t), ("max_batch", ctypes.c_int), ("slots", ctypes.c_int),
                ("queue_capacity", ctypes.c_int), ("keep_tables", ctypes.c_int),
                ("intra", ctypes.POINTER(ctypes.c_float)), ("inter", ctypes.POINTER(ctypes.c_float))]


class ServerStats(ctypes.Structure):
    """bpmax_server_stats."""
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong),
                ("batches", ctypes.c_longlong), ("max_batch_seen", ctypes.c_longlong),
                ("max_inflight_pairs", ctypes.c_longlong), ("budget", ctypes.c_size_t),
                ("bytes_in_use", ctypes.c_size_t), ("peak_bytes_in_use", ctypes.c_size_t)]


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def reference_table() -> np.ndarray:
    """The reference's default 5x5 weight table (GC=3, AU=1, GU=1.5), from the library."""
    L = load_library()
    intra = np.zeros(25, np.float32)
    inter = np.zeros(25, np.float32)
    L.bpmax_engine_score_table(_fp(intra), _fp(inter))
    return intra.reshape(5, 5)


def encode(seq: str) -> np.ndarray:
    """ACGU/T/N letters -> codes 0..4 (src/external_functions.c:74-85)."""
    table = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3, "0": 0, "1": 1, "2": 2, "3": 3}
    return np.array([table.get(ch.upper(), 4) for ch in seq], np.int32)


def cell_updates(M: int, N: int) -> float:
    """Cell updates of the full recurrence: C(M+1,3)C(N+1,3)+2C(M+1,2)C(N+1,3)+2C(M+1,3)C(N+1,2)."""
    return float(load_library().bpmax_cell_updates(M, N))


def score_bound(seq1, seq2, intra=None, inter=None) -> float:
    """Upper bound on every score of the pair (bpmax_score_bound; host only).  R0 runs on FP16
    operands when it times 2^p is at most 2^11 (weights multiples of 2^-p)."""
    s1 = np.ascontiguousarray(seq1, np.int32)
    s2 = np.ascontiguousarray(seq2, np.int32)
    ti = None if intra is None else np.ascontiguousarray(intra, np.float32).reshape(25)
    te = None if inter is None else np.ascontiguousarray(inter, np.float32).reshape(25)
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    ip = ctypes.POINTER(ctypes.c_int)
    b = load_library().bpmax_score_bound(ptr(ti), ptr(te), s1.ctypes.data_as(ip), len(s1), s2.ctypes.data_as(ip),
                                         len(s2))
    if b < 0:
        raise BPMaxError(EINVAL, "score_bound: empty or out-of-range sequence, or a negative or non-finite weight")
    return float(b)


def _table(t, default):
    if t is None:
        return None if default is None else np.ascontiguousarray(default, np.float32).reshape(25)
    return np.ascontiguousarray(t, np.float32).reshape(25)


class Engine:
    """One GPU context (HIP stream + device arena).  Not thread-safe."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.bpmax_ctx_create(ctypes.byref(h), device)
        if rc != OK:
            raise BPMaxError(rc, f"cannot create a context on HIP device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.bpmax_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise BPMaxError(rc, self.L.bpmax_last_error(self.h).decode(errors="replace"))

    # -------------------------------------------------------------- solve
    def solve(self, seq1, seq2, intra=None, inter=None, layout: int = LAYOUT_PACKED):
        """Returns (F, score).  F is the packed tri-tri table (1-D float32), the dense
        [M][M][N][N] table, or None for LAYOUT_NONE."""
        s1 = np.ascontiguousarray(seq1, np.int32)
        s2 = np.ascontiguousarray(seq2, np.int32)
        M, N = len(s1), len(s2)
        ti = _table(intra, None)
        te = _table(inter, ti)
        out = None
        if layout == LAYOUT_PACKED:
            out = np.empty((M * (M + 1) // 2) * (N * (N + 1) // 2), np.float32)
        elif layout == LAYOUT_DENSE_4D:
            out = np.empty((M, M, N, N), np.float32)
        score = ctypes.c_float()
        rc = self.L.bpmax_engine_solve(self.h, _ip(s1), M, _ip(s2), N,
                                       None if ti is None else _fp(ti), None if te is None else _fp(te),
                                       None if out is None else _fp(out), layout, ctypes.byref(score))
        self._check(rc)
        return out, float(score.value)

    def solve_noalloc(self, seq1, seq2, intra=None, inter=None, layout: int = LAYOUT_PACKED):
        """As solve() (LAYOUT_PACKED or LAYOUT_NONE) without any device allocation, in the
        library's fixed per-GPU scratch (bpmax_engine_solve_noalloc)."""
        s1 = np.ascontiguousarray(seq1, np.int32)
        s2 = np.ascontiguousarray(seq2, np.int32)
        M, N = len(s1), len(s2)
        ti = _table(intra, None)
        te = _table(inter, ti)
        out = np.empty((M * (M + 1) // 2) * (N * (N + 1) // 2), np.float32) if layout == LAYOUT_PACKED else None
        score = ctypes.c_float()
        rc = self.L.bpmax_engine_solve_noalloc(self.h, _ip(s1), M, _ip(s2), N,
                                               None if ti is None else _fp(ti), None if te is None else _fp(te),
                                               None if out is None else _fp(out), layout, ctypes.byref(score))
        self._check(rc)
        return out, float(score.value)

    def solve_batch(self, seqs1: Sequence, seqs2: Sequence, intra=None, inter=None) -> np.ndarray:
        """Scores of many independent pairs, scheduled together on the device."""
        n = len(seqs1)
        assert n == len(seqs2) and n > 0
        a1 = [np.ascontiguousarray(s, np.int32) for s in seqs1]
        a2 = [np.ascontiguousarray(s, np.int32) for s in seqs2]
        P = ctypes.POINTER(ctypes.c_int)
        p1 = (P * n)(*[_ip(a) for a in a1])
        p2 = (P * n)(*[_ip(a) for a in a2])
        Ms = np.array([len(a) for a in a1], np.int32)
        Ns = np.array([len(a) for a in a2], np.int32)
        ti = _table(intra, None)
        te = _table(inter, ti)
        scores = np.empty(n, np.float32)
        rc = self.L.bpmax_engine_solve_batch(self.h, n, p1, _ip(Ms), p2, _ip(Ns),
                                             None if ti is None else _fp(ti),
                                             None if te is None else _fp(te), _fp(scores))
        self._check(rc)
        return scores

    def fetch(self, pair: int, M: int, N: int, layout: int = LAYOUT_PACKED) -> np.ndarray:
        """F of pair `pair` of the last solve / solve_batch (bpmax_engine_fetch)."""
        if layout == LAYOUT_PACKED:
            out = np.empty((M * (M + 1) // 2) * (N * (N + 1) // 2), np.float32)
        else:
            out = np.empty((M, M, N, N), np.float32)
        self._check(self.L.bpmax_engine_fetch(self.h, pair, layout, _fp(out)))
        return out

    # ------------------------------------------------------- isolated R0 kernel
    def tri_tiled_floats(self, n: int) -> int:
        return self.L.bpmax_tri_tiled_floats(n)

    def tri_alloc(self, n: int, batch: int) -> int:
        floats = self.L.bpmax_tri_tiled_floats(n) * batch
        p = self.L.bpmax_device_alloc(self.h, floats * 4)
        if not p:
            raise BPMaxError(ENOMEM, self.L.bpmax_last_error(self.h).decode())
        return p

    def tri_free(self, p: int):
        self._check(self.L.bpmax_device_free(self.h, p))

    def tri_upload(self, dense: np.ndarray, dev: int):
        d = np.ascontiguousarray(dense, np.float32)
        batch, n, _ = d.shape
        self._check(self.L.bpmax_tri_upload(self.h, n, batch, _fp(d), dev))

    def tri_download(self, dev: int, n: int, batch: int) -> np.ndarray:
        out = np.empty((batch, n, n), np.float32)
        self._check(self.L.bpmax_tri_download(self.h, n, batch, dev, _fp(out)))
        return out

    def tri_fill_random(self, dev: int, n: int, batch: int, seed: int, lo: int, hi: int):
        """Synthetic integer operands in [lo, hi] on/above the diagonal, -inf below."""
        self._check(self.L.bpmax_tri_fill_random(self.h, n, batch, seed, lo, hi, dev))

    def r0_tri_batch(self, n: int, batch: int, A: int, B: int, C: int):
        self._check(self.L.bpmax_r0_tri_batch_device(self.h, n, batch, A, B, C))

    def tri_to_half(self, src: int, dst: int, n: int, batch: int) -> bool:
        """FP16 copies of tiled triangles (dst from tri_alloc holds twice what they need);
        True when every value is -inf or an integer of magnitude <= 1024, so that the FP16
        R0 gives exact results (bpmax_tri_to_half)."""
        exact = ctypes.c_int()
        self._check(self.L.bpmax_tri_to_half(self.h, n, batch, src, dst, ctypes.byref(exact)))
        return bool(exact.value)

    def r0_tri_batch_half(self, n: int, batch: int, Ah: int, Bh: int, C: int):
        """C max= A (max,+) B on FP16 operand copies, C in FP32 (bpmax_r0_tri_batch_device_half)."""
        self._check(self.L.bpmax_r0_tri_batch_device_half(self.h, n, batch, Ah, Bh, C))

    # -------------------------------------------------------------- schedule / timing
    def set_schedule(self, schedule, chunk: int = 0):
        """SCHED_PIPE (default: per-phase launches on two streams, graph-replayed),
        SCHED_LAUNCH (the same launches on one stream) or SCHED_TASKS (one dependency-tile
        scheduler launch per solve), or their names; chunk = K-steps per task."""
        if isinstance(schedule, str):
            schedule = SCHEDULES[schedule]
        self._check(self.L.bpmax_ctx_set_schedule(self.h, schedule, chunk))

    def set_r0_half(self, on: bool):
        """FP16 R0 operands where they are exact (default on; bpmax_ctx_set_r0_half)."""
        self._check(self.L.bpmax_ctx_set_r0_half(self.h, 1 if on else 0))

    def r0_half_used(self) -> bool:
        """Whether the last solve ran R0 on FP16 operands (bpmax_ctx_r0_half_used)."""
        return bool(self.L.bpmax_ctx_r0_half_used(self.h))

    def set_graphs(self, on: bool):
        """Capture and replay the pipelined schedule as a HIP graph (default off)."""
        self._check(self.L.bpmax_ctx_set_graphs(self.h, 1 if on else 0))

    def set_timing(self, on: bool, classes=None):
        """HIP events around every launch (or only those of the named kernel classes)."""
        if classes is None or not on:
            self._check(self.L.bpmax_ctx_set_timing(self.h, 1 if on else 0))
        else:
            mask = 0
            for c in classes:
                mask |= 1 << KERNEL_CLASSES.index(c)
            self._check(self.L.bpmax_ctx_set_timing_mask(self.h, mask))

    def task_stats(self) -> dict:
        """Tasks run and polls spent waiting in the last task-scheduler run."""
        n = ctypes.c_longlong()
        p = ctypes.c_double()
        self._check(self.L.bpmax_ctx_task_stats(self.h, ctypes.byref(n), ctypes.byref(p)))
        return {"tasks": n.value, "wait_polls": p.value}

    def reset_stats(self):
        self._check(self.L.bpmax_ctx_reset_stats(self.h))

    def kernel_stats(self) -> dict:
        out = {}
        for k, name in enumerate(KERNEL_CLASSES):
            n = ctypes.c_longlong()
            ms = ctypes.c_double()
            upd = ctypes.c_double()
            self._check(self.L.bpmax_ctx_kernel_stats(self.h, k, ctypes.byref(n), ctypes.byref(ms),
                                                      ctypes.byref(upd)))
            out[name] = {"launches": n.value, "ms": ms.value, "updates": upd.value}
        return out

    # ---------------------------------------------------------------- autotuning
    def tune(self, M: int, N: int, npairs: int = 1, reps: int = 2) -> dict:
        """Time the candidate launch schedules on synthetic M x N pairs and keep the fastest
        for the problem class of (M, N, npairs) (bpmax_tune_shape)."""
        t = Tuning()
        ms = ctypes.c_double()
        self._check(self.L.bpmax_tune_shape(self.h, M, N, npairs, reps, ctypes.byref(t), ctypes.byref(ms)))
        return dict(t.as_dict(), ms=ms.value)

    def tune_batch(self, seqs1, seqs2, reps: int = 1) -> dict:
        """As tune(), on these pairs (bpmax_tune_batch)."""
        n = len(seqs1)
        a1 = [np.ascontiguousarray(x, np.int32) for x in seqs1]
        a2 = [np.ascontiguousarray(x, np.int32) for x in seqs2]
        P = ctypes.POINTER(ctypes.c_int)
        p1 = (P * n)(*[_ip(a) for a in a1])
        p2 = (P * n)(*[_ip(a) for a in a2])
        Ms = np.array([len(a) for a in a1], np.int32)
        Ns = np.array([len(a) for a in a2], np.int32)
        t = Tuning()
        ms = ctypes.c_double()
        self._check(self.L.bpmax_tune_batch(self.h, n, p1, _ip(Ms), p2, _ip(Ns), reps, ctypes.byref(t),
                                            ctypes.byref(ms)))
        return dict(t.as_dict(), ms=ms.value)

    def set_tuning(self, M: int, N: int, npairs: int, tuning):
        """Set the settings of a problem class (dict as returned by tune(); None clears)."""
        if tuning is None:
            self._check(self.L.bpmax_ctx_set_tuning(self.h, M, N, npairs, None))
            return
        t = Tuning.from_dict(tuning)
        self._check(self.L.bpmax_ctx_set_tuning(self.h, M, N, npairs, ctypes.byref(t)))

    def set_default_tuning(self, tuning: dict):
        """Settings for problem classes without an entry (bpmax_ctx_set_default_tuning)."""
        t = Tuning.from_dict(tuning)
        self._check(self.L.bpmax_ctx_set_default_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self, M: int, N: int, npairs: int = 1):
        t = Tuning()
        rc = self.L.bpmax_ctx_get_tuning(self.h, M, N, npairs, ctypes.byref(t))
        if rc == ENOTFOUND:
            return None
        self._check(rc)
        return t.as_dict()

    # -------------------------------------------------------------- serving loop
    def server_start(self, memory_budget: int = 0, max_batch: int = 0, slots: int = 0, queue_capacity: int = 0,
                     keep_tables: bool = False, intra=None, inter=None):
        """Start the continuous-batching server on this context (bpmax_server_start)."""
        ti = _table(intra, None)
        te = _table(inter, ti)
        self._server_tables = (ti, te)  # kept alive while the server runs
        cfg = ServerConfig(memory_budget, max_batch, slots, queue_capacity, 1 if keep_tables else 0,
                           None if ti is None else _fp(ti), None if te is None else _fp(te))
        self._check(self.L.bpmax_server_start(self.h, ctypes.byref(cfg)))

    def server_submit(self, seq1, seq2) -> int:
        s1 = np.ascontiguousarray(seq1, np.int32)
        s2 = np.ascontiguousarray(seq2, np.int32)
        rid = ctypes.c_longlong()
        self._check(self.L.bpmax_server_submit(self.h, _ip(s1), len(s1), _ip(s2), len(s2), ctypes.byref(rid)))
        return rid.value

    def server_poll(self, rid: int):
        """(score, latency_ms), or None while the request is queued or running."""
        sc = ctypes.c_float()
        lat = ctypes.c_double()
        rc = self.L.bpmax_server_poll(self.h, rid, ctypes.byref(sc), ctypes.byref(lat))
        if rc == ENOTREADY:
            return None
        self._check(rc)
        return float(sc.value), float(lat.value)

    def server_wait(self, rid: int, timeout_ms: float = -1.0):
        sc = ctypes.c_float()
        lat = ctypes.c_double()
        rc = self.L.bpmax_server_wait(self.h, rid, timeout_ms, ctypes.byref(sc), ctypes.byref(lat))
        if rc == ENOTREADY:
            return None
        self._check(rc)
        return float(sc.value), float(lat.value)

    def server_fetch(self, rid: int, M: int, N: int, layout: int = LAYOUT_PACKED) -> np.ndarray:
        if layout == LAYOUT_PACKED:
            out = np.empty((M * (M + 1) // 2) * (N * (N + 1) // 2), np.float32)
        else:
            out = np.empty((M, M, N, N), np.float32)
        self._check(self.L.bpmax_server_fetch(self.h, rid, layout, _fp(out)))
        return out

    def server_release(self, rid: int):
        self._check(self.L.bpmax_server_release(self.h, rid))

    def server_stats(self) -> dict:
        st = ServerStats()
        self._check(self.L.bpmax_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def server_stop(self):
        self._check(self.L.bpmax_server_stop(self.h))

    def synchronize(self):
        self._check(self.L.bpmax_ctx_synchronize(self.h))

    def stream(self) -> int:
        return self.L.bpmax_ctx_stream(self.h)


def header_functions(path: str = HEADER) -> list:
    """Names of the functions declared in include/bpmax.h."""
    import re
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(bpmax\w*)\s*\(", text)
    return sorted(set(n for n in names if n not in ("bpmax_layout", "bpmax_kernel_class")))

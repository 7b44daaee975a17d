"""BPMax benchmark on one MI355X.

Headline: GCUPS of one full F-table computation at M=64, N=512.  A step is one complete
BPMax solve of one 64x512 pair (S1, S2, every outer diagonal's R0/R3/R4/P1 phase and every
inner R1/R2/P2 phase) with the sequences on the host and the F table left resident in HBM
(score-only output).  GCUPS = cells(M,N) / step time with
cells = C(M+1,3)C(N+1,3) + 2C(M+1,2)C(N+1,3) + 2C(M+1,3)C(N+1,2).

The JSON line also carries
  roofline     -- the dominant kernel, k_outer_h (R0 + init/P1/R3/R4 on FP16 operands, exact
                  here; k_outer<64> where FP16 is not exact), timed with HIP
                  events on the stream it is launched on during the timed steps:
                  algorithmic flops (2 per cell update: one add, one max) per launch over
                  its average launch duration, against the MI355X FP32 vector peak;
  cpu_baseline -- the reference's own bpmax() (oracle/_ref, compiled from the reference
                  sources) on this host's cores for the same 64x512 pair;
  batched      -- 256 mixed-size pairs (M~U[16,48], N~U[100,400], GC~U[0.3,0.7]) solved
                  in one batched schedule vs the same pairs solved one per call;
  r0_isolated  -- the R0 max-plus kernel alone on 4096 tiled 512x512 upper triangles;
  serving      -- the continuous-batching server on mixed requests (M log-uniform
                  [8,64], N log-uniform [64,1024]): saturated pairs/s and p50/p99 latency
                  under Poisson arrivals, against one solve call per pair.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

# launch schedules picked by the autotuner on an MI355X for the bench's classes
# (python -m bpmaxcpu_amd.autotune); --retune tunes them again instead
TUNING_TABLE = os.path.join(REPO, "bpmaxcpu_amd", "tuning_mi355x.json")
FP32_VALU_PEAK_TFLOPS = 157.3     # MI355X dense FP32 vector peak: 256 CUs x 2.4 GHz x 128 lanes x 2 (FMA)
VALU_ISSUE_PEAK = 2 * 256 * 2.4e9  # wave64 FP32 VALU instructions per second (2 per clock per CU)
VALU_ISSUE_PEAK_PK = 256 * 2.4e9   # wave64 packed FP16 VALU instructions per second (half the FP32 rate)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(M, N, seed=7):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 4, M).astype(np.int32), rng.integers(0, 4, N).astype(np.int32)


def gc_sequence(rng, n, gc):
    """i.i.d. bases with P(G)=P(C)=gc/2, P(A)=P(U)=(1-gc)/2 (codes A0 C1 G2 U3)."""
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return rng.choice(4, size=n, p=p).astype(np.int32)


def mixed_pairs(npairs=256, seed=11):
    rng = np.random.default_rng(seed)
    s1, s2 = [], []
    for _ in range(npairs):
        M = int(rng.integers(16, 49))
        N = int(rng.integers(100, 401))
        s1.append(gc_sequence(rng, M, rng.uniform(0.3, 0.7)))
        s2.append(gc_sequence(rng, N, rng.uniform(0.3, 0.7)))
    return s1, s2


def cpu_info():
    """CPU model and topology of this host (lscpu), and the CPUs this process may use."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    d = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            d[k.strip()] = v.strip()
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": d.get("Model name"), "sockets": d.get("Socket(s)"), "cores_per_socket": d.get("Core(s) per socket"),
            "threads_per_core": d.get("Thread(s) per core"), "cpus_online": os.cpu_count(), "cpus_usable": usable}


def cpu_baseline(M, N, s1, s2, threads, single_M=32):
    """Reference bpmax() on the host (kind 'reference'): 1 warm-up run, then the median of 3,
    with OMP_PROC_BIND=close and OMP_PLACES=cores, on `threads` threads for the full pair and
    on one thread for a single_M x N sample (the first single_M bases of seq1).  The oracle
    restatement on a bounded sample when the reference binary is absent (kind 'port')."""
    from oracle import oracle as O
    import bpmaxcpu_amd as B
    info = cpu_info()
    if O.have_reference():
        env = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
        log(f"cpu baseline: reference bpmax() {M}x{N} on {threads} threads, 1 warm-up + median of 3")
        _, score, secs, _ = O.reference_run(s1, s2, table=False, threads=threads, warmup=1, repeat=3, env_extra=env)
        cells = B.cell_updates(M, N)
        Ms = min(single_M, M)
        log(f"cpu baseline: reference bpmax() {Ms}x{N} on 1 thread, 1 warm-up + median of 3")
        _, _, secs1, _ = O.reference_run(s1[:Ms], s2, table=False, threads=1, warmup=1, repeat=3, env_extra=env)
        return {"value": round(cells / secs / 1e9, 3), "unit": "GCUPS", "cores": threads, "kind": "reference",
                "sample": f"one {M}x{N} pair (the full workload), reference [MPT+RT]:v3 bpmax() built with "
                          f"g++ -O3 -fopenmp -mavx2 -mfma + FTZ, N_tile 48, 3x24 register tile, "
                          f"OMP_NUM_THREADS={threads}, OMP_PROC_BIND=close, OMP_PLACES=cores; 1 warm-up run, "
                          f"median of 3: {secs:.2f} s, score {score}",
                "cpu": info,
                "single_core": {"value": round(B.cell_updates(Ms, N) / secs1 / 1e9, 3), "unit": "GCUPS", "cores": 1,
                                "sample": f"{Ms}x{N} (first {Ms} bases of seq1), 1 warm-up run, median of 3: "
                                          f"{secs1:.2f} s"}}
    Ms = 16
    log(f"cpu baseline: oracle port on a {Ms}x{N} sample")
    os.environ["OMP_NUM_THREADS"] = str(threads)
    t0 = time.time()
    O.bpmax(s1[:Ms], s2)
    secs = time.time() - t0
    return {"value": round(B.cell_updates(Ms, N) / secs / 1e9, 3), "unit": "GCUPS", "cores": threads,
            "kind": "port", "sample": f"oracle restatement on the first {Ms} bases of seq1 x all {N} of seq2; "
                                      f"{secs:.2f} s", "cpu": info}


def parse_schedule(text):
    """'pipe:2:4096:0:64:2[:1[:32]]' -> tuning dict (schedule, outer_split, outer_items, graphs,
    intra_wide, outer_kgroup[, closure_ahead[, tile]])."""
    f = text.split(":")
    return {"schedule": f[0], "outer_split": int(f[1]), "outer_items": int(f[2]), "graphs": bool(int(f[3])),
            "intra_wide": int(f[4]), "outer_kgroup": int(f[5]), "closure_ahead": len(f) > 6 and f[6] == "1",
            "tile": int(f[7]) if len(f) > 7 else 0}


def pair_leg(B, eng, M, N, steps, warmup, schedule=None, retune=False):
    s1, s2 = workload(M, N)
    cells = B.cell_updates(M, N)
    t0 = time.perf_counter()
    tuned = None if retune else eng.get_tuning(M, N, 1)
    if schedule:  # a fixed launch schedule (profiling runs: the same launches every run)
        tuned = parse_schedule(schedule)
        eng.set_tuning(M, N, 1, tuned)
        log(f"fixed schedule {tuned}")
    elif tuned is not None:  # the class's entry in the shipped tuning table: the same launches every run
        log(f"tuning table entry for {M}x{N}: {tuned}")
    else:
        tuned = eng.tune(M, N, 1, reps=2)  # launch schedule of this problem class (setup, untimed)
        log(f"autotuned {M}x{N} in {time.perf_counter() - t0:.1f} s: {tuned}")
    log(f"warmup {warmup} x {M}x{N}")
    score = None
    for _ in range(warmup):
        _, score = eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    eng.synchronize()
    # Timed steps.  The last one records HIP events around every R0 launch, on the stream it
    # runs on (the roofline's launch durations); the others run as a caller runs them (events
    # between launches turn graph replay off and cost ~2 ms per 64x512 step).
    eng.reset_stats()
    log(f"timed {steps} steps (R0 events in the last)")
    t0 = time.perf_counter()
    for i in range(steps):
        if i == steps - 1:
            eng.set_timing(True, classes=["outer"])
        _, score = eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    eng.synchronize()
    wall = time.perf_counter() - t0
    r0_half = eng.r0_half_used()  # R0 of these solves on FP16 operands (exact for this table and shape)
    st = eng.kernel_stats()
    eng.set_timing(False)
    timed_solves = 1
    ms_per_step = wall * 1e3 / steps
    gcups = cells / (ms_per_step / 1e3) / 1e9
    # the same steps with the events in every step, to show what the instrumentation costs
    eng.set_timing(True, classes=["outer"])
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    eng.synchronize()
    ms_events = (time.perf_counter() - t0) * 1e3 / steps
    eng.set_timing(False)
    log(f"{ms_per_step:.2f} ms/step ({ms_events:.2f} with R0 events in every step), {gcups:.1f} GCUPS, score {score}")
    # every class timed (costs ~15% on this many launches): per-class durations while overlapped
    eng.set_timing(True)
    eng.reset_stats()
    eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    phases = {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in eng.kernel_stats().items()
              if v["launches"]}
    eng.set_timing(False)

    o = st["outer"]
    flops_per_launch = 2.0 * o["updates"] / max(1, o["launches"])
    avg_s = o["ms"] / max(1, o["launches"]) / 1e3
    achieved = flops_per_launch / avg_s / 1e12 if avg_s > 0 else 0.0
    traffic = None
    kname = "k_outer_h" if r0_half else "k_outer"
    tpath = os.path.join(REPO, "profiles", f"{kname}_64x512_traffic.json")
    if (M, N) == (64, 512) and os.path.exists(tpath) and o["launches"]:
        # rocprofv3 FETCH_SIZE + WRITE_SIZE of one solve's k_outer launches (copy-calibrated),
        # spread over this run's launches per solve (the tuned R0 split changes their number)
        with open(tpath) as fh:
            traffic = int(round(json.load(fh)["bytes_per_solve"] * timed_solves / o["launches"]))
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": ("k_outer_h" if r0_half else "k_outer<64>") +
                          ": R0 double max-plus + init/P1/R3/R4 (2 flops per cell update)",
                "launches": o["launches"], "avg_launch_ms": round(avg_s * 1e3, 4),
                "flops_per_launch": flops_per_launch}
    vpath = os.path.join(REPO, "profiles", f"{kname}_64x512_valu.json")
    if (M, N) == (64, 512) and os.path.exists(vpath) and o["launches"] and avg_s > 0:
        # rocprofv3 SQ_INSTS_VALU of one solve (calibrated to count wave instructions exactly)
        # spread over this run's launches per solve, against the issue peak of the kernel's
        # instruction kind: two wave64 FP32 instructions per clock per CU, or one packed FP16
        # instruction (half rate; bench.py --calibrate-valu measures both)
        with open(vpath) as fh:
            per_solve = json.load(fh)["sq_insts_valu_per_solve"]
        per_launch = per_solve * timed_solves / o["launches"]
        rate = per_launch / avg_s
        peak = VALU_ISSUE_PEAK_PK if r0_half else VALU_ISSUE_PEAK
        roofline["valu_issue"] = {"wave_instr_per_launch": per_launch, "achieved": rate,
                                  "peak": peak, "unit": "wave VALU instructions/s",
                                  "frac": round(rate / peak, 4),
                                  "source": f"profiles/{kname}_64x512_valu.md"}

    # one-stream schedule, for the phase breakdown without overlap
    eng.set_tuning(M, N, 1, {"schedule": "launch", "outer_split": 1, "outer_items": tuned["outer_items"],
                             "graphs": False, "intra_wide": tuned["intra_wide"],
                             "outer_kgroup": tuned.get("outer_kgroup", 1),
                             "closure_ahead": tuned.get("closure_ahead", False)})
    eng.set_schedule(B.SCHED_LAUNCH)
    eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    eng.set_timing(True)
    eng.reset_stats()
    eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    st2 = eng.kernel_stats()
    eng.set_timing(False)
    eng.set_schedule(B.SCHED_PIPE)
    serial = {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in st2.items() if v["launches"]}
    o2 = st2["outer"]
    r0_serial = 2.0 * o2["updates"] / (o2["ms"] / 1e3) / 1e12 if o2["ms"] > 0 else 0.0
    roofline["k_outer_alone"] = {"achieved": round(r0_serial, 3), "frac": round(r0_serial / FP32_VALU_PEAK_TFLOPS, 4),
                                 "note": "one-stream schedule, one R0 launch per outer diagonal, no overlap"}
    log(f"k_outer {achieved:.2f} TFLOP/s in the pipelined run, {r0_serial:.2f} alone; phases {serial}")
    eng.set_tuning(M, N, 1, tuned)  # the one-stream breakdown above overrode the class
    # the dependency-tile scheduler (one persistent launch per solve) beside the graph path
    eng.set_schedule("tasks", 64)
    eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    t0 = time.perf_counter()
    _, sc_t = eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    ms_tasks = (time.perf_counter() - t0) * 1e3
    ts = eng.task_stats()
    eng.set_schedule(B.SCHED_PIPE)
    tasks = {"ms": round(ms_tasks, 2), "score_equal": sc_t == score, "chunk": 64, "tasks": ts["tasks"],
             "wait_polls": ts["wait_polls"],
             "note": "one persistent launch; workgroups take tasks in a topological order and wait on "
                     "per-tile counters (agent-scope acquire/release per task)"}
    log(f"dependency-tile scheduler: {ms_tasks:.1f} ms ({ts['tasks']} tasks, {ts['wait_polls']:.3g} waiting polls)")
    return {"score": score, "ms_per_step": ms_per_step, "ms_events": ms_events, "gcups": gcups,
            "r0_half": r0_half,
            "roofline": roofline, "phases_pipelined": phases, "phases_one_stream": serial, "tuning": tuned,
            "dependency_tiles": tasks}


def tree6(M, N, N_tile):
    """The reference driver's 6-level F pointer tree over one contiguous buffer
    (src/bpmax-wrapper.c:175-197), built with numpy address arithmetic: returns (buffer,
    top-level pointer array, keep-alive list)."""
    N_sec = (N + N_tile - 1) // N_tile
    sec = (N_tile + 1) * N_tile
    lin = np.empty(M * M * N_sec * N_sec * sec, np.float32)  # pages touched only where written
    base = lin.ctypes.data
    n4 = M * M * N_sec * N_sec
    l5 = (base + 4 * (np.arange(n4, dtype=np.uint64)[:, None] * sec +
                      np.arange(N_tile + 1, dtype=np.uint64)[None, :] * N_tile)).astype(np.uint64)
    l4 = l5.ctypes.data + 8 * (N_tile + 1) * np.arange(n4, dtype=np.uint64)
    l3 = l4.ctypes.data + 8 * N_sec * np.arange(M * M * N_sec, dtype=np.uint64)
    l2 = l3.ctypes.data + 8 * N_sec * np.arange(M * M, dtype=np.uint64)
    l1 = l2.ctypes.data + 8 * M * np.arange(M, dtype=np.uint64)
    return lin, l1, [l5, l4, l3, l2]


def dropin_leg(B, eng, M, N, s1, s2, N_tile=48):
    """What a caller of the reference's bpmax() pays: the drop-in symbol (GPU solve, packed F
    to the host, scatter into the reference's 6-D tree), and the packed copy alone."""
    import ctypes
    N_sec = (N + N_tile - 1) // N_tile
    R = N_sec * N_tile - N
    lin, top, keep = tree6(M, N, N_tile)
    fn = B.load_library().bpmax
    fn.restype = None
    a1, a2 = np.ascontiguousarray(s1, np.int32), np.ascontiguousarray(s2, np.int32)
    args = [ctypes.c_long(M), ctypes.c_long(N), ctypes.c_long(N_sec), ctypes.c_long(N_tile), ctypes.c_long(R),
            ctypes.c_long(3), ctypes.c_long(24), ctypes.c_long(16), B._ip(a1), B._ip(a2), ctypes.c_void_p(top.ctypes.data)]
    fn(*args)  # first call: context, buffers, page faults of the tree
    t0 = time.perf_counter()
    fn(*args)
    ms_6d = (time.perf_counter() - t0) * 1e3
    sc = float(lin[(((0 * M + M - 1) * N_sec + R // N_tile) * N_sec + (N - 1 + R) // N_tile) * (N_tile + 1) * N_tile +
                   (R % N_tile) * N_tile + (N - 1 + R) % N_tile])
    eng.solve(s1, s2, layout=B.LAYOUT_PACKED)
    t0 = time.perf_counter()
    _, sp = eng.solve(s1, s2, layout=B.LAYOUT_PACKED)
    ms_packed = (time.perf_counter() - t0) * 1e3
    cells = B.cell_updates(M, N)
    del keep, lin
    log(f"drop-in bpmax() with the 6-D tree {ms_6d:.1f} ms; packed F to host {ms_packed:.1f} ms")
    return {"bpmax_6d_tree": {"ms": round(ms_6d, 2), "gcups": round(cells / ms_6d / 1e6, 1), "score": sc,
                              "n_tile": N_tile},
            "packed_to_host": {"ms": round(ms_packed, 2), "gcups": round(cells / ms_packed / 1e6, 1), "score": sp,
                               "bytes": int(M * (M + 1) // 2 * (N * (N + 1) // 2) * 4)}}


def batched_leg(B, eng, reps=2, retune=False):
    s1, s2 = mixed_pairs()
    n = len(s1)
    cells = sum(B.cell_updates(len(a), len(b)) for a, b in zip(s1, s2))
    # the launch schedule of the batch class: the shipped table's entry (tuned on this batch),
    # or tuned on this batch now (setup, untimed)
    tuned = None if retune else eng.get_tuning(max(len(a) for a in s1), max(len(b) for b in s2), n)
    if tuned is not None:
        log(f"tuning table entry for the batch class: {tuned}")
    else:
        t0 = time.perf_counter()
        tuned = eng.tune_batch(s1, s2, reps=1)
        log(f"autotuned the batch class in {time.perf_counter() - t0:.1f} s: {tuned}")
    log(f"batched: {n} mixed pairs, {cells / 1e12:.2f} T cells")
    eng.solve_batch(s1, s2)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        sb = eng.solve_batch(s1, s2)
        ts.append(time.perf_counter() - t0)
    tb = min(ts)
    log(f"batched {tb * 1e3:.1f} ms ({n / tb:.1f} pairs/s); now one solve per pair")
    t0 = time.perf_counter()
    single = np.array([eng.solve(a, b, layout=B.LAYOUT_NONE)[1] for a, b in zip(s1, s2)], np.float32)
    ts_ = time.perf_counter() - t0
    same = bool(np.array_equal(single, sb))
    log(f"per-pair {ts_ * 1e3:.1f} ms ({n / ts_:.1f} pairs/s); scores identical: {same}")
    return {"workload": f"{n} pairs, M~U[16,48], N~U[100,400], GC~U[0.3,0.7], seed 11",
            "pairs_per_s": round(n / tb, 2), "gcups": round(cells / tb / 1e9, 2), "ms": round(tb * 1e3, 2),
            "per_pair_calls": {"pairs_per_s": round(n / ts_, 2), "ms": round(ts_ * 1e3, 2)},
            "speedup_vs_per_pair": round(ts_ / tb, 3), "scores_identical": same, "tuning": tuned}


def r0_leg(B, eng, n=512, batch=4096, reps=3):
    tri = eng.tri_tiled_floats(n)
    log(f"r0 isolated: {batch} x {n}x{n} tiled triangles ({3 * tri * batch * 4 / 2**30:.1f} GiB)")
    A = eng.tri_alloc(n, batch)
    Bm = eng.tri_alloc(n, batch)
    C = eng.tri_alloc(n, batch)
    try:
        eng.tri_fill_random(A, n, batch, 1, 0, 200)
        eng.tri_fill_random(Bm, n, batch, 2, 0, 200)
        eng.tri_fill_random(C, n, batch, 3, 0, 200)
        # two output matrices checked against a numpy max-plus of their operands
        picks = [0, batch - 1]
        ops = {m: [eng.tri_download(d + 4 * m * tri, n, 1)[0] for d in (A, Bm, C)] for m in picks}
        eng.r0_tri_batch(n, batch, A, Bm, C)
        checked = True
        for m in picks:
            a, b, c = ops[m]
            want = c.copy()
            for i in range(n):
                want[i, i:] = np.maximum(want[i, i:], (a[i, i:, None] + b[i:, i:]).max(axis=0))
            got = eng.tri_download(C + 4 * m * tri, n, 1)[0]
            checked = checked and bool(np.array_equal(got.view(np.uint32), want.view(np.uint32)))
        eng.set_timing(True)
        eng.reset_stats()
        for _ in range(reps):
            eng.r0_tri_batch(n, batch, A, Bm, C)
        st = eng.kernel_stats()["r0tri"]
        avg_ms = st["ms"] / st["launches"]
        upd = st["updates"] / st["launches"]
        tf = 2.0 * upd / (avg_ms / 1e3) / 1e12
        log(f"r0 isolated {avg_ms:.2f} ms/launch, {tf:.2f} TFLOP/s")
        out = {"workload": f"{batch} upper-triangular {n}x{n} fp32 triples, integers U[0,200], -inf below",
               "ms_per_launch": round(avg_ms, 3), "achieved": round(tf, 3), "unit": "TFLOP/s",
               "frac": round(tf / FP32_VALU_PEAK_TFLOPS, 4), "kernel": "k_r0tri<64>",
               "checked": {"matrices": picks, "equal_to_numpy_maxplus": checked}}
        # the same product on FP16 copies of A and B (exact: integers <= 1024), copies timed apart
        Ah, Bh = eng.tri_alloc(n, (batch + 1) // 2), eng.tri_alloc(n, (batch + 1) // 2)
        try:
            eng.reset_stats()
            exact = all(eng.tri_to_half(src, dst, n, batch) for src, dst in ((A, Ah), (Bm, Bh)))
            for m in picks:
                ops[m][2] = eng.tri_download(C + 4 * m * tri, n, 1)[0]
            eng.r0_tri_batch_half(n, batch, Ah, Bh, C)
            ok = exact
            for m in picks:
                a, b, c = ops[m]
                want = c.copy()
                for i in range(n):
                    want[i, i:] = np.maximum(want[i, i:], (a[i, i:, None] + b[i:, i:]).max(axis=0))
                got = eng.tri_download(C + 4 * m * tri, n, 1)[0]
                ok = ok and bool(np.array_equal(got.view(np.uint32), want.view(np.uint32)))
            copy_ms = eng.kernel_stats()["convert"]["ms"] / 2
            eng.reset_stats()
            for _ in range(reps):
                eng.r0_tri_batch_half(n, batch, Ah, Bh, C)
            st = eng.kernel_stats()["r0tri"]
            h_ms = st["ms"] / st["launches"]
            h_tf = 2.0 * upd / (h_ms / 1e3) / 1e12
            log(f"r0 isolated on FP16 copies {h_ms:.2f} ms/launch ({h_tf:.2f} TFLOP/s), copies {copy_ms:.2f} ms each")
            out["fp16_operands"] = {
                "kernel": "k_r0tri_h", "ms_per_launch": round(h_ms, 3), "achieved": round(h_tf, 3),
                "frac": round(h_tf / FP32_VALU_PEAK_TFLOPS, 4), "copy_ms_per_operand": round(copy_ms, 3),
                "achieved_with_both_copies": round(2.0 * upd / ((h_ms + 2 * copy_ms) / 1e3) / 1e12, 3),
                "checked": {"matrices": picks, "exact_copies_and_equal_to_numpy_maxplus": ok},
                "note": "FP16 copies exact for integer operands of magnitude <= 1024; with FP32 inputs "
                        "the two copy passes cost more than the kernel saves, so FP32 operands stay the "
                        "figure of this workload"}
        finally:
            eng.set_timing(False)
            eng.tri_free(Ah)
            eng.tri_free(Bh)
        return out
    finally:
        for p in (A, Bm, C):
            eng.tri_free(p)


def serving_requests(n, seed=13):
    """M log-uniform in [8,64], N log-uniform in [64,1024], GC content 0.5."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        M = int(round(np.exp(rng.uniform(np.log(8), np.log(64)))))
        N = int(round(np.exp(rng.uniform(np.log(64), np.log(1024)))))
        out.append((gc_sequence(rng, M, 0.5), gc_sequence(rng, N, 0.5)))
    return out


def _pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if v else None


def serving_leg(B, eng, n_plain=160, n_sat=160, n_rate=160, fractions=(0.25, 0.5)):
    """Continuous-batching server on the context `eng`: saturated throughput and p50/p99
    latency at Poisson arrival rates below saturation, against the plain path (one solve call
    per pair on the same context) on the same requests."""
    reqs = serving_requests(max(n_plain, n_sat, n_rate))
    cells = [B.cell_updates(len(a), len(b)) for a, b in reqs]
    try:
        log(f"serving: plain per-pair path on {n_plain} requests")
        eng.solve(*reqs[0], layout=B.LAYOUT_NONE)
        t0 = time.perf_counter()
        plain_scores = [eng.solve(a, b, layout=B.LAYOUT_NONE)[1] for a, b in reqs[:n_plain]]
        t_plain = time.perf_counter() - t0
        plain_rate = n_plain / t_plain
        log(f"serving: plain {plain_rate:.1f} pairs/s")
        eng.server_start()  # budget: 80% of free HBM, admission by table footprint
        st0 = eng.server_stats()

        def run(schedule):
            """schedule: arrival offsets (s) per request index; returns (wall, latencies, scores)."""
            t_start = time.perf_counter()
            ids = []
            for k, t_arr in enumerate(schedule):
                delay = t_start + t_arr - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(eng.server_submit(*reqs[k]))
            res = []
            for i in ids:
                r = eng.server_wait(i, 5000.0)
                while r is None:  # still queued or running: report progress, bounded
                    st = eng.server_stats()
                    log(f"serving: waiting, {st['completed'] - st0['completed']} done of {st['submitted']}, "
                        f"{st['bytes_in_use'] / 1e9:.1f} GB in use")
                    if time.perf_counter() - t_start > 240:
                        raise RuntimeError("serving run exceeded 240 s")
                    r = eng.server_wait(i, 5000.0)
                res.append(r)
            wall = time.perf_counter() - t_start
            for i in ids:
                eng.server_release(i)
            return wall, [r[1] for r in res], [r[0] for r in res]

        wall, lat, sc = run([0.0] * n_sat)
        sat_rate = n_sat / wall
        same = bool(np.array_equal(np.array(sc[:n_plain], np.float32), np.array(plain_scores, np.float32)))
        log(f"serving: saturated {sat_rate:.1f} pairs/s, p50 {_pct(lat, 50):.1f} ms, p99 {_pct(lat, 99):.1f} ms, "
            f"scores identical to plain: {same}")
        rates = []
        rng = np.random.default_rng(17)
        for f in fractions:
            lam = f * sat_rate
            arr = np.cumsum(rng.exponential(1.0 / lam, n_rate))
            arr -= arr[0]
            wall_r, lat_r, _ = run(list(arr))
            rates.append({"offered_pairs_per_s": round(lam, 2), "achieved_pairs_per_s": round(n_rate / wall_r, 2),
                          "p50_ms": round(_pct(lat_r, 50), 2), "p99_ms": round(_pct(lat_r, 99), 2)})
            log(f"serving: offered {lam:.1f}/s -> p50 {rates[-1]['p50_ms']} ms, p99 {rates[-1]['p99_ms']} ms")
        st = eng.server_stats()
        eng.server_stop()
        return {"workload": "pair requests, M log-uniform [8,64], N log-uniform [64,1024], GC 0.5, seed 13; "
                            "admission by table footprint against 80% of free HBM",
                "plain_per_pair": {"pairs_per_s": round(plain_rate, 2), "requests": n_plain},
                "saturated": {"pairs_per_s": round(sat_rate, 2), "requests": n_sat,
                              "gcups": round(sum(cells[:n_sat]) / wall / 1e9, 1),
                              "p50_ms": round(_pct(lat, 50), 2), "p99_ms": round(_pct(lat, 99), 2),
                              "speedup_vs_plain": round(sat_rate / plain_rate, 3)},
                "poisson": rates, "scores_identical_to_plain": same,
                "server": {"budget_gb": round(st0["budget"] / 1e9, 1), "ticks": st["batches"],
                           "max_pairs_per_tick": st["max_batch_seen"], "max_resident_pairs": st["max_inflight_pairs"],
                           "peak_gb_in_use": round(st["peak_bytes_in_use"] / 1e9, 2)}}
    finally:
        if getattr(eng, "_server_tables", None) is not None:
            try:
                eng.server_stop()
            except B.BPMaxError:
                pass


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and isolated-R0 legs")
    ap.add_argument("--schedule", default=None, metavar="S:SPLIT:ITEMS:GRAPHS:WIDE:KGROUP[:AHEAD]",
                    help="fixed launch schedule of the pair leg instead of the tuning table's, e.g. pipe:2:4096:0:64:2")
    ap.add_argument("--retune", action="store_true",
                    help="autotune the pair and batch classes on this GPU instead of loading the shipped "
                         "tuning table (bpmaxcpu_amd/tuning_mi355x.json)")
    ap.add_argument("--write-tuning", default=None, metavar="PATH",
                    help="only autotune the pair class (MxN) and the batched leg's class (on its own 256 "
                         "pairs) on this GPU and write them as a tuning table to PATH")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel over MIB MiB (reads MIB, writes MIB): the "
                         "reference point that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="only run the VALU counter-calibration kernels (libbpmax_probe.so): FP32, then packed "
                         "FP16, two launches each of 2048 workgroups x 4 waves with 32 x 1000 and 32 x 4000 VALU "
                         "instructions per wave, the reference points that scale rocprofv3 SQ_INSTS_VALU to wave "
                         "instructions and measure each kind's issue rate")
    ap.add_argument("--r0-only", action="store_true",
                    help="only run the isolated R0 leg (4096 tiled 512x512 triangles, FP32 then FP16 "
                         "operands) and print its JSON")
    args = ap.parse_args()
    if args.calibrate_valu:
        import ctypes
        P = ctypes.CDLL(os.path.join(REPO, "bpmaxcpu_amd", "libbpmax_probe.so"))
        out = []
        for name in ("bpmax_probe_valu", "bpmax_probe_valu_pk"):
            fn = getattr(P, name)
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
            for iters in (1000, 4000):
                ms = ctypes.c_double()
                rc = fn(2048, iters, ctypes.byref(ms))
                if rc != 0:
                    raise SystemExit(f"probe failed: {rc}")
                waves = 2048 * 4
                out.append({"probe": "packed fp16" if name.endswith("_pk") else "fp32", "iters": iters,
                            "waves": waves, "valu_per_wave_loop": 32 * iters, "ms": round(ms.value, 4),
                            "wave_valu_per_clock_per_cu": round(waves * 32 * iters / (ms.value * 1e-3) / 2.4e9 / 256, 3)})
        print(json.dumps({"valu_probe": out}), flush=True)
        return
    if args.calibrate_copy:
        import torch
        n = (args.calibrate_copy << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; pairs are independent, so a multi-GPU run is "
              "pair-level data parallelism with no communication (not measured here).", file=sys.stderr)
        sys.exit(2)

    import bpmaxcpu_amd as B
    M, N = args.M, args.N
    eng = B.Engine(0)
    if args.r0_only:
        print(json.dumps({"r0_isolated": r0_leg(B, eng)}), flush=True)
        eng.close()
        return
    if args.write_tuning:
        s1, s2 = mixed_pairs()
        table = [{"M": M, "N": N, "pairs": 1, **eng.tune(M, N, 1, reps=3)},
                 {"M": max(len(a) for a in s1), "N": max(len(b) for b in s2), "pairs": len(s1),
                  **eng.tune_batch(s1, s2, reps=2)}]
        with open(args.write_tuning, "w") as fh:
            json.dump(table, fh, indent=1)
        print(json.dumps(table), flush=True)
        eng.close()
        return
    if not args.retune and os.path.exists(TUNING_TABLE):
        from bpmaxcpu_amd.autotune import load_table
        load_table(eng, TUNING_TABLE)
    res = pair_leg(B, eng, M, N, args.steps, args.warmup, args.schedule, args.retune)
    batched = r0 = serving = dropin = None
    if not args.no_extra:
        s1d, s2d = workload(M, N)
        dropin = dropin_leg(B, eng, M, N, s1d, s2d)
        batched = batched_leg(B, eng, retune=args.retune)
        r0 = r0_leg(B, eng)
        serving = serving_leg(B, eng)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        s1, s2 = workload(M, N)
        cpu = cpu_baseline(M, N, s1, s2, threads)
        log(f"cpu baseline {cpu['value']:.2f} GCUPS on {threads} threads")

    line = {"metric": "BPMax GCUPS at M=64,N=512", "value": round(res["gcups"], 3), "unit": "GCUPS", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res["ms_per_step"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f16/f32" if res["r0_half"] else "f32",
            "dtype_note": ("R0 (k_outer_h) multiplies packed FP16 operands, exact here: scores are multiples of "
                           "2^-1 and at most 3 * 289 < 2^10; the rest is FP32; F and the score are "
                           "bit-identical to the reference's FP32") if res["r0_half"] else "FP32 throughout",
            "data": "synthetic: uniform i.i.d. ACGU from seed 7, reference weight table (GC=3, AU=1, GU=1.5)",
            "config": {"workload": f"plain BPMax, one pair M={M} N={N}, score-only output, F resident in HBM",
                       "M": M, "N": N, "pairs_per_step": 1, "score": res["score"],
                       "schedule": res["tuning"]},
            "roofline": res["roofline"], "cpu_baseline": cpu,
            "ms_per_step_with_r0_events_every_step": round(res["ms_events"], 3),
            "phases_pipelined": res["phases_pipelined"], "phases_one_stream": res["phases_one_stream"],
            "dependency_tiles": res["dependency_tiles"],
            "dropin": dropin, "batched": batched, "r0_isolated": r0, "serving": serving}
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()

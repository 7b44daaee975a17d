"""BPMax benchmark on one MI355X: GCUPS of one full F-table computation at M=64, N=512.

A step is one complete BPMax solve of one 64x512 pair (S1, S2, every outer diagonal's
R0/R3/R4/P1 phase and every inner R1/R2/P2 phase) with sequences already on the host and
the F table left resident in HBM (score-only output).  GCUPS = cells(M,N) / step time,
cells = C(M+1,3)C(N+1,3) + 2C(M+1,2)C(N+1,3) + 2C(M+1,3)C(N+1,2).

Besides the JSON line's headline value the line carries
  roofline     -- the dominant kernel (k_outer: R0 + init/P1/R3/R4) from HIP events on the
                  engine's stream: algorithmic flops (2 per cell update) / kernel time,
                  against the MI355X FP32 vector peak;
  cpu_baseline -- the reference's own bpmax() (oracle/_ref, compiled from the reference
                  sources) on this host's cores for the same 64x512 pair.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP32_VALU_PEAK_TFLOPS = 157.3     # MI355X dense FP32 vector peak (counts FMA as 2 flops)
MAXPLUS_CEILING_TFLOPS = 78.6     # one add + one max per lane-slot per clock


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(M, N, seed=7):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 4, M).astype(np.int32), rng.integers(0, 4, N).astype(np.int32)


def cpu_baseline(M, N, s1, s2, threads):
    """Reference bpmax() on the host (kind 'reference'), or the oracle restatement on a
    bounded sample when the reference binary is absent (kind 'port')."""
    from oracle import oracle as O
    cells = None
    import bpmaxcpu_amd as B
    if O.have_reference():
        log(f"cpu baseline: reference bpmax() {M}x{N} on {threads} threads")
        _, score, secs, _ = O.reference_run(s1, s2, table=False, threads=threads)
        cells = B.cell_updates(M, N)
        return {"value": cells / secs / 1e9, "unit": "GCUPS", "cores": threads, "kind": "reference",
                "sample": f"one {M}x{N} pair (the full workload), reference [MPT+RT]:v3 bpmax() built with "
                          f"g++ -O3 -fopenmp -mavx2 -mfma + FTZ, N_tile 48, 3x24 register tile, "
                          f"OMP_NUM_THREADS={threads}; {secs:.2f} s, score {score}"}
    Ms = 16
    log(f"cpu baseline: oracle port on a {Ms}x{N} sample")
    os.environ["OMP_NUM_THREADS"] = str(threads)
    t0 = time.time()
    O.bpmax(s1[:Ms], s2)
    secs = time.time() - t0
    return {"value": B.cell_updates(Ms, N) / secs / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"oracle restatement on the first {Ms} bases of seq1 x all {N} of seq2; {secs:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; pairs are independent, so a multi-GPU run is "
              "pair-level data parallelism with no communication (not measured here).", file=sys.stderr)
        sys.exit(2)

    import bpmaxcpu_amd as B
    M, N = args.M, args.N
    s1, s2 = workload(M, N)
    cells = B.cell_updates(M, N)
    eng = B.Engine(0)
    log(f"warmup {args.warmup} x {M}x{N}")
    score = None
    for _ in range(args.warmup):
        _, score = eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    eng.synchronize()
    log(f"timed {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, score = eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    eng.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    gcups = cells / (ms_per_step / 1e3) / 1e9
    log(f"{ms_per_step:.2f} ms/step, {gcups:.1f} GCUPS, score {score}")

    # instrumented steps: HIP events on the engine stream around every launch
    eng.set_timing(True)
    eng.reset_stats()
    eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    st = eng.kernel_stats()
    tk = st["tasks"]
    achieved = 2.0 * tk["updates"] / (tk["ms"] / 1e3) / 1e12 if tk["ms"] > 0 else 0.0
    roofline = {"bound": "valu", "kernel": "k_tasks<64> (dependency-tile scheduler: every phase of the solve)",
                "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4),
                "frac_of_maxplus_ceiling": round(achieved / MAXPLUS_CEILING_TFLOPS, 4),
                "launches": tk["launches"], "avg_launch_ms": round(tk["ms"] / max(1, tk["launches"]), 4),
                "flops_per_launch": 2.0 * tk["updates"] / max(1, tk["launches"]),
                "traffic": None}
    # the same solve with one launch per phase and diagonal, to isolate the R0 phase
    eng.set_schedule(B.SCHED_LAUNCH)
    eng.reset_stats()
    eng.solve(s1, s2, layout=B.LAYOUT_NONE)
    st2 = eng.kernel_stats()
    eng.set_schedule(B.SCHED_TASKS)
    eng.set_timing(False)
    o = st2["outer"]
    r0 = 2.0 * o["updates"] / (o["ms"] / 1e3) / 1e12 if o["ms"] > 0 else 0.0
    roofline["r0_phase"] = {"kernel": "k_outer<64> (R0 + init/P1/R3/R4, one launch per outer diagonal)",
                            "achieved": round(r0, 3), "frac": round(r0 / FP32_VALU_PEAK_TFLOPS, 4),
                            "frac_of_maxplus_ceiling": round(r0 / MAXPLUS_CEILING_TFLOPS, 4),
                            "launches": o["launches"], "ms": round(o["ms"], 3)}
    phases = {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in st2.items() if v["launches"]}
    log(f"k_tasks {tk['ms']:.2f} ms ({achieved:.1f} TFLOP/s); per-phase launches: {phases}")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1, 128)
        cpu = cpu_baseline(M, N, s1, s2, threads)
        log(f"cpu baseline {cpu['value']:.2f} GCUPS on {threads} threads")

    line = {"metric": "BPMax GCUPS at M=64,N=512", "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: uniform i.i.d. ACGU from seed 7, reference weight table (GC=3, AU=1, GU=1.5)",
            "config": {"workload": f"plain BPMax, one pair M={M} N={N}, score-only output, F resident in HBM",
                       "M": M, "N": N, "pairs_per_step": 1, "score": score},
            "roofline": roofline, "cpu_baseline": cpu, "launch_schedule_phases_ms": phases}
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()

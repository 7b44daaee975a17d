"""Autotuner front end: tune the launch schedule of problem classes on this GPU and save
the table as JSON, or load such a table into an engine.

    python -m bpmaxcpu_amd.autotune --out tuning.json 64x512 48x400x16

A shape is MxN (one pair) or MxNxP (batches of P pairs; the class covers every batch whose
largest M and N fall in the same power-of-two buckets).  Each entry records the settings
bpmax_tune_shape picked and its median time.
"""
import argparse
import json

from . import Engine


def parse_shape(text: str):
    parts = [int(x) for x in text.lower().split("x")]
    if len(parts) == 2:
        parts.append(1)
    if len(parts) != 3 or min(parts) < 1:
        raise ValueError(f"shape {text!r}: expected MxN or MxNxP")
    return tuple(parts)


def tune_shapes(eng: Engine, shapes, reps: int = 2) -> list:
    table = []
    for M, N, P in shapes:
        best = eng.tune(M, N, P, reps)
        table.append({"M": M, "N": N, "pairs": P, **best})
    return table


def load_table(eng: Engine, path: str) -> list:
    with open(path) as fh:
        table = json.load(fh)
    for e in table:
        eng.set_tuning(e["M"], e["N"], e["pairs"], e)
    return table


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("shapes", nargs="+", help="MxN or MxNxP")
    ap.add_argument("--out", required=True)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()
    with Engine(0) as eng:
        table = tune_shapes(eng, [parse_shape(s) for s in args.shapes], args.reps)
    with open(args.out, "w") as fh:
        json.dump(table, fh, indent=1)
    for e in table:
        print(e)


if __name__ == "__main__":
    main()

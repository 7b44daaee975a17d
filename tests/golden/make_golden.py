"""Writes the golden vectors under tests/golden/ from the reference's own bpmax().

The reference [MPT+RT]:v3 code is compiled from its sources by ``make -C oracle ref``
(g++ + FTZ/DAZ, see oracle/Makefile) and driven through oracle/ref_harness.cpp, which
supplies the weight table through the reference's external-function plug point and
canonicalizes F to the packed tri-tri order of the CHECKING dump
(src/bpmax-wrapper.c:318-342).  Sequences come from a seeded numpy generator and are
stored in each fixture, so nothing depends on a PRNG matching across languages.

Fixtures:
  full_<M>x<N>_<table>.npz   every defined cell (small shapes)
  hash_<M>x<N>_<table>.json  score, SHA-256 of the packed float32 table, strided samples

Usage:  python tests/golden/make_golden.py [--verify-oracle]
  --verify-oracle also runs the C restatement (oracle/bpmax_oracle.c) on every fixture and
  requires bit-identical tables (the 64x512 case takes about a minute on 8 cores).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

TABLES = {"ref": lambda: (O.reference_table(),) * 2, "321": lambda: (O.table_321(),) * 2,
          "asym": O.table_asym}

# (M, N, seed): small full tables, then shapes of the reference's own sweep
# (src/test.sh: N = 192..196, 384..389, 480 around N_tile 48/96/192), then the
# benchmark shape.
FULL = [(5, 17, 1), (16, 100, 1234), (1, 9, 5), (2, 33, 6), (4, 64, 7), (3, 65, 8)]
HASHED = [(6, 193, 21), (18, 196, 22), (7, 480, 23), (10, 385, 24), (33, 129, 25), (64, 512, 7)]


def seqs(M, N, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 4, M).astype(np.int32), rng.integers(0, 4, N).astype(np.int32)


def digits(a):
    return "".join(str(int(x)) for x in a)


def sha(F):
    return hashlib.sha256(np.ascontiguousarray(F, "<f4").tobytes()).hexdigest()


def samples(F, M, N, n=64, seed=99):
    rng = np.random.default_rng(seed)
    out = []
    # always include the score cell and the (0,0) corner cells
    cells = [(0, M - 1, 0, N - 1), (0, 0, 0, 0), (M - 1, M - 1, N - 1, N - 1)]
    while len(cells) < n:
        i1 = int(rng.integers(0, M)); j1 = int(rng.integers(i1, M))
        i2 = int(rng.integers(0, N)); j2 = int(rng.integers(i2, N))
        cells.append((i1, j1, i2, j2))
    for c in cells:
        out.append([*c, float(F[O.packed_index(*c, M, N)])])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--verify-oracle", action="store_true")
    ap.add_argument("--tables", default=",".join(TABLES), help="comma-separated subset of " + ",".join(TABLES))
    args = ap.parse_args()
    if not O.have_reference():
        sys.exit("oracle/_ref/bpmax_ref_run missing: run `make -C oracle ref` first")
    for (M, N, seed) in FULL + HASHED:
        s1, s2 = seqs(M, N, seed)
        for tname in args.tables.split(","):
            if tname == "asym" and (M, N) == (64, 512):
                continue
            intra, inter = TABLES[tname]()
            F, score, secs, ntile = O.reference_run(s1, s2, intra, inter)
            if args.verify_oracle:
                Fo, so = O.bpmax(s1, s2, intra, inter)
                assert np.array_equal(Fo.view(np.uint32), F.view(np.uint32)), (M, N, tname)
            meta = dict(M=M, N=N, seed=seed, table=tname, seq1=digits(s1), seq2=digits(s2),
                        intra=intra.reshape(-1).tolist(), inter=inter.reshape(-1).tolist(),
                        score=score, n_tile=ntile, sha256=sha(F),
                        source="reference bpmax() (src/bpmax.c:129) via oracle/ref_harness.cpp")
            if (M, N, seed) in FULL:
                path = os.path.join(HERE, f"full_{M}x{N}_{tname}.npz")
                np.savez_compressed(path, seq1=s1, seq2=s2, intra=intra, inter=inter, F=F,
                                    score=np.float32(score), meta=json.dumps(meta))
            else:
                meta["samples"] = samples(F, M, N)
                path = os.path.join(HERE, f"hash_{M}x{N}_{tname}.json")
                with open(path, "w") as f:
                    json.dump(meta, f, indent=1)
            print(f"{os.path.basename(path)}: score={score} ref_time={secs:.3f}s n_tile={ntile}", flush=True)


if __name__ == "__main__":
    main()

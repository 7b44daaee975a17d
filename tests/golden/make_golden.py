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
  set_<name>_<table>.json    a list of pairs drawn from one workload's distribution, each
                             with its score, SHA-256 and samples:
                               batch4  16 pairs, M~U[16,48], N~U[100,400], per-sequence
                                       GC~U[0.3,0.7] (the batched workload);
                               serve5  10 pairs, M log-uniform [8,64], N log-uniform
                                       [64,1024], GC 0.5 (the serving workload), with its
                                       corner shapes forced in (8x1024, 12x800, 40x300,
                                       48x700, 64x1024, 56x96, 8x64)

Usage:  python tests/golden/make_golden.py [--verify-oracle] [--sets batch4,serve5] [--no-single]
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


def gc_seq(rng, n, gc):
    """i.i.d. bases with P(G)=P(C)=gc/2, P(A)=P(U)=(1-gc)/2 (codes A0 C1 G2 U3)."""
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return rng.choice(4, size=n, p=p).astype(np.int32)


def set_pairs(name):
    """The pairs of fixture set `name` (sequences drawn here, stored in the fixture)."""
    if name == "batch4":
        rng = np.random.default_rng(404)
        out = []
        for _ in range(16):
            M = int(rng.integers(16, 49))
            N = int(rng.integers(100, 401))
            out.append((gc_seq(rng, M, rng.uniform(0.3, 0.7)), gc_seq(rng, N, rng.uniform(0.3, 0.7))))
        return out
    if name == "serve5":
        rng = np.random.default_rng(505)
        shapes = [(8, 1024), (12, 800), (40, 300), (48, 700), (64, 1024), (56, 96), (8, 64)]
        for _ in range(3):
            shapes.append((int(round(np.exp(rng.uniform(np.log(8), np.log(64))))),
                           int(round(np.exp(rng.uniform(np.log(64), np.log(1024)))))))
        return [(gc_seq(rng, M, 0.5), gc_seq(rng, N, 0.5)) for M, N in shapes]
    raise ValueError(name)


# per set: the tables written, and pairs left out of a table (reference CPU time)
SET_TABLES = {"batch4": ("ref", "321"), "serve5": ("ref", "321")}
SET_SKIP = {("serve5", "321"): {(64, 1024)}}


def write_set(name, tables, verify):
    for tname in tables:
        if tname not in SET_TABLES[name]:
            continue
        intra, inter = TABLES[tname]()
        entries = []
        for k, (s1, s2) in enumerate(set_pairs(name)):
            M, N = len(s1), len(s2)
            if (M, N) in SET_SKIP.get((name, tname), ()):
                continue
            F, score, secs, ntile = O.reference_run(s1, s2, intra, inter)
            if verify and M * N <= 48 * 400:
                Fo, so = O.bpmax(s1, s2, intra, inter)
                assert np.array_equal(Fo.view(np.uint32), F.view(np.uint32)), (name, k, tname)
            entries.append(dict(index=k, M=M, N=N, seq1=digits(s1), seq2=digits(s2), score=score,
                                n_tile=ntile, sha256=sha(F), samples=samples(F, M, N, n=16)))
            print(f"set {name} {tname} pair {k} {M}x{N}: score={score} ref_time={secs:.2f}s", flush=True)
            del F
        path = os.path.join(HERE, f"set_{name}_{tname}.json")
        with open(path, "w") as f:
            json.dump(dict(set=name, table=tname, intra=intra.reshape(-1).tolist(),
                           inter=inter.reshape(-1).tolist(), pairs=entries,
                           source="reference bpmax() (src/bpmax.c:129) via oracle/ref_harness.cpp"), f, indent=1)
        print(f"{os.path.basename(path)}: {len(entries)} pairs", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--verify-oracle", action="store_true")
    ap.add_argument("--tables", default=",".join(TABLES), help="comma-separated subset of " + ",".join(TABLES))
    ap.add_argument("--sets", default="", help="comma-separated fixture sets to write (batch4, serve5)")
    ap.add_argument("--no-single", action="store_true", help="skip the full_* and hash_* fixtures")
    args = ap.parse_args()
    if not O.have_reference():
        sys.exit("oracle/_ref/bpmax_ref_run missing: run `make -C oracle ref` first")
    for name in filter(None, args.sets.split(",")):
        write_set(name, args.tables.split(","), args.verify_oracle)
    if args.no_single:
        return
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

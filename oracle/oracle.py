"""CPU checker for the BPMax engine.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this
module.  It wraps

* ``libbpmax_oracle.so`` -- the C restatement of the reference recurrence
  (oracle/bpmax_oracle.c, following compilation_scripts/bpmax_verify.ab:16-71), and
* ``_ref/bpmax_ref_run`` -- the reference's own ``bpmax()`` (src/bpmax.c:129) compiled
  from its sources by oracle/Makefile, driven through oracle/ref_harness.cpp.

Weight tables are 5x5 float32 arrays indexed [code_x, code_y] with codes
0..4 = A, C, G, U/T, N (src/external_functions.c:74-85).
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libbpmax_oracle.so")
REF_RUN = os.path.join(HERE, "_ref", "bpmax_ref_run")
REF_CLI = os.path.join(HERE, "_ref", "bpmax")
REF_CHECK = os.path.join(HERE, "_ref", "bpmax.check")
# the reference's CHECKING driver linked against libbpmax_hip.so (make -C oracle dropin)
REF_DROPIN_CHECK = os.path.join(HERE, "_ref", "bpmax_dropin.check")


def _pair_table(au: float, gc: float, gu: float) -> np.ndarray:
    t = np.zeros((5, 5), np.float32)
    A, C, G, U = 0, 1, 2, 3
    t[A, U] = t[U, A] = au
    t[C, G] = t[G, C] = gc
    t[G, U] = t[U, G] = gu
    return t


def reference_table() -> np.ndarray:
    """The table hard-coded in the reference: GC=3, AU=1, GU=1.5, the same for intra- and
    inter-strand pairs (src/external_functions.c:21-68)."""
    return _pair_table(1.0, 3.0, 1.5)


def table_321() -> np.ndarray:
    """GC=3, AU=2, GU=1 (the integer table the benchmark description names)."""
    return _pair_table(2.0, 3.0, 1.0)


def table_asym():
    """(intra, inter): deliberately asymmetric tables (x,y) != (y,x), with non-Watson-Crick
    entries, so argument order and strand orientation are pinned by the fixtures."""
    intra = np.zeros((5, 5), np.float32)
    inter = np.zeros((5, 5), np.float32)
    A, C, G, U = 0, 1, 2, 3
    intra[A, U], intra[U, A], intra[C, G], intra[G, C], intra[G, U], intra[U, G] = 1.0, 2.0, 3.0, 2.5, 1.5, 0.5
    intra[A, A] = 0.25
    inter[A, U], inter[U, A], inter[C, G], inter[G, C], inter[G, U], inter[U, G] = 2.0, 1.0, 3.5, 3.0, 1.0, 1.5
    inter[C, C] = 0.75
    return intra, inter


MUT = dict(NO_P1=1, NO_P2=2, NO_R0=4, R0_NOSHIFT=8, NO_R1=16, NO_R2=32, NO_R3=64, NO_R4=128,
           INTER_TRANSPOSED=256, SEQ2_UNREVERSED=512, HAIRPIN_2=1024)


def tri(n: int) -> int:
    return n * (n + 1) // 2


def utri(i, j, n):
    """Row-major upper-triangle index of (i, j), i <= j < n."""
    return i * n - i * (i - 1) // 2 + (j - i)


def packed_index(i1, j1, i2, j2, M, N):
    return utri(i1, j1, M) * tri(N) + utri(i2, j2, N)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (run `make -C oracle oracle`)")
        L = ctypes.CDLL(ORACLE_SO)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_bpmax.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, fp, fp, fp, fp]
        L.oracle_bpmax.restype = ctypes.c_int
        L.oracle_bpmax_mut.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, fp, fp, fp, fp, ctypes.c_int]
        L.oracle_bpmax_mut.restype = ctypes.c_int
        L.oracle_single_strand.argtypes = [ctypes.c_int, ip, fp, fp]
        L.oracle_single_strand.restype = None
        L.oracle_maxplus_tri.argtypes = [ctypes.c_int, fp, fp, fp]
        L.oracle_maxplus_tri.restype = None
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def bpmax(seq1, seq2, intra=None, inter=None, mutation=0):
    """Oracle F table (packed tri-tri, float32) and optimal score F(0,M-1,0,N-1).

    ``mutation`` (an OR of MUT values) plants a mistake, for tests of the checks."""
    s1 = np.ascontiguousarray(seq1, np.int32)
    s2 = np.ascontiguousarray(seq2, np.int32)
    intra = np.ascontiguousarray(reference_table() if intra is None else intra, np.float32)
    inter = np.ascontiguousarray(intra if inter is None else inter, np.float32)
    M, N = len(s1), len(s2)
    F = np.empty(tri(M) * tri(N), np.float32)
    score = ctypes.c_float()
    rc = lib().oracle_bpmax_mut(M, N, _ip(s1), _ip(s2), _fp(intra), _fp(inter), _fp(F),
                                ctypes.byref(score), int(mutation))
    if rc != 0:
        raise ValueError("oracle_bpmax rejected its arguments")
    return F, float(score.value)


def single_strand(x, intra=None):
    """Dense n*n single-strand table (upper triangle; -inf below)."""
    x = np.ascontiguousarray(x, np.int32)
    intra = np.ascontiguousarray(reference_table() if intra is None else intra, np.float32)
    S = np.empty((len(x), len(x)), np.float32)
    lib().oracle_single_strand(len(x), _ip(x), _fp(intra), _fp(S))
    return S


def maxplus_tri(A, B, C):
    """C[i][j] = max(C[i][j], max_{i<=k<=j} A[i][k] + B[k][j]) on one triangular n*n pair."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    C = np.array(C, np.float32, copy=True, order="C")
    lib().oracle_maxplus_tri(A.shape[0], _fp(A), _fp(B), _fp(C))
    return C


# ----------------------------------------------------------------- reference binary
def have_reference() -> bool:
    return os.path.exists(REF_RUN)


def write_input(path, seq1, seq2, intra, inter):
    s1 = np.ascontiguousarray(seq1, np.int32)
    s2 = np.ascontiguousarray(seq2, np.int32)
    with open(path, "wb") as f:
        f.write(struct.pack("<iii", 0x42504D58, len(s1), len(s2)))
        f.write(s1.tobytes())
        f.write(s2.tobytes())
        f.write(np.ascontiguousarray(intra, np.float32).tobytes())
        f.write(np.ascontiguousarray(inter, np.float32).tobytes())


def reference_run(seq1, seq2, intra=None, inter=None, table=True, repeat=1, threads=None,
                  workdir=None, warmup=0, env_extra=None):
    """Run the reference's bpmax() through the harness: `warmup` untimed runs, then the
    median of `repeat` timed ones.  env_extra adds environment variables (OMP_PROC_BIND ...).

    Returns (F_packed or None, score, median_seconds, n_tile)."""
    if not have_reference():
        raise RuntimeError(f"{REF_RUN} is not built (run `make -C oracle ref`)")
    intra = reference_table() if intra is None else intra
    inter = intra if inter is None else inter
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    env.update(env_extra or {})
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_input(fin, seq1, seq2, intra, inter)
        args = [REF_RUN, fin, fout, "--repeat", str(repeat), "--warmup", str(warmup)]
        if not table:
            args.append("--no-table")
        subprocess.run(args, check=True, env=env, stdout=subprocess.DEVNULL)
        with open(fout, "rb") as f:
            score, secs, _rep, ntile = struct.unpack("<fdii", f.read(20))
            F = np.fromfile(f, np.float32) if table else None
    return F, float(score), float(secs), int(ntile)

"""Loading helpers for the golden vectors under tests/golden/ (written by make_golden.py
from the reference's own bpmax())."""
import glob
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def full_fixtures():
    return sorted(glob.glob(os.path.join(GOLDEN, "full_*.npz")))


def hash_fixtures():
    return sorted(glob.glob(os.path.join(GOLDEN, "hash_*.json")))


def load_full(path):
    z = np.load(path, allow_pickle=False)
    return dict(seq1=z["seq1"], seq2=z["seq2"], intra=z["intra"], inter=z["inter"], F=z["F"],
                score=float(z["score"]), meta=json.loads(str(z["meta"])))


def load_hash(path):
    with open(path) as f:
        m = json.load(f)
    m["seq1"] = np.array([int(c) for c in m["seq1"]], np.int32)
    m["seq2"] = np.array([int(c) for c in m["seq2"]], np.int32)
    m["intra"] = np.array(m["intra"], np.float32).reshape(5, 5)
    m["inter"] = np.array(m["inter"], np.float32).reshape(5, 5)
    return m


def sha256_f32(F):
    return hashlib.sha256(np.ascontiguousarray(F, "<f4").tobytes()).hexdigest()


def fid(path):
    return os.path.basename(path).rsplit(".", 1)[0]


def load_set(name):
    """A fixture set (set_<workload>_<table>.json): the table and a list of pairs, each
    with seq1/seq2 as code arrays, score, sha256 and samples."""
    with open(os.path.join(GOLDEN, name)) as f:
        s = json.load(f)
    s["intra"] = np.array(s["intra"], np.float32).reshape(5, 5)
    s["inter"] = np.array(s["inter"], np.float32).reshape(5, 5)
    for p in s["pairs"]:
        p["seq1"] = np.array([int(c) for c in p["seq1"]], np.int32)
        p["seq2"] = np.array([int(c) for c in p["seq2"]], np.int32)
    return s


def set_fixtures(workload):
    return sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, f"set_{workload}_*.json")))

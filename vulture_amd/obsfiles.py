"""Readers for Vulture observer files (observer.c:760-995, 1137-1198)."""
import glob
import os

import numpy as np


def read_td(path):
    """eh_*_td.asc / wf_*_td.asc -> (steps, times, values[n, ncomp])."""
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    a = np.array(rows, dtype=np.float64).reshape(len(rows), -1) if rows else np.zeros((0, 2))
    return a[:, 0].astype(np.int64), a[:, 1], a[:, 2:]


def read_fd(path):
    """eh_*_fd.asc / wf_*_fd.asc -> (freqs, complex values[n, ncomp])."""
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    a = np.array(rows, dtype=np.float64).reshape(len(rows), -1) if rows else np.zeros((0, 1))
    v = a[:, 1:]
    return a[:, 0], v[:, 0::2] + 1j * v[:, 1::2]


def read_impulse(path):
    """impulse.dat -> (numOutTimeSteps, dt, records[int ijk(3), float fields(6)])."""
    raw = open(path, "rb").read()
    n = np.frombuffer(raw[:4], dtype=np.int32)[0]
    dt = np.frombuffer(raw[4:8], dtype=np.float32)[0]
    rec = np.dtype([("ijk", np.int32, 3), ("f", np.float32, 6)])
    return int(n), float(dt), np.frombuffer(raw[8:], dtype=rec)


def observer_files(outdir, pattern="eh_*_td.asc"):
    return sorted(glob.glob(os.path.join(outdir, pattern)))


def rel_l2(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / d) if d > 0 else float(np.linalg.norm(x - ref))

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


def waveform_spectrum(outdir):
    """Raw waveform DFT from the single wf_*_fd.asc in `outdir` (None if absent or ambiguous)."""
    files = observer_files(outdir, "wf_*_fd.asc")
    if len(files) != 1:
        return None
    return read_fd(files[0])[1][:, 0]


def compare_outputs(test_dir, ref_dir, band_db=60.0):
    """Compare every eh_* observer file of `test_dir` with `ref_dir`.

    Returns a dict with the worst time-series relative L2 (`td`), the worst
    relative L2 of the raw field DFTs (`fd_raw`: the normalised spectra
    multiplied back by the waveform DFT, so that frequencies where the
    excitation has no energy do not amplify rounding noise), and the worst
    magnitude error in dB over the band where the waveform spectrum lies
    within `band_db` of its peak (`fd_db`).  `fd_db` leaves out bins where the
    reference magnitude is below 1e-3 of its in-band maximum (the nulls of a
    cavity or standing-wave spectrum, where a dB ratio measures rounding
    noise); `fd_db_all` is the same maximum over every in-band bin with a
    non-zero reference.  Missing
    or misshapen files, and any non-finite value in a tested file, give
    infinity.
    """
    res = {"td": 0.0, "fd_raw": 0.0, "fd_db": 0.0, "fd_db_all": 0.0, "files": 0}
    inf = float("inf")

    def worst(key, m):
        # max() drops NaN (max(0.0, nan) == 0.0): a non-finite metric fails
        res[key] = inf if not np.isfinite(m) else max(res[key], m)
    for f in observer_files(ref_dir, "eh_*_td.asc"):
        g = os.path.join(test_dir, os.path.basename(f))
        if not os.path.exists(g):
            res["td"] = inf
            continue
        _, _, va = read_td(g)
        _, _, vb = read_td(f)
        finite = bool(np.all(np.isfinite(va)))
        worst("td", rel_l2(va, vb) if va.shape == vb.shape and finite else inf)
        res["files"] += 1
    wf = waveform_spectrum(ref_dir)
    for f in observer_files(ref_dir, "eh_*_fd.asc"):
        g = os.path.join(test_dir, os.path.basename(f))
        if not os.path.exists(g):
            res["fd_raw"] = inf
            continue
        _, va = read_fd(g)
        _, vb = read_fd(f)
        res["files"] += 1
        if va.shape != vb.shape or not np.all(np.isfinite(va)):
            res["fd_raw"] = res["fd_db"] = res["fd_db_all"] = inf
            continue
        w = wf[:, None] if wf is not None and len(wf) == len(vb) else np.ones((len(vb), 1))
        worst("fd_raw", float(np.linalg.norm(va * w - vb * w) / max(np.linalg.norm(vb * w), 1e-300)))
        band = np.abs(w[:, 0]) >= np.abs(w[:, 0]).max() * 10 ** (-band_db / 20) if wf is not None else \
            np.ones(len(vb), bool)
        ma, mb = np.abs(va[band]), np.abs(vb[band])
        floor = mb.max() * 1e-3 if mb.size else 0.0
        sel = mb > floor
        with np.errstate(divide="ignore", invalid="ignore"):
            if sel.any():
                worst("fd_db", float(np.abs(20 * np.log10(ma[sel] / mb[sel])).max()))
            nz = mb > 0
            if nz.any():
                worst("fd_db_all", float(np.abs(20 * np.log10(ma[nz] / mb[nz])).max()))
    return res

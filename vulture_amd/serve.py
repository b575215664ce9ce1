"""Serving many solves on one GPU: ctypes front end of the C solve server
(vulture_amd/host/vh_serve.c) and the serving-sweep driver.

    with Server(precision="fp32", max_active=8) as sv:
        rid = sv.submit("a.mesh", "out_a")
        info = sv.wait(rid)

`sweep()` generates a seeded stream of solve requests (meshgen.serving_request),
measures each request's isolated service time (submitted alone to an idle
server), replays the stream with Poisson arrivals at a fraction of the
isolated capacity (capacity = 1 / mean isolated service time) for p50/p99
latency, queues the whole stream at once for the sustained throughput, and
checks that every request's observer files are byte-identical alone and in
the mix.
"""
import ctypes
import filecmp
import math
import os
import time

import numpy as np

from . import VgpuError, load_host, meshgen


class RequestInfo(ctypes.Structure):
    """vh_request_info (vulture_amd/host/vh_serve.h)."""
    _fields_ = [("id", ctypes.c_int), ("status", ctypes.c_int), ("t_submit", ctypes.c_double),
                ("t_ready", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_gpu_done", ctypes.c_double),
                ("t_done", ctypes.c_double), ("gpu_ms", ctypes.c_double), ("cells", ctypes.c_longlong),
                ("steps", ctypes.c_ulong), ("bytes", ctypes.c_size_t), ("err", ctypes.c_char * 256)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "err"}
        d["err"] = self.err.decode(errors="replace")
        return d


def _bind(h):
    if getattr(h, "_serve_bound", False):
        return h
    h.vh_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_ulong]
    h.vh_server_create.restype = ctypes.c_void_p
    h.vh_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    h.vh_server_submit.restype = ctypes.c_int
    h.vh_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(RequestInfo)]
    h.vh_server_wait.restype = ctypes.c_int
    h.vh_server_drain.argtypes = [ctypes.c_void_p]
    h.vh_server_drain.restype = ctypes.c_int
    h.vh_server_now.argtypes = [ctypes.c_void_p]
    h.vh_server_now.restype = ctypes.c_double
    h.vh_server_destroy.argtypes = [ctypes.c_void_p]
    h.vh_server_destroy.restype = None
    h._serve_bound = True
    return h


class Server:
    """A solve server on one device: many solves resident and in flight."""

    def __init__(self, precision="fp32", device=0, budget_gb=200.0, max_active=8, workers=8, chunk=200,
                 full_precision=True):
        self._h = _bind(load_host())
        self._p = self._h.vh_server_create(device, 1 if precision == "fp32" else 0, 1 if full_precision else 0,
                                           budget_gb * 1e9, max_active, workers, chunk)
        if not self._p:
            raise VgpuError("could not create the solve server")

    def submit(self, mesh, outdir):
        os.makedirs(outdir, exist_ok=True)
        rid = self._h.vh_server_submit(self._p, os.fsencode(mesh), os.fsencode(outdir))
        if rid < 0:
            raise VgpuError("submit failed")
        return rid

    def wait(self, rid):
        info = RequestInfo()
        self._h.vh_server_wait(self._p, rid, ctypes.byref(info))
        return info.as_dict()

    def now(self):
        return self._h.vh_server_now(self._p)

    def close(self):
        if self._p:
            self._h.vh_server_destroy(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def make_requests(outdir, count, seed=5, **kw):
    """Write `count` seeded serving requests; returns [(mesh path, cells, steps)]."""
    rng = np.random.default_rng(seed)
    os.makedirs(outdir, exist_ok=True)
    out = []
    for q in range(count):
        text, n, steps = meshgen.serving_request(rng, **kw)
        p = os.path.join(outdir, f"req{q:04d}.mesh")
        with open(p, "w") as f:
            f.write(text)
        out.append((p, n, steps))
    return out


def _same_outputs(a, b):
    fa = sorted(f for f in os.listdir(a) if f.endswith(".asc"))
    fb = sorted(f for f in os.listdir(b) if f.endswith(".asc"))
    return fa == fb and len(fa) > 0 and all(filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False)
                                            for f in fa)


def sweep(workdir, count=40, load=0.7, seed=5, precision="fp32", max_active=8, chunk=200, log=None, workers=8,
          **kw):
    """Isolated pass, then the Poisson-arrival mixed pass; returns a report dict."""
    reqs = make_requests(os.path.join(workdir, "meshes"), count, seed=seed, **kw)
    rng = np.random.default_rng(seed + 1000)
    iso = []
    with Server(precision=precision, max_active=1, chunk=chunk) as sv:
        for q, (mesh, _, _) in enumerate(reqs):
            info = sv.wait(sv.submit(mesh, os.path.join(workdir, "alone", f"{q:04d}")))
            if info["status"] != 1:
                raise VgpuError(f"request {q} failed alone: {info['err']}")
            iso.append(info)
            if log and q % 8 == 0:
                log(f"alone {q}/{count}")
    service = np.array([i["t_done"] - i["t_submit"] for i in iso])
    work = np.array([i["cells"] * i["steps"] for i in iso], dtype=np.float64)
    rate = load / service.mean()                      # arrivals per second
    gaps = rng.exponential(1.0 / rate, size=count)
    mix = []
    with Server(precision=precision, max_active=max_active, chunk=chunk, workers=workers) as sv:
        t0 = time.monotonic()
        arrive = np.cumsum(gaps) - gaps[0]
        ids = []
        for q, (mesh, _, _) in enumerate(reqs):
            dt = arrive[q] - (time.monotonic() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(sv.submit(mesh, os.path.join(workdir, "mix", f"{q:04d}")))
            if log and q % 8 == 0:
                log(f"mixed: submitted {q}/{count}")
        for q, rid in enumerate(ids):
            info = sv.wait(rid)
            if info["status"] != 1:
                raise VgpuError(f"request {q} failed in the mix: {info['err']}")
            mix.append(info)
            if log and q % 8 == 0:
                log(f"mixed {q}/{count}")
        span = max(i["t_done"] for i in mix) - min(i["t_submit"] for i in mix)
    # saturation: every request queued at once, for the sustained throughput
    with Server(precision=precision, max_active=max_active, chunk=chunk, workers=workers) as sv:
        ids = [sv.submit(mesh, os.path.join(workdir, "sat", f"{q:04d}")) for q, (mesh, _, _) in enumerate(reqs)]
        sat = []
        for q, rid in enumerate(ids):
            sat.append(sv.wait(rid))
            if log and q % 8 == 0:
                log(f"saturated {q}/{count}")
        sat_span = max(i["t_done"] for i in sat) - min(i["t_submit"] for i in sat)
    lat = np.array([i["t_done"] - i["t_submit"] for i in mix])
    # where the mixed pass's time goes: waiting for and doing the host set-up,
    # waiting for admission, on the device, writing the files
    phase = {"setup": [i["t_ready"] - i["t_submit"] for i in mix],
             "admission": [i["t_admit"] - i["t_ready"] for i in mix],
             "device": [i["t_gpu_done"] - i["t_admit"] for i in mix],
             "write": [i["t_done"] - i["t_gpu_done"] for i in mix]}
    phases = {k: {"p50_s": float(np.percentile(v, 50)), "p99_s": float(np.percentile(v, 99))}
              for k, v in phase.items()}
    # device efficiency of each request run alone (its loop's device time),
    # by size: how much of the stream's work small grids hold, and how well
    # the device runs them
    gpu_s = np.array([i["gpu_ms"] for i in iso]) / 1e3
    cells = np.array([i["cells"] for i in iso], dtype=np.float64)
    by_size = []
    for lo, hi in ((0, 5e5), (5e5, 4e6), (4e6, np.inf)):
        m = (cells >= lo) & (cells < hi)
        if m.any():
            by_size.append({"cells": f"[{lo:.0e}, {hi:.0e})", "requests": int(m.sum()),
                            "work_share": float(work[m].sum() / work.sum()),
                            "device_mcells_per_s": float(work[m].sum() / gpu_s[m].sum() / 1e6)})
    ratio = lat / service
    same = [_same_outputs(os.path.join(workdir, "alone", f"{q:04d}"), os.path.join(workdir, "mix", f"{q:04d}")) and
            _same_outputs(os.path.join(workdir, "alone", f"{q:04d}"), os.path.join(workdir, "sat", f"{q:04d}"))
            for q in range(count)]
    return {
        "requests": count, "load": load, "arrival_rate_per_s": rate, "max_active": max_active,
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "p50_isolated_s": float(np.percentile(service, 50)), "p99_isolated_s": float(np.percentile(service, 99)),
        "p99_latency_over_isolated": float(np.percentile(ratio, 99)),
        "p50_latency_over_isolated": float(np.percentile(ratio, 50)),
        "mixed_phases": phases,
        "worst": [{"request": int(q), "cells": int(iso[q]["cells"]), "steps": int(iso[q]["steps"]),
                   "isolated_s": float(service[q]), "latency_s": float(lat[q]),
                   **{k: float(v[q]) for k, v in phase.items()}}
                  for q in np.argsort(-ratio)[:3]],
        "workers": workers,
        "offered_mcells_per_s": float(work.sum() / span / 1e6),
        "isolated_mcells_per_s": float(work.sum() / service.sum() / 1e6),
        "isolated_device_mcells_per_s": float(work.sum() / gpu_s.sum() / 1e6),
        "isolated_by_size": by_size,
        "saturated_mcells_per_s": float(work.sum() / sat_span / 1e6),
        "saturated_over_isolated": float(service.sum() / sat_span),
        "alone_equals_mix": bool(all(same)), "mismatched": [q for q, s in enumerate(same) if not s],
        "makespan_s": float(span),
    }

"""rocprofv3 helpers for the roofline figures.

  python -m vulture_amd.profiling copy [--mib N]
      A plain elementwise copy of N MiB (default 1024), five times: the
      calibration kernel for the FETCH_SIZE / WRITE_SIZE counters (it reads
      and writes exactly N MiB per dispatch).  Then twenty more, timed with
      events, for the bandwidth a plain streaming kernel reaches (read plus
      write bytes per second).
  python -m vulture_amd.profiling stats DIR
      Print the kernel_stats.csv that `rocprofv3 --kernel-trace --stats -d DIR`
      wrote, as a text table.
  python -m vulture_amd.profiling requests --kernel NAME --pmc DIR --copy DIR [--mib N]
      From `rocprofv3 --pmc TCC_EA0_RDREQ_sum TCC_EA0_WRREQ_sum -d DIR` runs of
      the workload and of the copy kernel, print the L2-to-memory bytes per
      launch of kernel NAME, each request counter scaled by the bytes the copy
      moves per request.
  python -m vulture_amd.profiling traffic --kernel NAME --fetch DIR --write DIR
          --copy-fetch DIR --copy-write DIR [--mib N]
      From four `rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -d DIR` runs (the
      workload and the copy kernel), print the HBM bytes per launch of
      kernel NAME: each counter (KiB per dispatch) scaled by what it reads
      for the copy of known size.
"""
import argparse
import csv
import glob
import os
import re
import sys


def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _is_kernel(name, kernel):
    """The dispatch's kernel is `kernel` itself (any template arguments), not
    another kernel whose name starts with it (k_corr, not k_corr_gather)."""
    return re.search(r"(^|[\s:])" + re.escape(kernel) + r"[<(]", name) is not None


def counter_per_dispatch(d, kernel, counter, exact=True):
    """Mean value of `counter` over the dispatches of kernel `kernel` (with
    exact=False, of every kernel whose name contains `kernel`)."""
    per = {}
    for r in _rows(d, "counter_collection.csv"):
        name = r.get("Kernel_Name", "")
        if not (_is_kernel(name, kernel) if exact else kernel in name) or r.get("Counter_Name") != counter:
            continue
        key = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"no {counter} rows for kernels matching {kernel!r} under {d}")
    return sum(per.values()) / len(per), len(per)


def cmd_copy(args):
    import torch
    n = args.mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)  # one elementwise kernel: reads and writes N MiB
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(20):
        torch.mul(a, 1.0, out=b)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / 20
    print(f"copy_ms {ms:.4f} copy_GBps {2 * args.mib * (1 << 20) / (ms * 1e-3) / 1e9:.1f}", flush=True)
    print(f"copied {args.mib} MiB x5", flush=True)


def cmd_stats(args):
    rows = _rows(args.dir, "kernel_stats.csv")
    rows.sort(key=lambda r: -float(r.get("TotalDurationNs", 0)))
    print(f"{'calls':>7} {'avg_us':>10} {'total_ms':>10} {'pct':>6}  kernel")
    for r in rows:
        name = r["Name"]
        if len(name) > 90:
            name = name[:87] + "..."
        print(f"{int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:10.2f} {float(r['TotalDurationNs']) / 1e6:10.3f} "
              f"{float(r['Percentage']):6.2f}  {name}")


def cmd_traffic(args):
    copy_bytes = args.mib * (1 << 20)
    cf, _ = counter_per_dispatch(args.copy_fetch, args.copy_kernel, "FETCH_SIZE", exact=False)
    cw, _ = counter_per_dispatch(args.copy_write, args.copy_kernel, "WRITE_SIZE", exact=False)
    kf, nf = counter_per_dispatch(args.fetch, args.kernel, "FETCH_SIZE")
    kw, nw = counter_per_dispatch(args.write, args.kernel, "WRITE_SIZE")
    ff = copy_bytes / (cf * 1024.0)
    fw = copy_bytes / (cw * 1024.0)
    fetch = kf * 1024.0 * ff
    write = kw * 1024.0 * fw
    print(f"kernel {args.kernel}")
    print(f"dispatches_fetch {nf}")
    print(f"dispatches_write {nw}")
    print(f"copy_FETCH_SIZE_KiB {cf:.1f}")
    print(f"copy_WRITE_SIZE_KiB {cw:.1f}")
    print(f"copy_bytes {copy_bytes}")
    print(f"scale_fetch {ff:.4f}")
    print(f"scale_write {fw:.4f}")
    print(f"FETCH_SIZE_KiB_per_launch {kf:.1f}")
    print(f"WRITE_SIZE_KiB_per_launch {kw:.1f}")
    print(f"fetch_bytes_per_launch {fetch:.0f}")
    print(f"write_bytes_per_launch {write:.0f}")
    print(f"traffic_bytes_per_launch {fetch + write:.0f}")


def cmd_requests(args):
    """L2-to-memory requests per launch (one --pmc pass with both counters),
    each counter turned into bytes by what it reads for the copy of known
    size, which reads and writes exactly N MiB."""
    copy_bytes = args.mib * (1 << 20)
    cr, _ = counter_per_dispatch(args.copy, args.copy_kernel, args.read_counter, exact=False)
    cw, _ = counter_per_dispatch(args.copy, args.copy_kernel, args.write_counter, exact=False)
    kr, n = counter_per_dispatch(args.pmc, args.kernel, args.read_counter)
    kw, _ = counter_per_dispatch(args.pmc, args.kernel, args.write_counter)
    br, bw = copy_bytes / cr, copy_bytes / cw
    print(f"kernel {args.kernel}")
    print(f"dispatches {n}")
    print(f"copy_{args.read_counter} {cr:.0f}")
    print(f"copy_{args.write_counter} {cw:.0f}")
    print(f"bytes_per_read_request {br:.1f}")
    print(f"bytes_per_write_request {bw:.1f}")
    print(f"{args.read_counter}_per_launch {kr:.0f}")
    print(f"{args.write_counter}_per_launch {kw:.0f}")
    print(f"read_bytes_per_launch {kr * br:.0f}")
    print(f"write_bytes_per_launch {kw * bw:.0f}")
    print(f"traffic_bytes_per_launch {kr * br + kw * bw:.0f}")


def main(argv=None):
    ap = argparse.ArgumentParser(description="rocprofv3 helpers")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("copy")
    p.add_argument("--mib", type=int, default=1024)
    p = sub.add_parser("stats")
    p.add_argument("dir")
    p = sub.add_parser("traffic")
    p.add_argument("--kernel", required=True)
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--copy-fetch", required=True)
    p.add_argument("--copy-write", required=True)
    p.add_argument("--copy-kernel", default="elementwise")
    p.add_argument("--mib", type=int, default=1024)
    p = sub.add_parser("requests")
    p.add_argument("--kernel", required=True)
    p.add_argument("--pmc", required=True)
    p.add_argument("--copy", required=True)
    p.add_argument("--copy-kernel", default="elementwise")
    p.add_argument("--read-counter", default="TCC_EA0_RDREQ_sum")
    p.add_argument("--write-counter", default="TCC_EA0_WRREQ_sum")
    p.add_argument("--mib", type=int, default=1024)
    args = ap.parse_args(argv)
    {"copy": cmd_copy, "stats": cmd_stats, "traffic": cmd_traffic, "requests": cmd_requests}[args.cmd](args)


if __name__ == "__main__":
    main(sys.argv[1:])

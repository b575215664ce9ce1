#!/usr/bin/env python3
"""Regenerate the fixtures under tests/golden/ from the reference's own tests.

Reads (never runs) the reference test directories, tests/<name>/:
  ref_inputs/<name>/   the test's mesh and the data files it reads
                       (.pfe, .prm, .dat, .asc waveform tables)
  ref_outputs/<name>/  the reference's validation files eh_*_{td,fd}.asc and
                       wf_*_{td,fd}.asc for one axis permutation per family,
                       gzip-compressed; time series longer than 1000 rows
                       keep every stride-th row (the stride is recorded in
                       the first line) so the fixtures stay small.

Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""
import glob
import gzip
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PERMS = ("_xzy", "_yxz", "_yzx", "_zxy", "_zyx")
INPUT_EXT = (".mesh", ".pfe", ".prm", ".dat")


def main(ref):
    tdir = os.path.join(ref, "tests")
    for d in sorted(glob.glob(os.path.join(tdir, "*", ""))):
        name = os.path.basename(d.rstrip("/"))
        meshes = glob.glob(os.path.join(d, "*.mesh"))
        if not meshes:
            continue
        idir = os.path.join(HERE, "ref_inputs", name)
        os.makedirs(idir, exist_ok=True)
        for f in sorted(os.listdir(d)):
            p = os.path.join(d, f)
            data_asc = f.endswith(".asc") and not (f.startswith("eh_") or f.startswith("wf_"))
            if f.endswith(INPUT_EXT) or data_asc:
                shutil.copyfile(p, os.path.join(idir, f))
        if name.endswith(PERMS):
            continue
        outs = sorted(glob.glob(os.path.join(d, "eh_*_??.asc")) + glob.glob(os.path.join(d, "wf_*_??.asc")))
        if not outs:
            continue
        odir = os.path.join(HERE, "ref_outputs", name)
        os.makedirs(odir, exist_ok=True)
        for p in outs:
            lines = open(p).read().splitlines()
            head = [l for l in lines if l.startswith("#")]
            rows = [l for l in lines if l.strip() and not l.startswith("#")]
            stride = 1
            if p.endswith("_td.asc") and len(rows) > 1000:
                stride = -(-len(rows) // 1000)
            with gzip.open(os.path.join(odir, os.path.basename(p) + ".gz"), "wt") as g:
                g.write(f"# stride {stride}\n")
                g.write("\n".join(head + rows[::stride]) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

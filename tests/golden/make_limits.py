#!/usr/bin/env python3
"""Regenerate tests/golden/limits_unvisited.json.gz: which field cells the
reference's own field sweeps leave un-updated, from its limit-checking build.

oracle/Makefile builds the reference from its sources with CHECK_LIMITS
(oracle/_ref/vulture_ref_chk): every field starts at -1, each update in the
ordinary, UPML and Mur loops asserts the old value is -1 and sets the cell to
0 (grid.h:181-191), and after the run checkGrid (grid.c:951-1025) prints a
warning `FIELD[i][j][k] = value != 0` for every cell of the outer field
limits that is not 0.  Each mesh here runs with NT set to 1, so a cell is
updated at most once without tripping the assertion, and:
  value -1      the sweeps never updated the cell ("unvisited");
  other values  the sweeps updated it and a later step of the loop changed
                it again (SIBC H corrections, sources), so it is visited.
The reference's error count is a bool (grid.c:956), so its closing message
always says one element; the warnings are the record.

The fixture holds, per mesh: the run's exit status, the unvisited cells per
component as [i, j, k] lists, and the number of visited-then-changed cells
per component.  Meshes are the reference's own test inputs
(tests/golden/ref_inputs) and a few generated workloads
(vulture_amd.meshgen, recorded by generator and arguments).

Usage: python tests/golden/make_limits.py   (needs oracle/_ref/vulture_ref_chk)
"""
import gzip
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
EXE = os.path.join(ROOT, "oracle", "_ref", "vulture_ref_chk")
OUT = os.path.join(HERE, "limits_unvisited.json.gz")
FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")
WARN = re.compile(r"\*\*\* Warning: (EX|EY|EZ|HX|HY|HZ)\[(\d+)\]\[(\d+)\]\[(\d+)\] = (\S+) != 0")

# Generated workloads: (key, generator, keyword arguments)
GENERATED = [
    ("gen_debye_pml", "debye", dict(n=24, spheres=3, rmin=2, rmax=4, steps=1)),
    ("gen_debye_mur", "debye", dict(n=20, spheres=2, rmin=2, rmax=4, steps=1, boundary="MUR")),
    ("gen_debye_pec", "debye", dict(n=20, spheres=2, rmin=2, rmax=4, steps=1, boundary="PEC")),
    ("gen_enclosure", "enclosure", dict(nx=28, ny=24, nz=20, pml=4, steps=1, nfreq=2)),
    ("gen_sibc_box", "sibc_box", dict(nx=24, ny=22, nz=20, blocks=3, steps=1, observers=1, nfreq=2)),
]


def run_one(mesh_path):
    """Run the check build on a mesh (NT forced to 1) in its own directory."""
    d = os.path.dirname(mesh_path)
    text = re.sub(r"^NT \d+", "NT 1", open(mesh_path).read(), flags=re.M)
    with open(mesh_path, "w") as f:
        f.write(text)
    r = subprocess.run([EXE, os.path.basename(mesh_path)], cwd=d, capture_output=True, text=True, timeout=600)
    unvisited = {f: [] for f in FIELDS}
    changed = {f: 0 for f in FIELDS}
    for m in WARN.finditer(r.stdout + r.stderr):
        if float(m.group(5)) == -1.0:
            unvisited[m.group(1)].append([int(m.group(2)), int(m.group(3)), int(m.group(4))])
        else:
            changed[m.group(1)] += 1
    return {"rc": r.returncode, "unvisited": unvisited, "changed": changed}


def main():
    from vulture_amd import meshgen
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} is missing: make -C oracle ref")
    out = {}
    inp = os.path.join(HERE, "ref_inputs")
    for name in sorted(os.listdir(inp)):
        if name == "parser":  # deliberately malformed meshes
            continue
        work = tempfile.mkdtemp()
        shutil.copytree(os.path.join(inp, name), work, dirs_exist_ok=True)
        for f in sorted(os.listdir(work)):
            if f.endswith(".mesh"):
                out[f"{name}/{f}"] = run_one(os.path.join(work, f))
        shutil.rmtree(work)
    for key, gen, kw in GENERATED:
        work = tempfile.mkdtemp()
        mesh = meshgen.write(gen, work, **kw)
        rec = run_one(mesh)
        rec["generator"], rec["kwargs"] = gen, kw
        out[key] = rec
        shutil.rmtree(work)
    with gzip.open(OUT, "wt") as g:
        json.dump(out, g, sort_keys=True, separators=(",", ":"))
    for k, v in out.items():
        print(k, v["rc"], {f: len(c) for f, c in v["unvisited"].items() if c}, {f: n for f, n in v["changed"].items() if n})


if __name__ == "__main__":
    main()

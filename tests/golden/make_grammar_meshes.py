#!/usr/bin/env python3
"""Generated meshes for the parts of the card grammar that the reference's
own test meshes leave out, with the reference build's outputs as fixtures.

The reference's test meshes use 16 of the 84 source tags (source.c:97-280)
and neither the BR card (surface.c:140-190, the obsolete six-face boundary
card) nor the BE card (boundary.c:74-104, PML depths per face).  These six
meshes (22^3 cells of 1 mm, 200 steps) cover every source tag -- E, H, J,
JM, I, IM, ID, IMD with soft and hard (=) forms, V and =V, the lumped VR and
IG sources above and below the 0.01 ohm threshold, and the 24 surface
current tags JS/JMS/=JS/=JMS -- under BR-only, BR+BE, and BT boundaries
(PML, PEC, PMC, Mur).  Each source sits at its own spot with a size scaled
to its type, so each one is visible in the point observers.

Written to tests/golden/grammar/<name>/<name>.mesh.  When the reference's
float build is present (oracle/_ref/vulture_ref, built from the reference's
own sources), its ASCII observer files for each mesh are stored gzipped
beside the mesh as ref/<file>.gz: the fixtures the float oracle must match
byte for byte (tests/test_grammar.py).

Usage: python tests/golden/make_grammar_meshes.py
"""
import gzip
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "grammar")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "vulture_ref")

N = 22
AX = "XYZ"

# tag stem -> (bbox kind, size) ; sizes give fields of like magnitude on 1 mm cells
STEMS = {
    "E": ("line", 1.0), "=E": ("line", 0.5), "H": ("vol", 2.6e-3), "=H": ("vol", 1.3e-3),
    "J": ("vol", 1.0), "=J": ("vol", 0.7), "JM": ("vol", 380.0), "=JM": ("vol", 250.0),
    "I": ("vol", 2e-6), "=I": ("vol", 1.5e-6), "IM": ("vol", 7e-4), "=IM": ("vol", 5e-4),
    "ID": ("vol", 3e-9), "=ID": ("vol", 2e-9), "IMD": ("vol", 1e-6), "=IMD": ("vol", 8e-7),
    "V": ("line", 1e-3), "=V": ("line", 8e-4),
}
SURF = {"JS": 1e-3, "=JS": 8e-4, "JMS": 0.38, "=JMS": 0.3}
PAIRS = ["XY", "YZ", "ZX", "XZ", "YX", "ZY"]


def all_tags():
    """(tag, kind, size, resistance) for the 84 tags."""
    out = []
    for stem, (kind, size) in STEMS.items():
        for a in AX:
            out.append((stem + a, kind, size, None))
    for a, r in zip(AX, (50.0, 0.005, 10.0)):
        out.append(("VR" + a, "edge", 1.0, r))
    for a, r in zip(AX, (0.001, 75.0, 25.0)):
        out.append(("IG" + a, "edge", 0.02, r))
    for stem, size in SURF.items():
        for p in PAIRS:
            out.append((stem + p, "surf", size, None))
    return out


def spots():
    """Source spots in the interior, apart from each other and the observers."""
    return [(4 + 5 * a, 4 + 5 * b, 4 + 5 * c) for a in range(3) for b in range(3) for c in range(3)
            if (a, b, c) != (1, 1, 1)]


def bbox(kind, tag, s):
    i, j, k = s
    pol = AX.index(tag[-2] if kind == "surf" else tag[-1])
    if kind in ("line", "edge"):
        b = [i, i, j, j, k, k]
        b[2 * pol + 1] += 1 if kind == "edge" else 2
        return b
    if kind == "surf":
        # electric surface currents on a surface normal to the normal axis;
        # magnetic ones, whose H components sit half a cell off the mesh
        # planes, on a slab one cell thick (as the reference's hardsrc_jmsxy)
        nrm = AX.index(tag[-1])
        b = [i, i + 2, j, j + 2, k, k + 2]
        b[2 * nrm + 1] = b[2 * nrm] + (1 if "JMS" in tag else 0)
        return b
    return [i, i + 2, j, j + 2, k, k + 2]


MESHES = [
    ("g_br_pml", ["BR 0 0 0 0 0 0"], []),
    ("g_br_be", ["BR 0 0 0 0 0 0", "BE 4 5 6 7 4 5"], []),
    ("g_br_mixed", ["BR -1 0 1 0 0 -1", "BE 0 5 0 6 4 0"], []),
    ("g_br_pmc", ["BR 1 1 0 0 -1 -1", "BE 0 0 5 6 0 0"], []),
    ("g_bt_mur", [], [f"BT {f} MUR" for f in ("XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI")]),
    ("g_bt_pml_pec", [], [f"BT {f} PML 5" for f in ("XLO", "XHI", "YLO", "YHI", "ZLO")] + ["BT ZHI PEC"]),
]


def mesh_text(name, sec1, bts, tags):
    lines = ["VM 1.0.0", f"CE generated grammar mesh {name}: {len(tags)} source tags", f"DM {N} {N} {N}"]
    lines += sec1 + ["GS"] + bts
    lines += ["WF wg GAUSSIAN_PULSE 1.0 4.0e-11 1.0e-11", "WF wr RICKER_WAVELET 1.0 5.0e-11 1.2e-11"]
    sp = spots()
    for q, (tag, kind, size, r) in enumerate(tags):
        b = bbox(kind, tag, sp[q % len(sp)])
        wf = "wg" if q % 2 == 0 else "wr"
        delay = f"{1e-12 * (q % 5):.1e}"
        src = f"EX {b[0]} {b[1]} {b[2]} {b[3]} {b[4]} {b[5]} s{q} {tag} {wf}"
        lines.append(src + (f" {r} {size} {delay}" if r is not None else f" {size} {delay}"))
    lines += ["OP 11 11 11 11 11 11 oc TDOM_ASCII", "OP 7 7 15 15 10 10 oa TDOM_ASCII",
              "OP 16 16 6 6 13 13 ob TDOM_ASCII", "OP 12 12 9 9 17 17 of FDOM_ASCII",
              "GE", "NT 200", "OF 1e9 4e10 16", "CN 0.95", "MS 1e-3", "EN"]
    return "\n".join(lines) + "\n"


def plan():
    tags = all_tags()
    per = -(-len(tags) // len(MESHES))
    return [(name, s1, bt, tags[m * per:(m + 1) * per]) for m, (name, s1, bt) in enumerate(MESHES)]


def main():
    for name, s1, bt, tags in plan():
        d = os.path.join(OUT, name)
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        with open(os.path.join(d, f"{name}.mesh"), "w") as f:
            f.write(mesh_text(name, s1, bt, tags))
        if not os.path.exists(REF_BIN):
            print(f"{name}: mesh only (no reference build)")
            continue
        work = tempfile.mkdtemp(prefix=f"gram_{name}_")
        shutil.copy(os.path.join(d, f"{name}.mesh"), work)
        r = subprocess.run([REF_BIN, f"{name}.mesh"], cwd=work, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            sys.exit(f"{name}: reference build failed: {r.stderr[-400:]}")
        os.makedirs(os.path.join(d, "ref"))
        outs = sorted(f for f in os.listdir(work) if f.startswith(("eh_", "wf_")))
        for f in outs:
            with open(os.path.join(work, f), "rb") as src, \
                    gzip.GzipFile(os.path.join(d, "ref", f + ".gz"), "wb", mtime=0) as dst:
                dst.write(src.read())
        shutil.rmtree(work)
        print(f"{name}: {len(tags)} tags, {len(outs)} reference files")


if __name__ == "__main__":
    main()

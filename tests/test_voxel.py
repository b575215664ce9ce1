"""The voxel object loader (vulture_amd/host/vh_voxel.c: MT <name> VOXEL
"<file.vox>" and MB <node> <name> [theta [phi]], the loader the reference
plans in doc/ToDo.md "Voxel object loader" and does not implement).

The reference cannot run a voxel mesh, so the loader is pinned through MB
cards, which it can run:
  * the Debye-sphere workload written as one voxel object gives the same
    observer files, byte for byte, as the same spheres written as one MB card
    per (i, j) run -- and that MB mesh gives the same bytes as the reference
    build (when present);
  * geometry is pinned by what rotations fix: an array placed with a quarter
    or half turn equals the array turned in numpy and placed without one; a
    moved reference point equals a moved node; an object cut at the mesh
    boundary equals the array cut in numpy;
  * an unrotated placement equals explicit MB cards for its k-runs, with the
    literal key lines equal to MT SIMPLE cards of the same values;
  * 1-, 2- and 4-byte keys give the same result;
  * malformed objects and placements are rejected with a message.
All runs use the CPU oracle, which shares the product's host set-up.
"""
import filecmp
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from vulture_amd import meshgen

H = 1e-3
N = 24


def _outputs_equal(a, b):
    fa = sorted(f for f in os.listdir(a) if f.endswith((".asc", ".dat")))
    fb = sorted(f for f in os.listdir(b) if f.endswith((".asc", ".dat")))
    assert fa == fb and fa
    return [f for f in fa if not filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False)]


def _mesh(body, steps=120):
    lines = ["VM 1.0.0", "CE voxel loader test", f"DM {N} {N} {N}", "GS"]
    lines += [f"BT {f} PML 4" for f in ("XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI")]
    lines += ["MT dd DEBYE 4.0 1.0e-2 1.0 1.0e+11 -5.0e+10"]
    lines += body
    lines += [meshgen._gauss_card("wf1", 60e9), "EX 6 6 7 7 6 7 src EZ wf1 1.0",
              "EX 17 17 15 16 16 16 src2 =EY wf1 0.5",
              "OP 12 12 12 12 12 12 o1 TDOM_ASCII", "OP 9 9 14 14 10 10 o2 TDOM_ASCII",
              "OP 15 15 8 8 13 13 o3 TDOM_ASCII", "GE", f"NT {steps}", f"MS {H}", "EN"]
    return "\n".join(lines) + "\n"


KEYS = ["1 2.5 0.01 1.0", "2 dd", "3 7.0 0.0 1.0"]


def _object(d, name, arr, width=1, ref=(0.0, 0.0, 0.0), keys=KEYS, vox=None):
    dt = {1: np.uint8, 2: "<u2", 4: "<u4"}[width]
    np.asarray(arr).astype(dt).tofile(os.path.join(d, f"{name}.keys"))
    nx, ny, nz = np.asarray(arr).shape
    text = vox if vox is not None else "\n".join(
        [f'"{name}.keys"', str(width), f"{nx} {ny} {nz}", f"{H} {H} {H}", f"{ref[0]} {ref[1]} {ref[2]}"] + keys)
    with open(os.path.join(d, f"{name}.vox"), "w") as f:
        f.write(text + "\n")
    return f'MT {name} VOXEL "{name}.vox"'


def _run(d, body, precision="fp64"):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "t.mesh"), "w") as f:
        f.write(_mesh(body))
    out = os.path.join(d, "out")
    oracle.run_mesh(os.path.join(d, "t.mesh"), out, precision=precision, threads=2)
    return out


def _changes_fields(out, tmp_path):
    """The object is seen: the run differs from one without it."""
    return bool(_outputs_equal(out, _run(str(tmp_path / "empty"), [])))


def _array(seed=7, shape=(5, 4, 3)):
    return np.random.default_rng(seed).integers(0, 4, size=shape)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_sphere_voxels_equal_mb_cards(precision, tmp_path):
    outs = []
    for v in (False, True):
        d = tmp_path / f"v{int(v)}"
        mesh = meshgen.write("debye", str(d), n=40, spheres=6, steps=150, observers=3, voxel=v)
        outs.append(str(d / "out"))
        oracle.run_mesh(mesh, outs[-1], precision=precision, threads=4)
    assert "MT spheres VOXEL" in open(tmp_path / "v1" / "debye.mesh").read()
    assert not _outputs_equal(*outs)


@pytest.mark.skipif(not os.path.exists(oracle.REF_BIN), reason="reference build not present")
def test_sphere_voxels_equal_reference_build_on_mb_cards(tmp_path):
    """The voxel form through the float oracle against the MB form through
    the reference's own float build: ASCII observer files byte for byte."""
    a, b = tmp_path / "vox", tmp_path / "ref"
    ma = meshgen.write("debye", str(a), n=40, spheres=6, steps=150, observers=3, voxel=True)
    mb = meshgen.write("debye", str(b), n=40, spheres=6, steps=150, observers=3)
    oracle.run_mesh(ma, str(a), precision="fp32", full_precision=False, threads=4)
    r = subprocess.run([oracle.REF_BIN, os.path.basename(mb)], cwd=b, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-500:]
    outs = sorted(f for f in os.listdir(b) if f.startswith("eh_"))
    assert outs
    for f in outs:
        assert filecmp.cmp(a / f, b / f, shallow=False), f


# (theta, phi, turn of the array in numpy, node shift for the turned array)
TURNS = [
    (0, 90, lambda A: A.transpose(1, 0, 2)[::-1], lambda s: (-s[1], 0, 0)),
    (0, -90, lambda A: A.transpose(1, 0, 2)[:, ::-1], lambda s: (0, -s[0], 0)),
    (180, 0, lambda A: A[::-1, :, ::-1], lambda s: (-s[0], 0, -s[2])),
    (90, 90, lambda A: A.transpose(1, 2, 0)[::-1, :, ::-1], lambda s: (-s[1], 0, -s[0])),
    (0, 180, lambda A: A[::-1, ::-1], lambda s: (-s[0], -s[1], 0)),
]


@pytest.mark.parametrize("theta,phi,turn,shift", TURNS, ids=[f"t{t[0]}p{t[1]}" for t in TURNS])
def test_quarter_turns_equal_turned_arrays(theta, phi, turn, shift, tmp_path):
    A = _array()
    node = (12, 11, 12)
    d1 = str(tmp_path / "placed")
    os.makedirs(d1)
    o1 = _run(d1, [_object(d1, "obj", A), f"MB {node[0]} {node[0]} {node[1]} {node[1]} {node[2]} {node[2]} obj "
                                           f"{theta} {phi}"])
    B = np.ascontiguousarray(turn(A))
    sh = shift(A.shape)
    n2 = [node[a] + sh[a] for a in range(3)]
    d2 = str(tmp_path / "turned")
    os.makedirs(d2)
    o2 = _run(d2, [_object(d2, "obj", B), f"MB {n2[0]} {n2[0]} {n2[1]} {n2[1]} {n2[2]} {n2[2]} obj"])
    assert not _outputs_equal(o1, o2)
    assert _changes_fields(o1, tmp_path)


def test_reference_point_equals_moved_node(tmp_path):
    A = _array(seed=3)
    d1, d2 = str(tmp_path / "a"), str(tmp_path / "b")
    os.makedirs(d1)
    os.makedirs(d2)
    o1 = _run(d1, [_object(d1, "obj", A, ref=(2 * H, 1 * H, 3 * H)), "MB 12 12 10 10 13 13 obj"])
    o2 = _run(d2, [_object(d2, "obj", A), "MB 10 10 9 9 10 10 obj"])
    assert not _outputs_equal(o1, o2)
    assert _changes_fields(o1, tmp_path)


def test_object_cut_at_mesh_boundary(tmp_path):
    A = _array(seed=11, shape=(6, 5, 4))
    ds = [str(tmp_path / c) for c in "abcd"]
    for d in ds:
        os.makedirs(d)
    # 6 voxels from x node 21 on a 24-cell mesh: the last three are cut
    o1 = _run(ds[0], [_object(ds[0], "obj", A), "MB 21 21 3 3 9 9 obj"])
    o2 = _run(ds[1], [_object(ds[1], "obj", A[:3]), "MB 21 21 3 3 9 9 obj"])
    assert not _outputs_equal(o1, o2)
    # a half turn about z at node (2, 2, 2): voxel (a, b, c) goes to cell
    # (1 - a, 1 - b, 2 + c), so only a, b in {0, 1} stay inside the mesh
    o3 = _run(ds[2], [_object(ds[2], "obj", A), "MB 2 2 2 2 2 2 obj 0 180"])
    o4 = _run(ds[3], [_object(ds[3], "obj", np.ascontiguousarray(A[1::-1, 1::-1, :])), "MB 0 0 0 0 2 2 obj"])
    assert not _outputs_equal(o3, o4)
    assert _changes_fields(o1, tmp_path) and _changes_fields(o3, tmp_path)


def _runs_as_cards(A, node, names):
    """MB cards of an unrotated placement, one per k-run of a mapped key, with
    the loader's face masks (low x, y, z faces; high z face when the next cell
    along k is not in another run)."""
    cards = []
    nx, ny, nz = A.shape
    for i in range(nx):
        for j in range(ny):
            k = 0
            while k < nz:
                if A[i, j, k] not in names:
                    k += 1
                    continue
                e = k
                while e + 1 < nz and A[i, j, e + 1] == A[i, j, k]:
                    e += 1
                zhi = e + 1 >= nz or A[i, j, e + 1] not in names
                x, y, z = node[0] + i, node[1] + j, node[2] + k
                cards.append(f"MB {x} {x + 1} {y} {y + 1} {z} {node[2] + e + 1} {names[A[i, j, k]]} "
                             f"10101{1 if zhi else 0}")
                k = e + 1
    return cards


def test_unrotated_placement_equals_mb_cards(tmp_path):
    A = _array(seed=5, shape=(5, 6, 7))
    d1, d2 = str(tmp_path / "vox"), str(tmp_path / "mb")
    os.makedirs(d1)
    os.makedirs(d2)
    o1 = _run(d1, [_object(d1, "obj", A), "MB 8 8 7 7 9 9 obj"])
    mts = ["MT s1 SIMPLE 2.5 0.01 1.0", "MT s3 SIMPLE 7.0 0.0 1.0"]
    o2 = _run(d2, mts + _runs_as_cards(A, (8, 7, 9), {1: "s1", 2: "dd", 3: "s3"}))
    assert not _outputs_equal(o1, o2)


@pytest.mark.parametrize("width", [2, 4])
def test_key_widths(width, tmp_path):
    """Keys scaled into the top byte of a wider key give the same result."""
    A = _array(seed=9)
    scale = 1 << (8 * (width - 1))
    d1, d2 = str(tmp_path / "w1"), str(tmp_path / f"w{width}")
    os.makedirs(d1)
    os.makedirs(d2)
    keys = [f"{int(k) * scale} {rest}" for k, rest in (line.split(" ", 1) for line in KEYS)]
    o1 = _run(d1, [_object(d1, "obj", A), "MB 9 9 9 9 9 9 obj 30 40"])
    o2 = _run(d2, [_object(d2, "obj", A.astype(np.int64) * scale, width=width, keys=keys),
                   "MB 9 9 9 9 9 9 obj 30 40"])
    assert not _outputs_equal(o1, o2)


def test_unlisted_keys_leave_cells(tmp_path):
    """Keys with no line leave the material as it was: an object whose
    keys are all unlisted changes nothing."""
    A = _array(seed=2)
    d1, d2 = str(tmp_path / "a"), str(tmp_path / "b")
    os.makedirs(d1)
    os.makedirs(d2)
    o1 = _run(d1, [_object(d1, "obj", A + 10), "MB 9 9 9 9 9 9 obj"])
    o2 = _run(d2, [])
    assert not _outputs_equal(o1, o2)


BAD = {
    "missing file": (None, 'MT obj VOXEL "nothere.vox"', "cannot open voxel file"),
    "short data": ("short", None, "expected exactly"),
    "key width 3": ("width3", None, "bytes per key"),
    "unknown medium": ("unknown", None, "bad key line"),
    "duplicate key": ("dup", None, "listed twice"),
    "bad eps_r": ("eps", None, "bad key line"),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_bad_objects_rejected(case, tmp_path):
    kind, card, msg = BAD[case]
    d = str(tmp_path)
    A = _array()
    if kind is None:
        body = [card, "MB 9 9 9 9 9 9 obj"]
    else:
        vox = {
            "short": None,
            "width3": "\n".join(['"obj.keys"', "3", "5 4 3", f"{H} {H} {H}", "0 0 0"] + KEYS),
            "unknown": "\n".join(['"obj.keys"', "1", "5 4 3", f"{H} {H} {H}", "0 0 0", "1 nosuch"]),
            "dup": "\n".join(['"obj.keys"', "1", "5 4 3", f"{H} {H} {H}", "0 0 0", "1 dd", "1 dd"]),
            "eps": "\n".join(['"obj.keys"', "1", "5 4 3", f"{H} {H} {H}", "0 0 0", "1 0.5 0 1"]),
        }[kind]
        body = [_object(d, "obj", A, vox=vox), "MB 9 9 9 9 9 9 obj"]
        if kind == "short":
            with open(os.path.join(d, "obj.keys"), "ab") as f:
                f.write(b"\0")
    with pytest.raises(RuntimeError, match=msg):
        _run(d, body)


@pytest.mark.parametrize("card", ["MB 9 10 9 9 9 9 obj", "MB 9 9 9 9 9 9 obj 10 20 30", "MB 9 9 9 9 9 9 obj x",
                                  "MB 9 9 9 9 30 30 obj"])
def test_bad_placements_rejected(card, tmp_path):
    d = str(tmp_path)
    with pytest.raises(RuntimeError):
        _run(d, [_object(d, "obj", _array()), card])


def test_voxel_name_clash_rejected(tmp_path):
    d = str(tmp_path)
    with pytest.raises(RuntimeError):
        _run(d, [_object(d, "dd", _array()), "MB 9 9 9 9 9 9 dd"])

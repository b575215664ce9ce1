"""Card grammar beyond the reference's own test meshes: the BR and BE cards
and all 84 source tags (tests/golden/grammar, written by
tests/golden/make_grammar_meshes.py).

The float oracle must write the same observer files, byte for byte, as the
reference's own float build did for these meshes (the gzipped fixtures
beside each mesh); the FP64 oracle must match the reference's FP64 build
byte for byte where that build exists.  Every source in these meshes changes
the observer files when it is removed (checked when the meshes were made),
so each tag's set-up and update is pinned.
"""
import filecmp
import gzip
import os
import shutil
import subprocess

import pytest

import oracle

from conftest import GOLDEN

G = os.path.join(GOLDEN, "grammar")
NAMES = sorted(os.listdir(G)) if os.path.isdir(G) else []
REF64_BIN = os.path.join(os.path.dirname(oracle.REF_BIN), "vulture_ref64")


def _stage(name, dst):
    os.makedirs(dst)
    shutil.copy(os.path.join(G, name, f"{name}.mesh"), dst)
    return os.path.join(dst, f"{name}.mesh")


def test_every_source_tag_is_covered():
    """The meshes hold each of the 84 source tags once, and BR and BE cards."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLDEN, "make_grammar_meshes.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    tags, cards = [], set()
    for n in NAMES:
        for line in open(os.path.join(G, n, f"{n}.mesh")):
            f = line.split()
            if f and f[0] == "EX":
                tags.append(f[8])
            if f:
                cards.add(f[0])
    assert sorted(tags) == sorted(t[0] for t in mg.all_tags()) and len(set(tags)) == 84
    assert {"BR", "BE", "BT"} <= cards


@pytest.mark.parametrize("name", NAMES)
def test_float_oracle_matches_reference_fixtures(name, tmp_path):
    mesh = _stage(name, str(tmp_path / "in"))
    out = tmp_path / "out"
    oracle.run_mesh(mesh, str(out), precision="fp32", full_precision=False, threads=2)
    ref = os.path.join(G, name, "ref")
    files = sorted(f[:-3] for f in os.listdir(ref))
    assert files and sorted(f for f in os.listdir(out) if f.startswith(("eh_", "wf_"))) == files
    for f in files:
        with gzip.open(os.path.join(ref, f + ".gz"), "rb") as z:
            assert (out / f).read_bytes() == z.read(), f"{name}/{f} differs from the reference build's"


@pytest.mark.skipif(not os.path.exists(REF64_BIN), reason="FP64 reference build not present")
@pytest.mark.parametrize("name", NAMES)
def test_fp64_oracle_matches_fp64_reference_build(name, tmp_path):
    a, b = tmp_path / "oracle", tmp_path / "ref"
    mesh = _stage(name, str(a))
    _stage(name, str(b))
    oracle.run_mesh(mesh, str(a), precision="fp64", full_precision=True, threads=2)
    r = subprocess.run([REF64_BIN, f"{name}.mesh"], cwd=b, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-500:]
    outs = sorted(f for f in os.listdir(b) if f.startswith(("eh_", "wf_")))
    assert outs
    for f in outs:
        assert filecmp.cmp(a / f, b / f, shallow=False), f"{name}/{f} differs from the FP64 reference build"

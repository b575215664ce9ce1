"""The CPU oracle pinned to the reference.

1. Byte identity: the float build of the oracle writes the same observer
   files, byte for byte, as the reference compiled from its own sources
   (oracle/Makefile, target `ref`) on every reference test mesh; the FP64
   build writes the same files as the reference compiled in double
   precision from its own sources (target _ref/vulture_ref64).  Runs only
   where those builds exist.
2. The reference's validation files (tests/<name>/eh_*_td.asc, wf_*_td.asc;
   one axis permutation per family kept as fixtures, long series
   subsampled): the float oracle must pass util/pydiff.py's acceptance rule
   at its "marginal" level (numpy.allclose rtol=1e-4, atol=1e-1) and stay
   within a relative L2 bound.  The validation files were written by another
   build of the reference (the reference compiled here reproduces them no
   better than the oracle does, and byte-identically to it): for four meshes
   they differ by 5e-3..8e-2, which the bounds below record.  The FP64 oracle
   is held to the same bounds, widened only where the signal is
   cancellation-dominated (sgref_pml6: 3e-8 peak field).
"""
import filecmp
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

from conftest import GOLDEN

OUT = os.path.join(GOLDEN, "ref_outputs")
INP = os.path.join(GOLDEN, "ref_inputs")
NAMES = sorted(os.listdir(OUT)) if os.path.isdir(OUT) else []
ALL_INPUTS = sorted(n for n in os.listdir(INP) if n != "parser") if os.path.isdir(INP) else []
SLOW = {"sgref_mur", "sgref_pml12", "sgref_pml6", "dgref_ref", "dgrefmb_ref"}

# Relative L2 bounds of the oracle against the validation files.
DEFAULT_TOL = 1e-3
STALE = {"parplate1D_debye_xyz": 0.1, "parplate1D_planewave_xyz": 1e-2, "planewave_empty_xyz": 1e-2,
         "planewave_pec_xyz": 1e-2}
FP64_EXTRA = {"sgref_pml6": 5e-2}


def load_golden(path):
    with gzip.open(path, "rt") as f:
        stride = int(f.readline().split()[2])
        rows = [l.split() for l in f if l.strip() and not l.startswith("#")]
    return stride, np.array(rows, dtype=np.float64)


def load_asc(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    return np.array(rows, dtype=np.float64)


def run_oracle(name, tmp_path, precision):
    work = tmp_path / name
    shutil.copytree(os.path.join(INP, name), work)
    mesh = [f for f in os.listdir(work) if f.endswith(".mesh")][0]
    oracle.run_mesh(str(work / mesh), str(work), precision=precision, full_precision=False, threads=4)
    return work


RUN_SLOW = os.environ.get("VULTURE_SLOW_TESTS") == "1"


def _params(names, opt_in=False):
    """Slow meshes are marked `slow`; with opt_in they also need VULTURE_SLOW_TESTS=1."""
    out = []
    for n in names:
        if n in SLOW:
            marks = [pytest.mark.slow]
            if opt_in and not RUN_SLOW:
                marks.append(pytest.mark.skip(reason="15-60 s each; set VULTURE_SLOW_TESTS=1"))
            out.append(pytest.param(n, marks=marks))
        else:
            out.append(n)
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("name", _params(NAMES))
def test_oracle_matches_validation_files(name, precision, tmp_path):
    work = run_oracle(name, tmp_path, precision)
    tol = STALE.get(name, DEFAULT_TOL)
    if precision == "fp64":
        tol = max(tol, FP64_EXTRA.get(name, 0.0))
    files = [g for g in sorted(os.listdir(os.path.join(OUT, name))) if g.endswith("_td.asc.gz")]
    assert files
    for g in files:
        stride, valid = load_golden(os.path.join(OUT, name, g))
        test = load_asc(str(work / g[:-3]))[::stride]
        assert test.shape == valid.shape, g
        rel = np.linalg.norm(valid[:, 2:] - test[:, 2:]) / max(np.linalg.norm(valid[:, 2:]), 1e-300)
        assert rel <= tol, f"{name}/{g}: relative L2 {rel:.2e} > {tol:.0e}"
        if precision == "fp32":
            assert np.allclose(valid, test, rtol=1e-4, atol=1e-1), f"{name}/{g}: pydiff marginal rule fails"


REF_BIN = oracle.REF_BIN


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference build not present (needs the reference sources)")
@pytest.mark.parametrize("name", _params(ALL_INPUTS, opt_in=True))
def test_float_oracle_is_byte_identical_to_reference_build(name, tmp_path):
    a = tmp_path / "oracle"
    b = tmp_path / "reference"
    shutil.copytree(os.path.join(INP, name), a)
    shutil.copytree(os.path.join(INP, name), b)
    mesh = [f for f in os.listdir(a) if f.endswith(".mesh")][0]
    oracle.run_mesh(str(a / mesh), str(a), precision="fp32", full_precision=False, threads=4)
    r = subprocess.run([REF_BIN, mesh], cwd=b, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-500:]
    outs = sorted(f for f in os.listdir(b) if f.startswith(("eh_", "wf_")) or f.endswith(".dat") and
                  f in ("impulse.dat", "excite.dat", "process.dat"))
    assert outs
    for f in outs:
        assert filecmp.cmp(a / f, b / f, shallow=False), f"{name}/{f} differs from the reference build"


REF64_BIN = os.path.join(os.path.dirname(oracle.REF_BIN), "vulture_ref64")


@pytest.mark.skipif(not os.path.exists(REF64_BIN), reason="FP64 reference build not present (needs the reference sources)")
@pytest.mark.parametrize("name", _params(NAMES, opt_in=True) +
                         ([n for n in ALL_INPUTS if n not in NAMES] if RUN_SLOW else []))
def test_fp64_oracle_is_byte_identical_to_fp64_reference_build(name, tmp_path):
    """The FP64 oracle (real = double, -O2 without contraction, values printed
    with 17 significant digits) against the reference built the same way from
    its own sources (oracle/Makefile target _ref/vulture_ref64: real double,
    %le scans, %24.16e observer values): every observer file byte for byte.
    This pins the set-up the device library's host shares with the oracle in
    the precision the product runs it in (VH_REAL=double).  One mesh per test
    family by default; every reference mesh with VULTURE_SLOW_TESTS=1."""
    a = tmp_path / "oracle"
    b = tmp_path / "reference"
    shutil.copytree(os.path.join(INP, name), a)
    shutil.copytree(os.path.join(INP, name), b)
    mesh = [f for f in os.listdir(a) if f.endswith(".mesh")][0]
    oracle.run_mesh(str(a / mesh), str(a), precision="fp64", full_precision=True, threads=4)
    r = subprocess.run([REF64_BIN, mesh], cwd=b, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-500:]
    outs = sorted(f for f in os.listdir(b) if f.startswith(("eh_", "wf_")) or f in ("impulse.dat", "excite.dat",
                                                                                   "process.dat"))
    assert outs
    for f in outs:
        if f == "impulse.dat":
            # The header's dt is written as fwrite(&dt, sizeof(float)) of a
            # `real` argument (observer.c:1137-1147): with real = double the
            # FP64 build writes the low four bytes of the double there.  The
            # oracle writes the float the float build writes; every other
            # byte (step count, node indices, float field samples) must match.
            x, y = (a / f).read_bytes(), (b / f).read_bytes()
            assert len(x) == len(y) and x[:4] == y[:4] and x[8:] == y[8:], f"{name}/{f} differs"
            continue
        assert filecmp.cmp(a / f, b / f, shallow=False), f"{name}/{f} differs from the FP64 reference build"


def test_parser_rejects_reference_bad_meshes(tmp_path):
    """tests/parser: a mesh without a VM card and one cut short both fail (mesh.c:58-191)."""
    for m in ("no_vm.mesh", "eof.mesh"):
        with pytest.raises(RuntimeError):
            oracle.run_mesh(os.path.join(INP, "parser", m), str(tmp_path / m), precision="fp64", steps=1)

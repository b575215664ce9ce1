"""The C-ABI library: it loads without a GPU and exports every entry point
that include/vgpu.h declares; the host library exports its driver API."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "vgpu.h")
LIB = os.path.join(ROOT, "vulture_amd", "lib")


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vgpu_[a-z_0-9]+)\s*\(", text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_header_declares_entry_points():
    names = declared_functions(HEADER)
    for n in ("vgpu_create", "vgpu_run", "vgpu_finalize", "vgpu_destroy", "vgpu_read_point_observer"):
        assert n in names


def test_libvgpu_exports_every_declared_symbol():
    lib = os.path.join(LIB, "libvgpu.so")
    assert os.path.exists(lib), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    missing = [n for n in declared_functions(HEADER) if n not in exported(lib)]
    assert not missing, f"declared but not exported: {missing}"


def test_libraries_load_without_gpu():
    ctypes.CDLL(os.path.join(LIB, "libvgpu.so"), mode=ctypes.RTLD_GLOBAL)
    host = ctypes.CDLL(os.path.join(LIB, "libvhost.so"))
    for n in ("vh_run_mesh_gpu", "vh_gpu_open", "vh_gpu_close", "vh_gpu_setup", "vh_solve_sizes", "vh_read_mesh",
              "vh_init", "vh_solve_field_limits"):
        assert hasattr(host, n), n


def test_null_context_is_rejected():
    v = ctypes.CDLL(os.path.join(LIB, "libvgpu.so"))
    v.vgpu_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
    v.vgpu_last_error.restype = ctypes.c_char_p
    assert v.vgpu_run(None, 0, 1, None) != 0
    assert v.vgpu_last_error(None) == b"null context"


def test_python_front_end_fails_loudly_without_library(monkeypatch, tmp_path):
    import vulture_amd
    monkeypatch.setattr(vulture_amd, "_LIB", str(tmp_path))
    monkeypatch.setattr(vulture_amd, "_vgpu", None)
    monkeypatch.setattr(vulture_amd, "_host", None)
    with pytest.raises(vulture_amd.VgpuError):
        vulture_amd.load_vgpu()


def test_command_line_solver_fails_loudly_without_a_device(tmp_path):
    """No CPU fallback: with no visible GPU the solver exits with the
    reference's fatal status (6) and says why; no arguments print the usage."""
    import shutil
    cli = os.path.join(ROOT, "vulture_amd", "bin", "vgvulture")
    assert os.path.exists(cli), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    shutil.copytree(os.path.join(ROOT, "tests", "golden", "ref_inputs", "freespace_pml"), tmp_path / "m")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([cli, str(tmp_path / "m" / "freespace_pml.mesh")], env=env, capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    assert r.returncode == 6, r.stdout + r.stderr
    assert "no usable GPU" in r.stderr
    r = subprocess.run([cli], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr

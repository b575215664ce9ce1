"""vulture_amd: the Vulture FDTD time-stepping core on one AMD MI355X.

The solver reads Vulture mesh files and writes Vulture observer files; the
time loop runs in hand-written gfx950 HIP kernels behind the C-ABI declared
in include/vgpu.h.  This module is a thin ctypes front end:

  run_mesh(mesh, outdir, precision="fp64")  -> device loop time in ms

It loads the in-tree libraries built by vulture_amd/build.py and fails
loudly when they are missing: there is no CPU fallback.
"""
import ctypes
import os

__all__ = ["run_mesh", "load_vgpu", "load_host", "library_path", "VgpuError"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib")
_vgpu = None
_host = None


class VgpuError(RuntimeError):
    pass


def library_path(name):
    return os.path.join(_LIB, name)


def load_vgpu():
    """Load libvgpu.so (the C-ABI).  Raises if it has not been built."""
    global _vgpu
    if _vgpu is None:
        p = library_path("libvgpu.so")
        if not os.path.exists(p):
            raise VgpuError(f"{p} is missing: run vulture_amd/build.py (no CPU fallback exists)")
        _vgpu = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    return _vgpu


def load_host():
    """Load libvhost.so (mesh parser, set-up, device driver)."""
    global _host
    if _host is None:
        load_vgpu()
        p = library_path("libvhost.so")
        if not os.path.exists(p):
            raise VgpuError(f"{p} is missing: run vulture_amd/build.py")
        _host = ctypes.CDLL(p)
        f = _host.vh_run_mesh_gpu
        f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                      ctypes.c_ulong, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_int]
        f.restype = ctypes.c_int
    return _host


def run_mesh(mesh, outdir, precision="fp64", full_precision=True, steps=-1, chunk=0, device=0):
    """Run a Vulture mesh on the GPU, writing observer files into outdir.

    precision: "fp64" (plain form) or "fp32" (single-precision storage).
    full_precision: write observer data with 17 significant digits.
    steps: override NT when >= 0.  chunk: steps per device run (0 = all).
    Returns the device time of the time loop in milliseconds.
    """
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32'")
    host = load_host()
    os.makedirs(outdir, exist_ok=True)
    ms = ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(1024)
    rc = host.vh_run_mesh_gpu(os.fsencode(mesh), os.fsencode(outdir), 1 if precision == "fp32" else 0, device,
                              1 if full_precision else 0, steps, chunk, ctypes.byref(ms), err, len(err))
    if rc != 0:
        raise VgpuError(err.value.decode(errors="replace"))
    return ms.value

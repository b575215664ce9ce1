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

__all__ = ["run_mesh", "DeviceSolve", "load_vgpu", "load_host", "library_path", "VgpuError"]

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
        f = _host.vh_run_meshes_gpu_group
        f.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_int, ctypes.c_ulong, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p,
                      ctypes.c_int]
        f.restype = ctypes.c_int
        f = _host.vh_gpu_open
        f.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_ulong,
                      ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_int]
        f.restype = ctypes.c_int
        _host.vh_gpu_close.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _host.vh_gpu_close.restype = None
        _host.vh_solve_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        _host.vh_solve_sizes.restype = None
        _host.vh_solve_shape.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        _host.vh_solve_shape.restype = None
        _host.vh_solve_field_limits.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        _host.vh_solve_field_limits.restype = None
        v = _vgpu
        v.vgpu_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
        v.vgpu_run.restype = ctypes.c_int
        v.vgpu_synchronize.argtypes = [ctypes.c_void_p]
        v.vgpu_synchronize.restype = ctypes.c_int
        v.vgpu_last_run_timing.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 3
        v.vgpu_last_run_timing.restype = ctypes.c_int
        v.vgpu_last_error.argtypes = [ctypes.c_void_p]
        v.vgpu_last_error.restype = ctypes.c_char_p
        v.vgpu_footprint_bytes.argtypes = [ctypes.c_void_p]
        v.vgpu_footprint_bytes.restype = ctypes.c_uint64
        v.vgpu_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        v.vgpu_get_stats.restype = ctypes.c_int
        v.vgpu_download_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        v.vgpu_download_field.restype = ctypes.c_int
        v.vgpu_check_limits.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(LimitsReport),
                                        ctypes.c_void_p]
        v.vgpu_check_limits.restype = ctypes.c_int
        v.vgpu_check_limits_noalloc.argtypes = v.vgpu_check_limits.argtypes
        v.vgpu_check_limits_noalloc.restype = ctypes.c_int
        v.vgpu_noalloc_scratch_bytes.argtypes = []
        v.vgpu_noalloc_scratch_bytes.restype = ctypes.c_size_t
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


GROUP_MAX = 16  # VGPU_GROUP_MAX (include/vgpu.h)


def run_meshes_group(meshes, outdirs, precision="fp32", full_precision=True, chunk=0, device=0):
    """Run several Vulture meshes on the GPU together, as one packed group.

    Every step kernel is launched once over all of them (vgpu_run_group)
    instead of once per mesh, for grids too small to fill the device alone;
    each mesh's observer files come out as run_mesh writes them.  At most
    GROUP_MAX meshes.  Returns the host's wall time of the time loop in ms.
    """
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32'")
    if len(meshes) != len(outdirs) or not 1 <= len(meshes) <= GROUP_MAX:
        raise ValueError(f"1 to {GROUP_MAX} meshes, each with an output directory")
    host = load_host()
    for d in outdirs:
        os.makedirs(d, exist_ok=True)
    n = len(meshes)
    m = (ctypes.c_char_p * n)(*[os.fsencode(x) for x in meshes])
    o = (ctypes.c_char_p * n)(*[os.fsencode(x) for x in outdirs])
    ms = ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(1024)
    rc = host.vh_run_meshes_gpu_group(m, o, n, 1 if precision == "fp32" else 0, device, 1 if full_precision else 0,
                                      chunk, ctypes.byref(ms), err, len(err))
    if rc != 0:
        raise VgpuError(err.value.decode(errors="replace"))
    return ms.value


class Stats(ctypes.Structure):
    """vgpu_stats (include/vgpu.h)."""
    _fields_ = [("cells", ctypes.c_uint64), ("mesh_cells", ctypes.c_uint64), ("debye_sites", ctypes.c_uint64),
                ("debye_poles", ctypes.c_int), ("debye_complex", ctypes.c_int), ("sibc_faces", ctypes.c_uint64),
                ("correction_rows", ctypes.c_uint64 * 2), ("bytes", ctypes.c_uint64),
                ("fast_cells", ctypes.c_uint64 * 2), ("debye_fused_sites", ctypes.c_uint64),
                ("indexed_media", ctypes.c_int), ("debye_fused", ctypes.c_int), ("sibc_state", ctypes.c_int),
                ("leap", ctypes.c_int), ("fix_cells", ctypes.c_uint64), ("fix_chunks", ctypes.c_uint64),
                ("graph_runs", ctypes.c_uint64)]


class LimitsReport(ctypes.Structure):
    """vgpu_limits_report (include/vgpu.h)."""
    _fields_ = [("missed", ctypes.c_uint64 * 6), ("repeated", ctypes.c_uint64 * 6),
                ("outside", ctypes.c_uint64 * 6), ("first", ctypes.c_int64 * 6)]


class DeviceSolve:
    """A mesh parsed, set up and resident on the device, stepped on demand.

    Wraps vh_gpu_open (parse + set-up + upload) and the vgpu C-ABI:
    ``run(first, n)`` enqueues n time steps, ``synchronize()`` waits and
    returns the device milliseconds of the last run (hipEvents on the run's
    stream).  With ``kernel_events=K`` the first K steps of each run also
    time the E and H volume kernels (``kernel_ms()``).
    """

    def __init__(self, mesh, precision="fp32", device=0, steps=-1, kernel_events=0, chunk=0):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        host = load_host()
        self._v = load_vgpu()
        if kernel_events:
            os.environ["VGPU_KERNEL_EVENTS"] = str(int(kernel_events))
        self._s = ctypes.c_void_p()
        self._c = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        try:
            rc = host.vh_gpu_open(os.fsencode(mesh), 1 if precision == "fp32" else 0, device, steps, chunk,
                                  ctypes.byref(self._s), ctypes.byref(self._c), err, len(err))
        finally:
            if kernel_events:
                del os.environ["VGPU_KERNEL_EVENTS"]
        if rc != 0:
            raise VgpuError(err.value.decode(errors="replace"))
        sizes = (ctypes.c_longlong * 3)()
        host.vh_solve_sizes(self._s, sizes)
        self.cells, self.mesh_cells, self.steps = int(sizes[0]), int(sizes[1]), int(sizes[2])
        self.precision = precision

    def _check(self, rc):
        if rc != 0:
            raise VgpuError(self._v.vgpu_last_error(self._c).decode(errors="replace"))

    def run(self, first, nsteps, stream=None):
        self._check(self._v.vgpu_run(self._c, first, nsteps, stream))

    def synchronize(self):
        self._check(self._v.vgpu_synchronize(self._c))
        t = ctypes.c_double()
        self._v.vgpu_last_run_timing(self._c, ctypes.byref(t), None, None)
        return t.value

    def kernel_ms(self):
        e, h = ctypes.c_double(), ctypes.c_double()
        self._v.vgpu_last_run_timing(self._c, None, ctypes.byref(e), ctypes.byref(h))
        return e.value, h.value

    def footprint_bytes(self):
        return int(self._v.vgpu_footprint_bytes(self._c))

    def stats(self):
        st = Stats()
        self._check(self._v.vgpu_get_stats(self._c, ctypes.byref(st)))
        d = {k: getattr(st, k) for k, _ in Stats._fields_}
        d["correction_rows"] = [int(x) for x in st.correction_rows]
        d["fast_cells"] = [int(x) for x in st.fast_cells]
        return d

    def download(self, component):
        """Field component 0..5 (Ex..Hz) as float64 [nx][ny][nz] including ghost cells."""
        import numpy as np
        n = self.grid_shape()
        out = np.empty(n, dtype=np.float64)
        self._check(self._v.vgpu_download_field(self._c, component, out.ctypes.data))
        return out

    def field_limits(self):
        """Outer field limits gfolim[component][face] of the solve (6x6 ints)."""
        lim = (ctypes.c_int * 36)()
        load_host().vh_solve_field_limits(self._s, lim)
        return [list(lim[6 * c:6 * c + 6]) for c in range(6)]

    def check_limits(self, counts=True, noalloc=False):
        """Update-once check of the field sweeps (vgpu_check_limits): one step
        of Mur E, the E sweep and the H sweep counts its updates per cell,
        tallied against the outer field limits; no field changes.  Returns a
        dict of per-component lists (missed, repeated, outside, first) and,
        with counts, the visits per cell as uint8 [6][nx][ny][nz].  With
        noalloc the entry point that allocates nothing runs the check
        (vgpu_check_limits_noalloc)."""
        import numpy as np
        lim = (ctypes.c_int * 36)()
        load_host().vh_solve_field_limits(self._s, lim)
        rep = LimitsReport()
        nx, ny, nz = self.grid_shape()
        buf = np.zeros(6 * nx * ny * nz, dtype=np.uint8) if counts else None
        fn = self._v.vgpu_check_limits_noalloc if noalloc else self._v.vgpu_check_limits
        self._check(fn(self._c, lim, ctypes.byref(rep), buf.ctypes.data if counts else None))
        out = {k: [int(x) for x in getattr(rep, k)] for k, _ in LimitsReport._fields_}
        if counts:
            return out, buf.reshape(6, nz, ny, nx).transpose(0, 3, 2, 1)
        return out

    def grid_shape(self):
        sizes = (ctypes.c_longlong * 3)()
        load_host().vh_solve_shape(self._s, sizes)
        return tuple(int(x) for x in sizes)

    def close(self):
        if self._s:
            load_host().vh_gpu_close(self._s, self._c)
            self._s = ctypes.c_void_p()
            self._c = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

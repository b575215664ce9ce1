"""Build the native parts of vulture_amd in-tree.

Outputs (git-ignored, shipped with the tree to the GPU host):
  vulture_amd/lib/libvgpu.so     HIP kernels + C-ABI (include/vgpu.h), gfx950
  vulture_amd/lib/libvhost.so    mesh parser / set-up / device driver (C)
  vulture_amd/bin/vgvulture      command-line solver

hipcc cross-compiles for gfx950 without a GPU present.  Every unit is
compiled with -ffp-contract=off: each path performs the reference's
operations one by one, rounded once each, so the FP64 kernels agree with
the FP64 restatement to the last bits and the FP32 kernels round the field
updates as the reference's own float build does (the kernels are bound by
memory traffic, so fused multiply-adds would buy no time).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
OBJ = os.path.join(HERE, "_build")

ARCH = os.environ.get("VULTURE_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are kept out of all device code.
HIP_COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I" + os.path.join(ROOT, "include"),
              "-I" + CSRC, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
C_COMMON = ["-std=gnu99", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-result",
            "-Wno-stringop-truncation", "-Wno-format-truncation", "-I" + HOST, "-I" + os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def _newer(target, deps):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build(verbose=False, force=True):
    """Compile every native component (always recompiles when force=True)."""
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + \
        [os.path.join(ROOT, "include", "vgpu.h")]
    units = [("vgpu.hip", ["-ffp-contract=off"]), ("vgpu_k64.hip", ["-ffp-contract=off"]),
             ("vgpu_k32.hip", ["-ffp-contract=off"])]
    objs = []
    for src, extra in units:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        if force or not _newer(o, [s] + hdrs):
            _run([HIPCC] + HIP_COMMON + extra + ["-c", s, "-o", o], verbose)
        objs.append(o)
    libvgpu = os.path.join(LIB, "libvgpu.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", libvgpu] + objs, verbose)

    hsrc = ["vh_parse.c", "vh_setup.c", "vh_voxel.c", "vh_observe.c", "vh_gpu.c", "vh_serve.c"]
    hobjs = []
    for src in hsrc:
        s = os.path.join(HOST, src)
        o = os.path.join(OBJ, src.replace(".c", ".o"))
        _run(["gcc"] + C_COMMON + ["-DVH_REAL=double", "-c", s, "-o", o], verbose)
        hobjs.append(o)
    libvhost = os.path.join(LIB, "libvhost.so")
    _run(["gcc", "-shared", "-o", libvhost] + hobjs + ["-L" + LIB, "-lvgpu", "-Wl,-rpath,$ORIGIN", "-lm", "-lpthread"],
         verbose)
    cli = os.path.join(BIN, "vgvulture")
    _run(["gcc"] + C_COMMON + [os.path.join(HOST, "vgvulture_main.c"), "-o", cli, "-L" + LIB, "-lvhost", "-lvgpu",
                               "-Wl,-rpath,$ORIGIN/../lib", "-lm"], verbose)
    return {"libvgpu": libvgpu, "libvhost": libvhost, "cli": cli}


def build_oracle(verbose=False):
    """Build the CPU oracle (and, where /root/reference exists, the reference)."""
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "-B" if False else "all"]
    _run(cmd, verbose)


if __name__ == "__main__":
    out = build(verbose=True)
    build_oracle(verbose=True)
    print(out)

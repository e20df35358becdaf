"""Build the engine in-tree: libccgpu.so (HIP for gfx950) and the Fortran host
program ccm_gpu (AMD flang, iso_c_binding over the C ABI)."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "libccgpu.so")
EXE = os.path.join(HERE, "host", "ccm_gpu")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
SOURCES = ["plan.cpp", "chiral.cpp", "kernels.hip", "triples.hip", "mlp.hip", "noalloc.hip", "api.cpp"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + res.stdout[-4000:] + res.stderr[-4000:])
    return res


def build_library():
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """Fortran drop-in driver; skipped (returns None) when flang is absent."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(HOST, "_mod")
    os.makedirs(mod, exist_ok=True)
    srcs = [os.path.join(HOST, f) for f in ("ccgpu_iface.f90", "ccm_gpu.f90")]
    _run([FLANG, "-O2", "-module-dir", mod, *srcs, "-o", EXE, "-L" + HERE, "-lccgpu",
          "-Wl,-rpath," + HERE])
    return EXE


def build_oracle():
    """The oracle's compiled reference (test infrastructure): only where the
    reference sources exist."""
    if not os.path.isdir("/root/reference/deltafull-CCMkspace"):
        return None
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"])
    return os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def build_all():
    out = {"lib": build_library()}
    out["host"] = build_host()
    out["oracle"] = build_oracle()
    return out


if __name__ == "__main__":
    print(build_all())

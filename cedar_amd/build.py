"""Builds the in-tree HIP library cedar_amd/lib/libcedar_hip.so for gfx950 with hipcc.

    python -m cedar_amd.build [--incremental]       (or __graft_entry__.build())

Flags: -O3, -ffp-contract=off (no fused multiply-add: the FP64 kernels keep the reference's
rounding order), and the packed-FP32 instructions disabled for the device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcedar_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["bmg2_kernels.hip", "capi2d.hip", "solver2d.cpp", "bmg3_kernels.hip", "capi3d.hip",
           "solver3d.cpp", "adjoint2d.hip", "matvec.hip", "capi_handles.cpp", "noalloc2d.hip"]
ARCH = os.environ.get("CEDAR_AMD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-mcode-object-version=5", "-mllvm", "-pragma-unroll-threshold=1000000",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def _headers_mtime():
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    inc = os.path.join(os.path.dirname(HERE), "include")
    hdrs += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in hdrs)


def _compile(src, incremental=False):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    if incremental and os.path.exists(obj):
        newest = max(os.path.getmtime(os.path.join(CSRC, src)), _headers_mtime())
        if os.path.getmtime(obj) > newest:
            return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    # the host pass does not know the packed-fp32 target feature; that warning is expected
    noise = [ln for ln in r.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(verbose=True, incremental=False):
    """Compiles every source (incremental=True skips objects newer than their source and every
    header) and links the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 4)) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)

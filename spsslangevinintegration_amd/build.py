"""Build libspss_hip.so (the HIP kernels + C ABI) for gfx950 with hipcc.

    python spsslangevinintegration_amd/build.py        # or __graft_entry__.build()

Every .hip file under csrc/ is compiled to an object in _build/ (in parallel)
and linked into spsslangevinintegration_amd/libspss_hip.so; cpp/shim_check (the
reference-named C++ classes driven like the reference's engine) links against it.  All device code is
compiled with packed-FP32 instructions disabled and FP contraction off (FMAs
are written explicitly where a kernel wants them).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspss_hip.so")
ARCH = os.environ.get("SPSS_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the C++ drop-in check program over the reference-named classes (host code, g++)
    cpp = os.path.join(HERE, "cpp")
    exe = os.path.join(cpp, "shim_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(cpp, "shim_check.cpp"), "-L" + HERE, "-lspss_hip", "-Wl,-rpath,$ORIGIN/..", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on shim_check.cpp:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", LIB, "and", exe, file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()

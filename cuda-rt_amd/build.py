"""Build librtamd.so (HIP for gfx950 + host C++) in-tree.

Usage: python cuda-rt_amd/build.py
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librtamd.so")
CLI = os.path.join(HERE, "rtamd_cli")

HIP_SOURCES = ["rt_build.hip", "rt_trace.hip", "rt_api.hip", "rt_batch.hip"]
CXX_SOURCES = ["rt_scenes.cpp", "rt_io.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RT_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FP64 parity arithmetic must be evaluated exactly as
# written. The packed-fp32 feature is disabled for every device compile.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o], o))
    for s in CXX_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append((["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o], o))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in jobs]
        for f in futs:
            r = f.result()
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp", "-o", tmp,
          *[o for _, o in jobs]])
    os.replace(tmp, LIB)
    # command-line renderer over the C++ facade (host code only)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-o", tmp, os.path.join(CSRC, "rt_cli.cpp"), "-L" + HERE, "-lrtamd",
          "-Wl,-rpath,$ORIGIN", "-Wno-unused-command-line-argument"])
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

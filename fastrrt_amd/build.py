"""Build libfastrrt.so (HIP, gfx950) in-tree: fastrrt_amd/lib/libfastrrt.so.

The kernels are compiled with FP contraction off (the device must round
every a*b+c twice, like the CPU oracle and the reference's operation order)
and without packed-FP32 instruction selection.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libfastrrt.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["fr_kernels.hip", "fr_apf.hip", "fr_frame.hip", "fr_engine.cpp", "fr_api.cpp", "fr_batch.cpp"]

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]
DEVICE = [
    "--offload-arch=gfx950", "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libfastrrt.so")


def build(verbose=False, jobs=4):
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs, procs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        host = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++"]
        lang = DEVICE + ["-x", "hip"] if src.endswith(".hip") else host
        cmd = [cc, "-c"] + COMMON + lang + [path, "-o", obj]
        objs.append(obj)
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError("libfastrrt build failed:\n" + msg)
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-Wl,-soname,libfastrrt.so.1"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""Build libcufhe_amd.so in-tree: HIP kernels for gfx950 + host runtime.

Usage: python cufhe_amd/build.py [--jobs N]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcufhe_amd.so")

HIP_SOURCES = ["kernels_fft.hip", "kernels_ntt.hip", "kernels_ntt64.hip", "kernels_annihilate.hip",
               "ks_gemm.hip", "runtime.hip", "server.hip"]
CXX_SOURCES = ["tfhe_host.cpp", "circuits_host.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops are disabled on purpose (see the build notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "unknown target feature" not in l and "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs +
         ["-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    # C++ drop-in tests (tests/cpp/*.cpp), linked against the library in-tree
    root = os.path.dirname(HERE)
    cpp_dir = os.path.join(root, "tests", "cpp")
    srcs = sorted(f for f in os.listdir(cpp_dir) if f.endswith(".cpp")) if os.path.isdir(cpp_dir) else []
    cmds = [[HIPCC, "-O2", "-std=c++17", "-I", os.path.join(root, "include"),
             os.path.join(cpp_dir, f), "-o", os.path.join(cpp_dir, f[:-4] + ".bin"),
             "-L", HERE, "-lcufhe_amd", "-Wl,-rpath,$ORIGIN/../../cufhe_amd"] for f in srcs]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))

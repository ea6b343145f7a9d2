"""Builds libsph_amd.so (gfx950 HIP kernels + C++ host layer + C-ABI) in-tree.

Usage: python build.py [--arch gfx950] [--jobs N]

The device code is compiled by hipcc for gfx950 only; host C++ by g++.  All HIP
code is compiled with packed FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsph_amd.so")
LIB_ASAN = os.path.join(HERE, "libsph_amd_asan.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["sph_device.hip", "sph_mlp.hip", "sph_noalloc.hip"]
CXX_SOURCES = ["sph_host.cpp", "sph_capi.cpp"]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return res.stdout


def build(arch="gfx950", jobs=4, verbose=False, sanitize=False):
    """Builds libsph_amd.so; with `sanitize`, libsph_amd_asan.so instead, whose
    host C++ (input, set-up, writers, C-ABI) is compiled by g++ with
    AddressSanitizer and UndefinedBehaviorSanitizer (the device objects are the
    normal ones)."""
    os.makedirs(BUILD, exist_ok=True)
    hip_flags = [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={arch}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
    ]
    cxx_flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
    if sanitize:
        cxx_flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                     "-fno-sanitize-recover=undefined"] + cxx_flags[1:]
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC] + hip_flags + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + (".asan.o" if sanitize else ".o"))
        cmds.append(["g++"] + cxx_flags + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            sys.stdout.write(o)
    lib = LIB_ASAN if sanitize else LIB
    tmp = lib + ".tmp"
    link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={arch}", "-o", tmp] + objs
    if sanitize:
        link += ["-Xarch_host", "-fsanitize=address,undefined"]
    _run(link)
    os.replace(tmp, lib)
    return lib


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--sanitize", action="store_true",
                    help="build libsph_amd_asan.so (host C++ under ASan + UBSan)")
    a = ap.parse_args()
    if a.clean:
        clean()
    print(build(a.arch, a.jobs, a.verbose, a.sanitize))

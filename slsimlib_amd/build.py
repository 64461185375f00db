"""Builds libslsim_amd.so in-tree with hipcc for gfx950.

Usage: python slsimlib_amd/build.py  (run by path: importing the package
first would load the library that is being rebuilt)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libslsim_amd.so")

SOURCES = ["tree_build.hip", "shoot.hip", "api.hip", "batch.hip", "smoothing.hip", "generator.cpp",
           "formats.cpp", "halos.hip", "images.hip", "crit.hip", "agrid.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=4):
    """Compiles every source to an object, then links the shared library."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [hipcc] + HIPCC_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if out.returncode != 0:
        raise RuntimeError("link failed:\n" + out.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


def build_cpp_example(verbose=False):
    """Compiles tests/cpp/dropin_gridmap.cpp, dropin_formats.cpp and dropin_halos.cpp,
    reference-style C++ programs against include/glamer_amd.h linked with
    the library."""
    root = os.path.dirname(HERE)
    exes = []
    for name in ("dropin_gridmap", "dropin_formats", "dropin_halos", "dropin_boundary",
                 "dropin_grid"):
        src = os.path.join(root, "tests", "cpp", name + ".cpp")
        exe = os.path.join(root, "tests", "cpp", name)
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src,
               "-L", HERE, "-lslsim_amd", "-Wl,-rpath,$ORIGIN/../../slsimlib_amd", "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if out.returncode != 0:
            raise RuntimeError("C++ example build failed:\n" + out.stdout.decode(errors="replace"))
        exes.append(exe)
    return exes


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError("compile of %s failed:\n%s" % (src, out.decode(errors="replace")))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_example(verbose="-v" in sys.argv))

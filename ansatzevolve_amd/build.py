"""Builds libansatzsynth.so in-tree: HIP sources with hipcc for gfx950, the
host C++ with g++, linked by hipcc.  Run: python -m ansatzevolve_amd.build"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libansatzsynth.so")
ARCH = os.environ.get("AE_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "lanes.hip", "window.hip", "sigg.hip", "engine.hip", "mlp.hip", "multi.hip"]
# per-file device flags: the lane sweeps keep their orbit arrays as separate
# registers (no promotion of private arrays to one vector register tuple)
HIP_FILE_FLAGS = {"lanes.hip": ["-mllvm", "-disable-promote-alloca-to-vector=1"],
                  "window.hip": ["-mllvm", "-disable-promote-alloca-to-vector=1"]}
CXX_SOURCES = ["manager.cpp", "newton.cpp", "projected.cpp", "spectral.cpp", "autotune.cpp", "abi.cpp", "runfiles.cpp"]
# command-line front end over the C ABI (cli.cpp), beside the library
CLI = os.path.join(HERE, "ansatzsynth")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             # packed FP32 instructions stay off on this hardware
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    # the grouped-target sigma programs are generated (sigg.py)
    from . import sigg
    sigg.emit()
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        jobs.append([HIPCC] + HIP_FLAGS + HIP_FILE_FLAGS.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    _run(["g++"] + CXX_FLAGS[:-1] + ["-I", os.path.join(HERE, "..", "include"), os.path.join(CSRC, "cli.cpp"),
                                     "-o", tmp, "-L" + HERE, "-lansatzsynth", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

"""Build recipe for the native libraries (run by __graft_entry__.build()).

Products (in-tree, next to this file):
  libswg.so         HIP kernels for gfx950 + the C ABI of include/swg.h
  libsalsa_host.so  C++ mirror of the reference's host-side block
                    decomposition, matrix writers and synthetic-data generator
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # the distance epilogue is defined by per-operation IEEE binary32 rounding
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include")]

HIP_SOURCES = ["swg_kernels.hip", "swg_api.hip"]
HOST_SOURCES = ["salsa_host.cpp", "synth.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj])
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj])
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    libswg = os.path.join(HERE, "libswg.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", libswg + ".tmp"]
         + [os.path.join(BUILD, s + ".o") for s in HIP_SOURCES], verbose)
    os.replace(libswg + ".tmp", libswg)
    libhost = os.path.join(HERE, "libsalsa_host.so")
    _run(["g++", "-shared", "-fPIC", "-o", libhost + ".tmp"]
         + [os.path.join(BUILD, s + ".o") for s in HOST_SOURCES]
         + ["-L" + HERE, "-lswg", "-Wl,-rpath,$ORIGIN", "-lpthread"], verbose)
    os.replace(libhost + ".tmp", libhost)
    return libswg, libhost


if __name__ == "__main__":
    build(verbose=True)

"""Build recipe for libselm_hip.so (gfx950) and the C++ driver executable.

Everything is compiled in-tree with hipcc; no JIT caches.  ``python -m
monomers2026_amd.build`` rebuilds from scratch.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libselm_hip.so")
HOST_LIB = os.path.join(PKG, "libselm_host.so")
DRIVER = os.path.join(PKG, "selm_hip_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-Wno-unused-command-line-argument"]

LIB_SOURCES = ["selm_kernels.hip", "selm_fft_kernels.hip", "selm_analysis.hip", "selm_engine.cpp"]
# host-only support library: Model.SELM reader, LAMMPS-subset deck, host
# force terms, DCD writer (no GPU code; testable without a GPU)
HOST_SOURCES = ["host/lammps_subset.cpp", "host/dcd_writer.cpp", "host/selm_host_api.cpp"]
# the deck driver: Selm / Eulerian_Period_HIP / Lagrangian_LAMMPS_HIP over both libraries
DRIVER_SOURCES = ["host/selm_main.cpp", "host/SELM.cpp"]
HOST_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-I", INCLUDE, "-I", os.path.join(CSRC, "host")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)


def build_lib(verbose=False):
    """Each source is compiled to an object in parallel, then linked."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o") for src in LIB_SOURCES]
    cmds = [[HIPCC] + DEVICE_FLAGS + ["-c", "-I", INCLUDE, "-I", CSRC, "-o", o, src]
            for src, o in zip(LIB_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC] + DEVICE_FLAGS + ["-shared", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_host(verbose=False):
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++"] + HOST_FLAGS + ["-fPIC", "-shared", "-o", tmp] + HOST_SOURCES
    _run(cmd, verbose)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build_driver(verbose=False):
    cmd = ["g++"] + HOST_FLAGS + ["-o", DRIVER] + DRIVER_SOURCES + [
        "-L", PKG, "-lselm_host", "-lselm_hip", "-Wl,-rpath,$ORIGIN",
        "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    _run(cmd, verbose)
    return DRIVER


def build(verbose=False):
    build_lib(verbose)
    build_host(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)

"""Build step of the package: compiles the HIP engine for gfx950 into
radlite_amd/librl.so (in-tree, so it travels with the source tree).

    python -m radlite_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LIB = os.path.join(HERE, "librl.so")
SOURCES = ["csrc/rl_engine.hip", "csrc/rl_prepass.hip", "csrc/rl_resample.hip",
           "csrc/rl_scratch.hip", "csrc/rl_api.cpp"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-pthread",
    # FP64 parity mode rounds every a*b+c as two operations, like the
    # reference on x86-64; FP32 code uses explicit intrinsics instead.
    "-ffp-contract=off",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_HOST = os.path.join(FORTRAN_DIR, "rl_host")


def build_fortran_host(verbose=False):
    """The Fortran host program (ISO_C_BINDING module + driver), linked
    against librl.so with an rpath to it."""
    import tempfile
    with tempfile.TemporaryDirectory() as mod:
        cmd = [FLANG, "-O2", "-module-dir", mod,
               os.path.join(FORTRAN_DIR, "radlite_gpu.f90"),
               os.path.join(FORTRAN_DIR, "rl_host.f90"),
               "-o", FORTRAN_HOST + ".tmp", "-L" + HERE, "-lrl",
               "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed building the Fortran host")
    os.replace(FORTRAN_HOST + ".tmp", FORTRAN_HOST)
    return FORTRAN_HOST


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I" + os.path.join(ROOT, "include"),
                             "-I" + os.path.join(HERE, "csrc")] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librl.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran_host(verbose=True))

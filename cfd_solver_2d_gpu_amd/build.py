"""Builds the in-tree shared library libfvmr.so (HIP for gfx950 + host C++).

Usage: python -m cfd_solver_2d_gpu_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfvmr.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_*_f32 generation (they have given wrong results on this platform).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose=True, defines=(), lib_path=None):
    """Compiles csrc/fvmr_solver.hip with hipcc and csrc/*.cpp with g++, links libfvmr.so.

    defines: extra -D tuning knobs for the device code; lib_path: output path."""
    lib_path = lib_path or LIB
    tag = "_".join(d.replace("=", "-") for d in defines) if defines else "default"
    out_dir = os.path.join(HERE, "_build", tag)
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for name in ("fvmr_mesh", "fvmr_tiles"):
        obj = os.path.join(out_dir, name + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, name + ".cpp"), "-o", obj])
        objs.append(obj)
    solver_o = os.path.join(out_dir, "fvmr_solver.o")
    _run([HIPCC, *HIP_FLAGS, *["-D" + d for d in defines], "-c", os.path.join(CSRC, "fvmr_solver.hip"), "-o", solver_o])
    objs.append(solver_o)
    tmp = lib_path + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib_path)
    return lib_path


C_API_DIR = os.path.join(os.path.dirname(HERE), "tests", "c_api")
C_CALLER = os.path.join(HERE, "_build", "fvmr_c_caller")      # include/fvmr.h only
CPP_MIRROR = os.path.join(HERE, "_build", "fvmr_cpp_mirror")  # include/fvmr.hpp


def build_c_caller(out=None, name="fvmr_c_caller"):
    """Compiles a caller of the C ABI from tests/c_api (fvmr_c_caller: include/fvmr.h only;
    fvmr_cpp_mirror: include/fvmr.hpp) against libfvmr.so; it finds the library from its own
    directory (rpath $ORIGIN/...)."""
    out = os.path.abspath(out or os.path.join(HERE, "_build", name))
    os.makedirs(os.path.dirname(out), exist_ok=True)
    rpath = "$ORIGIN/" + os.path.relpath(HERE, os.path.dirname(out))
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
          "-I" + os.path.join(os.path.dirname(HERE), "include"), os.path.join(C_API_DIR, name + ".cpp"),
          "-L" + HERE, "-lfvmr", "-Wl,-rpath," + rpath, "-o", out])
    return out


if __name__ == "__main__":
    build_library()
    build_c_caller()
    build_c_caller(name="fvmr_cpp_mirror")
    sys.exit(0)

"""Build the gfx950 HIP library in-tree with hipcc (no JIT cache)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")

# -ffp-contract=off: FP64 kernels round exactly like the CPU reference order;
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIPCC_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
               "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result"]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libuaamg_hip.so")
    tmp = out + ".tmp"
    cmd = ["hipcc"] + HIPCC_FLAGS + ["-o", tmp, os.path.join(CSRC, "uaamg.hip")]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


ROOT = os.path.dirname(HERE)
BUILD_DIR = os.path.join(ROOT, "build")


def build_host_tools(verbose=False):
    """Compile the C++ host-mirror test program against the built library."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    out = os.path.join(BUILD_DIR, "host_solver_test")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "host_solver_test.cpp"),
           "-L", LIBDIR, "-luaamg_hip", "-Wl,-rpath,$ORIGIN/../uaamg_amd/lib",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return out

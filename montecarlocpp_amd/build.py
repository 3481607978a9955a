"""Build the engine library in-tree: montecarlocpp_amd/libmcg.so (gfx950).

    python -m montecarlocpp_amd.build

The HIP translation unit is compiled by hipcc for gfx950 only; the host
scene builder is plain C++ compiled by g++.  Both use -ffp-contract=off so
that the FP64 path performs the same separately rounded operations as the
CPU checker, which deterministic-mode parity relies on.  Packed FP32 VALU
instructions are disabled for the device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmcg.so")
DRIVER = os.path.join(HERE, "api", "mcg_main")
ROOT = os.path.dirname(HERE)

ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return path


def build(verbose=False, save_temps=False):
    os.makedirs(BUILD, exist_ok=True)
    host_obj = os.path.join(BUILD, "mcg_host.o")
    kern_obj = os.path.join(BUILD, "mcg_kernels.o")
    _run(["g++", *COMMON, "-c", os.path.join(CSRC, "mcg_host.cpp"), "-o", host_obj])
    extra = ["--save-temps"] if save_temps else []
    _run([hipcc(), *COMMON, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, "mcg_kernels.hip"),
          "-o", kern_obj])
    tmp = LIB + ".tmp"
    _run([hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp, kern_obj, host_obj])
    os.replace(tmp, LIB)
    # the reference-compatible command-line driver over the C ABI
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
          os.path.join(HERE, "api"), os.path.join(HERE, "api", "mcg_main.cpp"), "-L", HERE, "-lmcg",
          "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER])
    if verbose:
        print(f"built {LIB} and {DRIVER}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, save_temps="--save-temps" in sys.argv)

"""Build libmpm_hip.so (gfx950) in-tree with hipcc.

The library holds every HIP kernel of the MPM substep and the C ABI of include/mpm_c_abi.h.
It is written next to this file so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpm_hip.so")
SOURCES = ["mpm_kernels.hip", "mpm_mlp.hip", "mpm_noalloc.hip", "mpm_runtime.hip"]
# programs over the C++ host API (include/mpm/*.h), linked against the library
EXAMPLES = ["sand_column_example", "cloth_example"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmpm_hip.so")
    os.replace(tmp, LIB)
    for name in EXAMPLES:
        exe = os.path.join(HERE, name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I/opt/rocm/include", os.path.join(CSRC, "host", name + ".cpp"),
               "-L", HERE, "-lmpm_hip", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("g++ failed building " + name)
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

"""Build the CPU oracle (test infrastructure only) into oracle/_build/libmpm_oracle.so.

The oracle is a C++ restatement of the reference substep (mpm_oracle.cpp). The reference itself
cannot be compiled here: it needs Eigen 3 and libigl, neither of which is present, so there is no
oracle/_ref build.

Flags: -O3 with a portable x86-64-v3 target so that the library built here also runs on the host
CPU of the GPU machine. OpenMP is not used: the reference solver is single-threaded.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmpm_oracle.so")


def build(verbose=False, source=None, out=None, opt="-O3"):
    """Compile the oracle (or another copy of its source, `source`, into `out`)."""
    lib = out or LIB
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    tmp = lib + ".tmp"
    cmd = ["g++", opt, "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared",
           "-I", os.path.join(os.path.dirname(HERE), "include"),
           source or os.path.join(HERE, "mpm_oracle.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the oracle")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))

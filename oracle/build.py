"""Build the CPU oracle (test infrastructure only) into oracle/_build/libmpm_oracle.so.

The oracle is a C++ restatement of the reference substep (mpm_oracle.cpp). The reference itself
cannot be compiled here: it needs Eigen 3 and libigl, neither of which is present, so there is no
oracle/_ref build.

Flags: -O3 with a portable x86-64-v3 target so that the library built here also runs on the host
CPU of the GPU machine. The default library is serial, as the reference solver is; a second build
with OpenMP (oracle/_build/libmpm_oracle_omp.so) times the CPU path on all host cores.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmpm_oracle.so")
LIB_OMP = os.path.join(OUT_DIR, "libmpm_oracle_omp.so")


def build(verbose=False, source=None, out=None, opt="-O3", openmp=False):
    """Compile the oracle (or another copy of its source, `source`, into `out`). The default call
    builds both the serial library and the OpenMP one."""
    if source is None and out is None and not openmp:
        build(verbose, opt=opt, openmp=True)
    lib = out or (LIB_OMP if openmp else LIB)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    tmp = lib + ".tmp"
    cmd = ["g++", opt, "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared"] + (["-fopenmp"] if openmp else []) + [
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

"""mpm_surrogate_gradient_noalloc (device pointers, fixed library scratch, no allocation) against
mpm_surrogate_gradient on the same inputs and weights: gradient and loss equal bit for bit, on two
calls in a row, the second with a smaller batch (the scratch keeps the first call's contents); a
batch above the 2048-row maximum is refused with a clear error. Device buffers for the test come
from the HIP runtime the library itself loaded (ctypes), not from another runtime."""
import ctypes
import sys

import numpy as np
import pytest

from anisotropicelastoplasticity_amd import Solver, _lib
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hip():
    _lib.load()  # the library (and through it libamdhip64) is loaded first
    hip = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    inp = np.empty((n, 5), np.float32)
    inp[:, :3] = np.sort(rng.normal(0.0, 0.01, size=(n, 3)), axis=1)[:, ::-1]
    inp[:, 3] = rng.uniform(25.0, 40.0, size=n)
    inp[:, 4] = rng.uniform(1.5, 3.5, size=n)
    H = (inp[:, :3] + rng.normal(0.0, 2e-4, size=(n, 3))).astype(np.float32)
    return inp, H


def test_gradient_noalloc_matches_allocating_entry():
    hip = _hip()
    with Solver() as s:
        s.surrogate_init(7)
        npar = s.surrogate_nparam()
        n_max = 2048
        d_in, d_H = DevBuf(hip, 5 * 4 * n_max), DevBuf(hip, 3 * 4 * n_max)
        d_g, d_l = DevBuf(hip, 4 * npar), DevBuf(hip, 4)
        try:
            for call, n in enumerate((2048, 777)):
                inp, H = _inputs(n, 10 + call)
                say(f"[no-alloc gradient] call {call}, n={n}")
                g_ref, l_ref = s.surrogate_gradient(inp, H)
                d_in.put(inp)
                d_H.put(H)
                rc = s.lib.mpm_surrogate_gradient_noalloc(s.h, d_in.p, d_H.p, n, d_g.p, d_l.p)
                assert rc == 0, s.lib.mpm_last_error(s.h)
                s.synchronize()
                g = d_g.get(npar)
                l = float(d_l.get(1)[0])
                say(f"  loss {l:.6e} (allocating entry {l_ref:.6e}); max |dg| {np.max(np.abs(g - g_ref)):.3e}")
                assert np.array_equal(g.view(np.uint32), np.asarray(g_ref, np.float32).view(np.uint32))
                assert np.float32(l).view(np.uint32) == np.float32(l_ref).view(np.uint32)
                assert np.all(g[~mo.active_mask()] == 0.0) and np.any(g != 0.0)
            rc = s.lib.mpm_surrogate_gradient_noalloc(s.h, d_in.p, d_H.p, n_max + 1, d_g.p, d_l.p)
            msg = s.lib.mpm_last_error(s.h)
            msg = msg.decode() if isinstance(msg, bytes) else str(msg)
            say(f"  n = {n_max + 1}: rc {rc}, {msg}")
            assert rc == _lib.ERR_INVALID and "2048" in msg
        finally:
            for b in (d_in, d_H, d_g, d_l):
                b.free()

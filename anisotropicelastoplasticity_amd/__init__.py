"""MI355X-native explicit MPM substep (snow, Drucker-Prager sand) behind the C ABI of
include/mpm_c_abi.h. The compute path is libmpm_hip.so (hand-written gfx950 HIP kernels); the
Python modules here only bind and drive it."""
from . import _lib  # noqa: F401
from .api import MPMError, Solver, generate_block, generate_shape, levelset  # noqa: F401

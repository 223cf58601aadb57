// mpm_noalloc.h -- layout of the fixed scratch of the allocation-free surrogate gradient (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_mlp.h"

namespace mpm {

// Largest batch of mpm_surrogate_gradient_noalloc.
constexpr int NOALLOC_MAX_ROWS = 2048;
// workgroups of its backward pass (64 rows each), each with a partial gradient
constexpr int NOALLOC_GROUPS = (NOALLOC_MAX_ROWS + 63) / 64;
// scratch (floats): normalised targets (3 per row), outputs (3 per row), pre-activations
// (4 layers x 128 per row), partial gradients and losses of the backward workgroups
constexpr size_t NOALLOC_TGT = 0;
constexpr size_t NOALLOC_Y = NOALLOC_TGT + 3 * (size_t)NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_Z = NOALLOC_Y + 3 * (size_t)NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_GPART = NOALLOC_Z + 4 * (size_t)MLP_HIDDEN * NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_LPART = NOALLOC_GPART + (size_t)NOALLOC_GROUPS * MLP_NPARAM;
constexpr size_t NOALLOC_SCRATCH_FLOATS = NOALLOC_LPART + NOALLOC_GROUPS;

hipError_t noalloc_scratch(int device, float** ptr);

}  // namespace mpm

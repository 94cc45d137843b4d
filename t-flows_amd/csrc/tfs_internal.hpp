// tfs_internal.hpp -- what the library's other translation units need of a
// tfs_matrix beyond the C ABI (C++ linkage, not exported by the headers).
#pragma once

#include <hip/hip_runtime.h>

#include "tfs_solver.h"

namespace tfs {
// The matrix's own stream: work queued on it is ordered with its solves.
hipStream_t matrix_stream(const tfs_matrix *A);
// Workgroups of the heaviest persistent sweep kernel that fit on the GPU at
// once, and a cap on one matrix's persistent sweep grid (at least 1).
int sweep_capacity(const tfs_matrix *A);
void set_sweep_grid(tfs_matrix *A, int groups);
}  // namespace tfs

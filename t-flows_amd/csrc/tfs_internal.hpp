// tfs_internal.hpp -- what the library's other translation units need of a
// tfs_matrix beyond the C ABI (C++ linkage, not exported by the headers).
#pragma once

#include <hip/hip_runtime.h>

#include "tfs_solver.h"

namespace tfs {
// The matrix's own stream: work queued on it is ordered with its solves.
hipStream_t matrix_stream(const tfs_matrix *A);
}  // namespace tfs

// gc_dict.hpp -- dictionary search of the NEXI likelihood start
// (NEXI/gpuNEXI.m:822-849): for each voxel, the dictionary entry whose
// scaled copy fits the voxel's signal best.
#pragma once
#include "gc_host.hpp"

namespace gc {

constexpr int kDictMaxMeas = 32;  // measurements per voxel (8 MFMA k-steps of 4)

// y [m][nv] (voxel fastest), dict [m][nd], w [m] host weights or nullptr
// (ones); best [nv] (0-based entry, -1 when no entry has a finite score),
// sse [nv] or nullptr.  All pointers but w are device pointers.
int dict_search(gc_ctx* ctx, const float* y, int64_t nv, int m, const float* dict, int64_t nd, const float* w,
                int32_t* best, double* sse);

}  // namespace gc

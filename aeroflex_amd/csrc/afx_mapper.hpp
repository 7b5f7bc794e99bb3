// Launchers of the level-prolongation kernels (afx_mapper.hip), FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace afx {

// per tile of 256 coarse cells: bounding box of the centroids grown by sqrt(2 A)
hipError_t launch_map_tile_boxes(hipStream_t s, const double* cx, const double* cy, const double* ca, int32_t nc,
                                 double* box);
hipError_t launch_map_count(hipStream_t s, const double* fx, const double* fy, int32_t nf, const double* cx,
                            const double* cy, const double* ca, int32_t nc, const double* box, int32_t* count);
hipError_t launch_map_fill(hipStream_t s, const double* fx, const double* fy, int32_t nf, const double* cx,
                           const double* cy, const double* ca, int32_t nc, const double* box, const int64_t* rowptr,
                           int32_t* cols, double* vals);
hipError_t launch_map_cols(hipStream_t s, const int32_t* cols, const int32_t* old2new, int64_t nnz, int32_t* cols_dev);
hipError_t launch_map_prolong(hipStream_t s, const int64_t* rowptr, const int32_t* fnew2old, const int32_t* cols_dev,
                              const double* vals, const double* qc, int64_t ntc, double* qf, int32_t ntf, int nv);

}  // namespace afx

// bsar_mixed.h -- argument blocks of the mixed-precision back-projection.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsar_kernels.h"

namespace bsar {

hipError_t launch_r1_table(const R1Table& tab, const double* ax, const double* ay, const double* az, int n,
                           double* out, hipStream_t s);
hipError_t launch_bp_mixed(const BpArgs& a, const MixedArgs& m, hipStream_t s);
hipError_t launch_group_sum_mixed(const float4* acc, int groups, int64_t stride, int64_t nvox, double2* out,
                                  hipStream_t s);
hipError_t fill_zero4(float4* p, int64_t n, hipStream_t s);

}  // namespace bsar

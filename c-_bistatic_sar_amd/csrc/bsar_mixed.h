// bsar_mixed.h -- argument blocks of the mixed-precision back-projection.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsar_kernels.h"

namespace bsar {

// Lattice of the mean-distance tables: node (i, j, k) sits at
// (x0 + i h, y0 + j h, z0 + k h), i < nx, j < ny, k < nz; values [k][j][i].
struct R1Table {
    double x0, y0, z0, h, ih;
    int nx, ny, nz;
};

struct MixedArgs {
    R1Table tab;
    const double* ftx;               // Tx mean-distance table
    const double* frx;               // Rx mean-distance table
    float4* acc;                     // [groups][acc_stride] (sum.re, sum.im, comp.re, comp.im)
    unsigned long long* exact_count; // R1 evaluations that fell outside the table
    double two_pi, inv_two_pi;
};

hipError_t launch_r1_table(const R1Table& tab, const double* ax, const double* ay, const double* az, int n,
                           double* out, hipStream_t s);
hipError_t launch_bp_mixed(const BpArgs& a, const MixedArgs& m, hipStream_t s);
hipError_t launch_group_sum_mixed(const float4* acc, int groups, int64_t stride, int64_t nvox, double2* out,
                                  hipStream_t s);
hipError_t fill_zero4(float4* p, int64_t n, hipStream_t s);

}  // namespace bsar

// bsar_kernels.h -- argument blocks and host launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsar_device.h"

namespace bsar {

// DEMs up to this many 8-byte slots (X + Y + padded Z in FP64, then the FP32
// (height, x difference) pairs of Z) are staged in LDS per workgroup
// (DEM_LDS); up to kSplitMax
// slots of X, Y and the FP32 heights alone are staged with two workgroups per
// CU, the FP64 heights read from global memory (DEM_SPLIT); larger DEMs are
// read through the L1/L2 caches (DEM_GLOBAL).
constexpr int kLdsDemMax = 8192;   // 64 KiB
constexpr int kSplitMax = 9728;    // 76 KiB

struct DemPlan {
    int mode;       // DemMode
    size_t bytes;   // dynamic LDS
};
inline DemPlan plan_dem(const DemDev& d)
{
    if (dem_lds_slots(d) <= kLdsDemMax) return {DEM_LDS, (size_t)dem_lds_slots(d) * 8};
    if (dem_split_slots(d) <= kSplitMax) return {DEM_SPLIT, (size_t)dem_split_slots(d) * 8};
    return {DEM_GLOBAL, 0};
}

// Waves per SIMD a search kernel is compiled for: `waves` with the whole DEM
// in LDS, 2 otherwise (split mode holds two 256-thread workgroups per CU in
// LDS; measured faster than larger workgroups at 3 waves).  Workgroups are
// 256 threads in every mode.
constexpr int dem_block(int, int) { return 256; }
constexpr int dem_waves(int mode, int waves) { return mode == DEM_LDS ? waves : 2; }

inline unsigned dem_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Launches kernel template K<mode> (W waves per SIMD with the DEM in LDS)
// over n threads with the block size and dynamic LDS of plan.mode.
#define BSAR_LAUNCH_DEM(K, W, plan, n, s, ...)                                                                  \
    do {                                                                                                      \
        if ((plan).mode == DEM_LDS)                                                                            \
            hipLaunchKernelGGL(K<DEM_LDS>, dim3(dem_blocks(n, dem_block(DEM_LDS, W))),                          \
                               dim3(dem_block(DEM_LDS, W)), (plan).bytes, s, __VA_ARGS__);                      \
        else if ((plan).mode == DEM_SPLIT) {                                                                   \
            (void)hipFuncSetAttribute((const void*)K<DEM_SPLIT>, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                      (int)(plan).bytes);                                                     \
            hipLaunchKernelGGL(K<DEM_SPLIT>, dim3(dem_blocks(n, dem_block(DEM_SPLIT, W))),                      \
                               dim3(dem_block(DEM_SPLIT, W)), (plan).bytes, s, __VA_ARGS__);                    \
        }                                                                                                     \
        else                                                                                                  \
            hipLaunchKernelGGL(K<DEM_GLOBAL>, dim3(dem_blocks(n, dem_block(DEM_GLOBAL, W))),                    \
                               dim3(dem_block(DEM_GLOBAL, W)), 0, s, __VA_ARGS__);                             \
    } while (0)

// Antenna positions, structure of arrays, resident in HBM.
struct Traj {
    const double* tx; const double* ty; const double* tz;
    const double* rx; const double* ry; const double* rz;
    int n;
};

// Antennas about a local origin o for the plain form's R1 sums: per set
// (t = Tx, r = Rx) four arrays of stride entries, b = A - o per axis and
// c = |b|^2, so |A - q|^2 = c - 2 b.(q - o) + |q - o|^2 takes one add and
// three fmas per antenna.  bsq bounds every c from above; [zlo, zhi] is the
// antennas' height range.
struct R1Geo {
    const double* t;
    const double* r;
    int stride;
    double ox, oy, oz;
    double bsq;
    double zlo, zhi;
};

// One back-projection launch: thread positions [vox0, vox0 + nvox) of an
// nx*ny*nz grid (band_voxel maps a position to its voxel; accumulators and
// images are indexed by voxel, x fastest), pulse groups g = 0..groups-1 each
// covering pulses
// [g*glen, min((g+1)*glen, n)), of which this launch adds the window
// [woff, woff + wlen) to the group's running sums acc[g][v].
struct BpArgs {
    DemDev dem;
    Traj traj;
    R1Geo r1;
    const double2* rc;      // upsampled range-compressed data [ndata][S4]
    const float2* rc32;     // same in FP32 (mixed form), may be null
    int ndata, S4;
    double s4m1;            // S4 - 1 as double (window test)
    double n1, n2;
    double inv_c;           // 1.0 / c              (proc_test_script.cpp:567)
    double fs4;             // fs * upsample        (proc_test_script.cpp:223)
    double kphi;            // 2.0 * M_PI / lambda  (proc_test_script.cpp:575)
    const double* gx; const double* gy; const double* gz;
    int nx, ny, nz;
    int64_t vox0, nvox, nvox_pad;
    int groups, glen, woff, wlen;
    double2* acc;           // [groups][acc_stride]
    int64_t acc_stride;
};

// Lattice of the mean-distance tables.  The spacing h is a power of two and
// node (i, j, k) sits at ((ax + i) h, (ay + j) h, (az + k) h), i < nx, j < ny,
// k < nz; values [k][j][i].  A node's value depends only on its absolute
// lattice index, and a lookup at q uses u = q / h (exact), so any two tables
// of one spacing give bit-identical answers wherever both cover q.  Lookups
// answer only for u in [lo, hi) on every axis (the box of the image that
// uses the table); everywhere else the caller sums exactly.
struct R1Table {
    double h, ih;
    int ax, ay, az;
    int nx, ny, nz;
    double lo[3], hi[3];
};

struct MixedArgs {
    R1Table tab;
    const double* ftx;               // Tx mean-distance table
    const double* frx;               // Rx mean-distance table
    float4* acc;                     // [groups][acc_stride] (sum.re, sum.im, comp.re, comp.im)
    unsigned long long* exact_count; // R1 evaluations that fell outside the table
    double two_pi, inv_two_pi;
};

hipError_t launch_dem_interp(const DemDev& dem, const double* xy, int64_t m, double* z, hipStream_t s);
hipError_t launch_refraction(const DemDev& dem, const double* legs, int64_t m, double n1, double n2,
                             double* out, hipStream_t s);
// The exhaustive form: one wave per leg, lanes over the candidates.
hipError_t launch_refraction_wave(const DemDev& dem, const double* legs, int64_t m, double n1, double n2,
                                  double* out, hipStream_t s);
hipError_t launch_slant_range(const DemDev& dem, const Traj& tr, const double* pts, int64_t V, double n1,
                              double n2, double* ida, double* volta, hipStream_t s);
// out32 (may be null): the output rounded to complex64 as well; S * f < 2^30.
// complex64 input is widened exactly to complex128 on load.
hipError_t launch_spline(const double2* in, int S, int N, int f, double2* out, float2* out32, hipStream_t s);
hipError_t launch_spline(const float2* in, int S, int N, int f, double2* out, float2* out32, hipStream_t s);
hipError_t launch_bp_fp64(const BpArgs& a, hipStream_t s);
hipError_t launch_group_sum(const double2* acc, int groups, int64_t stride, int64_t nvox, double2* out,
                            hipStream_t s);

}  // namespace bsar

// kernels_mixed.hip -- mixed-precision back-projection (gfx950).
//
// Per voxel-pulse (proc_test_script.cpp:484-668, same quantities as the
// plain form):
//   - both refraction points: exact FP64 grid search (identical argmin);
//   - R2t, R2r, the delay sum and the phase argument 2*pi/lambda * path: FP64;
//   - R1t, R1r (the mean distance from all Tx / Rx positions to the
//     refraction point, proc_test_script.cpp:518-548): tricubic Lagrange
//     interpolation in a per-dataset FP64 table of the mean-distance
//     function over the region the refraction points occupy; exact FP64
//     sums for points outside the table;
//   - phase reduced mod 2*pi in FP64, then sin/cos, interpolation, complex
//     product and Kahan-compensated accumulation over pulses in FP32.
#include <math.h>

#include "bsar_bp.h"
#include "bsar_device.h"
#include "bsar_kernels.h"
#include "bsar_mixed.h"

namespace bsar {

// Mean distance table: value at lattice node (i, j, k), at ((ax + i) h,
// (ay + j) h, (az + k) h), is (1/N) sum_j |A_j - node| over the antenna set A.
__global__ __launch_bounds__(256) void k_r1_table(R1Table tab, const double* __restrict__ ax,
                                                  const double* __restrict__ ay, const double* __restrict__ az, int n,
                                                  double* __restrict__ out)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)tab.nx * tab.ny * tab.nz;
    if (t >= total) return;
    const int i = (int)(t % tab.nx);
    const int j = (int)((t / tab.nx) % tab.ny);
    const int k = (int)(t / ((int64_t)tab.nx * tab.ny));
    // integer times a power of two: the node position is exact, whatever the table's origin
    const double x = (double)(tab.ax + i) * tab.h, y = (double)(tab.ay + j) * tab.h, z = (double)(tab.az + k) * tab.h;
    double s = 0.0;
    for (int m = 0; m < n; ++m) {
        const double a = ax[m] - x, b = ay[m] - y, c = az[m] - z;
        s += __builtin_sqrt(a * a + b * b + c * c);
    }
    out[t] = s / (double)n;
}

template <int kMode>
__global__ __launch_bounds__(dem_block(kMode, 4), dem_waves(kMode, 4)) void k_bp_mixed(BpArgs a, MixedArgs m)
{
    extern __shared__ double lds[];
    const DemDev d = dem_stage<kMode>(a.dem, lds);
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int g = __builtin_amdgcn_readfirstlane((int)(t / a.nvox_pad));
    const int64_t vl = t - (int64_t)g * a.nvox_pad;
    if (g >= a.groups || vl >= a.nvox) return;
    const int gbeg = g * a.glen;
    const int gend = min(gbeg + a.glen, a.traj.n);
    const int p0 = gbeg + a.woff;
    const int p1 = min(p0 + a.wlen, gend);
    if (p0 >= p1) return;
    unsigned exact = 0;
    bp_mixed_thread<kMode == DEM_LDS, kMode == DEM_LDS>(d, a, m, g, a.vox0 + vl, p0, p1, exact);
    if (exact && m.exact_count) atomicAdd(m.exact_count, (unsigned long long)exact);
}

// Ordered FP64 sum of the per-group compensated FP32 sums.
// out may alias acc (group 0, same 16 bytes per voxel): each thread reads its
// voxel's groups before writing.
__global__ __launch_bounds__(256) void k_group_sum_mixed(const float4* acc, int groups, int64_t stride, int64_t nvox,
                                                         double2* out)
{
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    double r = 0.0, im = 0.0;
    for (int g = 0; g < groups; ++g) {
        const float4 x = acc[(int64_t)g * stride + v];
        r += (double)x.x - (double)x.z;
        im += (double)x.y - (double)x.w;
    }
    out[v] = make_double2(r, im);
}

__global__ void k_fill4(float4* p, int64_t n)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) p[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

static inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_r1_table(const R1Table& tab, const double* ax, const double* ay, const double* az, int n,
                           double* out, hipStream_t s)
{
    const int64_t total = (int64_t)tab.nx * tab.ny * tab.nz;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_r1_table, dim3(nb(total, 256)), dim3(256), 0, s, tab, ax, ay, az, n, out);
    return hipGetLastError();
}

hipError_t launch_bp_mixed(const BpArgs& a, const MixedArgs& m, hipStream_t s)
{
    const int64_t threads = (int64_t)a.groups * a.nvox_pad;
    if (threads <= 0) return hipSuccess;
    const DemPlan p = plan_dem(a.dem);
    BSAR_LAUNCH_DEM(k_bp_mixed, 4, p, threads, s, a, m);
    return hipGetLastError();
}

hipError_t launch_group_sum_mixed(const float4* acc, int groups, int64_t stride, int64_t nvox, double2* out,
                                  hipStream_t s)
{
    if (nvox <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_group_sum_mixed, dim3(nb(nvox, 256)), dim3(256), 0, s, acc, groups, stride, nvox, out);
    return hipGetLastError();
}

hipError_t fill_zero4(float4* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill4, dim3(nb(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace bsar

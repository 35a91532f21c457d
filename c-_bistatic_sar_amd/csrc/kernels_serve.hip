// kernels_serve.hip -- persistent back-projection kernel of the serving loop.
//
// A scheduling round hands the device one work list.  Every item is (job
// slot, pulse group, tile of 256 thread positions (band_voxel), pulse window
// [p0, p1)); the resident workgroups of each kernel variant pull that
// variant's items with an atomic counter until they are exhausted, so tiles
// of different jobs run side by side.  The per-voxel body is the one of the single-image kernels
// (bsar_bp.h): each voxel adds its pulses in increasing order with the job's
// own group split, so a job's image does not depend on what else shares the
// device.  Every workgroup exits once the counter passes the variant's end.
#include "bsar_bp.h"
#include "bsar_serve.h"

namespace bsar {

template <bool kMixed, bool kLds>
__global__ __launch_bounds__(256, kLds ? 4 : 2) void k_serve(const SrvJobDev* __restrict__ jobs, SrvList* __restrict__ wl)
{
    constexpr int kVar = 2 * (int)kMixed + (int)kLds;
    extern __shared__ double lds[];
    __shared__ int s_item[2];           // alternated: a slow wave may still read the previous one
    __shared__ const double* s_dem;     // X pointer of the DEM staged in LDS
    if (threadIdx.x == 0) s_dem = nullptr;
    const int end = (int)wl->end[kVar];
    for (int iter = 0;; ++iter) {
        if (threadIdx.x == 0) s_item[iter & 1] = (int)atomicAdd(&wl->next[kVar], 1u);
        __syncthreads();                 // the item is published; all waves are done with the last one
        const int it = __builtin_amdgcn_readfirstlane(s_item[iter & 1]);
        if (it >= end) return;           // uniform across the workgroup
        const SrvItem w = wl->items[it];
        const SrvJobDev& J = jobs[w.slot];
        DemDev d = J.a.dem;
        if (kLds) {
            const bool reload = (s_dem != J.a.dem.X);
            if (reload) {
                d = dem_to_lds(J.a.dem, lds);
                __syncthreads();         // staged, and every wave has read s_dem
                if (threadIdx.x == 0) s_dem = J.a.dem.X;
            } else {
                d = dem_lds_view(J.a.dem, lds);
            }
        }
        if ((int)threadIdx.x < w.nvox) {
            const int64_t v = w.vox0 + threadIdx.x;
            if (kMixed) {
                unsigned exact = 0;
                bp_mixed_thread<kLds, kLds, kLds>(d, J.a, J.m, w.g, v, w.p0, w.p1, exact);
                if (exact && J.m.exact_count) atomicAdd(J.m.exact_count, (unsigned long long)exact);
            } else {
                bp_fp64_thread<kLds, kLds>(d, J.a, w.g, v, w.p0, w.p1);
            }
        }
    }
}

const void* serve_kernel(int variant)
{
    switch (variant) {
    case 0: return (const void*)k_serve<false, false>;
    case 1: return (const void*)k_serve<false, true>;
    case 2: return (const void*)k_serve<true, false>;
    default: return (const void*)k_serve<true, true>;
    }
}

hipError_t launch_serve(int variant, int blocks, const SrvJobDev* jobs, SrvList* wl, hipStream_t s)
{
    const size_t lds = (variant & 1) ? (size_t)kSrvLdsDem * sizeof(double) : 0;
    switch (variant) {
    case 0: hipLaunchKernelGGL((k_serve<false, false>), dim3(blocks), dim3(256), 0, s, jobs, wl); break;
    case 1: hipLaunchKernelGGL((k_serve<false, true>), dim3(blocks), dim3(256), lds, s, jobs, wl); break;
    case 2: hipLaunchKernelGGL((k_serve<true, false>), dim3(blocks), dim3(256), 0, s, jobs, wl); break;
    default: hipLaunchKernelGGL((k_serve<true, true>), dim3(blocks), dim3(256), lds, s, jobs, wl); break;
    }
    return hipGetLastError();
}

}  // namespace bsar

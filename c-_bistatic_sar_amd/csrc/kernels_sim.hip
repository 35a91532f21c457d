// kernels_sim.hip -- the forward model of processRawData
// (raw_data_processing.cpp:292-452) on gfx950: per (scatterer, pulse) the
// refraction points, slant ranges, delay bin and phase of the response
// (:322-361), then the responses placed per pulse in scatterer order.
// Built with -ffp-contract=off: the ranges, bin and phase follow the
// reference's operation order; only cos/sin come from the device libm.
#include <math.h>

#include "bsar_device.h"
#include "bsar_kernels.h"
#include "bsar_sim.h"

namespace bsar {

__device__ __forceinline__ double dist3(double ax, double ay, double az, double bx, double by, double bz)
{
    const double a = ax - bx, b = ay - by, c = az - bz;
    return __builtin_sqrt(a * a + b * b + c * c);
}

// Thread per (scatterer s, pulse i), pulses fastest.  bins[s][i] = the row of
// the response (-1 when outside [0, S)), vals[s][i] its complex value.
template <int kMode>
__global__ __launch_bounds__(256, dem_waves(kMode, 3)) void k_sim_terms(DemDev dem, Traj tr,
                                                                        const double* __restrict__ scat,
                                                                        const double2* __restrict__ rcs, int ns,
                                                                        SimArgs sa, int* __restrict__ bins,
                                                                        double2* __restrict__ vals)
{
    extern __shared__ double lds[];
    const DemDev d = dem_stage<kMode>(dem, lds);
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= (int64_t)ns * tr.n) return;
    const int s = (int)(k / tr.n);
    const int i = (int)(k - (int64_t)s * tr.n);
    const double Px = scat[3 * s], Py = scat[3 * s + 1], Pz = scat[3 * s + 2];
    const double Tx = tr.tx[i], Ty = tr.ty[i], Tz = tr.tz[i];
    const double Rx = tr.rx[i], Ry = tr.ry[i], Rz = tr.rz[i];
    double ax, ay, az, bx, by, bz;
    fermat_solve<false, kMode == DEM_LDS>(d, Tx, Ty, Tz, Px, Py, Pz, sa.n1, sa.n2, ax, ay, az);      // rdp:166-172
    fermat_solve<false, kMode == DEM_LDS>(d, Px, Py, Pz, Rx, Ry, Rz, sa.n2, sa.n1, bx, by, bz);      // rdp:175-181
    const double r1Tx = dist3(Tx, Ty, Tz, ax, ay, az);                      // rdp:197-222
    const double r2Tx = dist3(Px, Py, Pz, ax, ay, az);
    const double r1Rx = dist3(Rx, Ry, Rz, bx, by, bz);
    const double r2Rx = dist3(Px, Py, Pz, bx, by, bz);
    const double totalRange = sa.n1 * r1Tx + sa.n2 * r2Tx + sa.n1 * r1Rx + sa.n2 * r2Rx;   // rdp:322-361
    const double t = totalRange / sa.c;
    const int rngBin = 1 + (int)round(t * sa.fs);
    const double phi = -(2.0 * M_PI / sa.lamb) * totalRange;
    double sn, cs;
    sincos(phi, &sn, &cs);
    const double2 a = rcs[s];
    bins[k] = (rngBin >= 0 && rngBin < sa.S) ? rngBin : -1;
    vals[k] = make_double2(a.x * cs - a.y * sn, a.x * sn + a.y * cs);
}

// Thread per pulse i: aux[i][bin] += value for every scatterer in order.
__global__ __launch_bounds__(256) void k_sim_place(const int* __restrict__ bins, const double2* __restrict__ vals,
                                                   int ns, int n, int S, double2* __restrict__ aux)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2* col = aux + (int64_t)i * S;
    for (int s = 0; s < ns; ++s) {
        const int64_t k = (int64_t)s * n + i;
        const int b = bins[k];
        if (b < 0) continue;
        const double2 v = vals[k];
        double2 x = col[b];
        x.x += v.x;
        x.y += v.y;
        col[b] = x;
    }
}

hipError_t launch_sim_terms(const DemDev& dem, const Traj& tr, const double* scat, const double2* rcs, int ns,
                            const SimArgs& sa, int* bins, double2* vals, hipStream_t s)
{
    const int64_t m = (int64_t)ns * tr.n;
    if (m <= 0) return hipSuccess;
    const DemPlan p = plan_dem(dem);
    BSAR_LAUNCH_DEM(k_sim_terms, 3, p, m, s, dem, tr, scat, rcs, ns, sa, bins, vals);
    return hipGetLastError();
}

hipError_t launch_sim_place(const int* bins, const double2* vals, int ns, int n, int S, double2* aux, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sim_place, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bins, vals, ns, n, S, aux);
    return hipGetLastError();
}

}  // namespace bsar

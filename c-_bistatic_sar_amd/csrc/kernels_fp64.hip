// kernels_fp64.hip -- FP64 kernels of the imaging hot path (gfx950).
//
//   k_dem_interp      interpolateDEMElevation, batched          (raw_data_processing.cpp:615-680)
//   k_refraction      calculateRefractionPoint, batched legs    (raw_data_processing.cpp:455-572)
//   k_refraction_wave the same search as a wavefront reduction over all 709 candidates
//   k_slant_range     refraction points of calculaSlantRangeFermat (raw_data_processing.cpp:153-193)
//   k_spline          splineInterpolation x factor per pulse    (interpolation.cpp:4-51)
//   k_bp_fp64         fused plain-FP64 back-projection          (proc_test_script.cpp:484-668)
//   k_group_sum       ordered sum of pulse-group partial images
//
// Built with -ffp-contract=off: the plain path evaluates the reference's
// operation order exactly, so refraction points, distances, delays and the
// interpolation weights are bit-identical to the reference's FP64 CPU build;
// only sin/cos (device libm vs glibc) may differ in the last bit.
#include "bsar_bp.h"
#include "bsar_device.h"
#include "bsar_kernels.h"

#include <math.h>

#include <algorithm>

namespace bsar {

__global__ __launch_bounds__(256) void k_dem_interp(DemDev dem, const double* __restrict__ xy, int64_t m,
                                                    double* __restrict__ z)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= m) return;
    z[k] = dem_interp(dem, xy[2 * k], xy[2 * k + 1]);
}

template <int kMode>
__global__ __launch_bounds__(dem_block(kMode, 3), dem_waves(kMode, 3)) void k_refraction(
    DemDev dem, const double* __restrict__ legs, int64_t m, double n1, double n2, double* __restrict__ out)
{
    extern __shared__ double lds[];
    const DemDev d = dem_stage<kMode>(dem, lds);
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double* s = legs + 6 * k;
    double qx, qy, qz;
    fermat_solve<false, kMode == DEM_LDS>(d, s[0], s[1], s[2], s[3], s[4], s[5], n1, n2, qx, qy, qz);
    out[3 * k] = qx;
    out[3 * k + 1] = qy;
    out[3 * k + 2] = qz;
}

// calculateRefractionPoint (raw_data_processing.cpp:455-572) as a wavefront-
// parallel reduction over the candidates: one wave per leg, its 64 lanes
// evaluate a level's G*G candidates exactly (lane l takes k = l, l + 64, ...
// in increasing order, so its own first minimum is the sequential scan's),
// then a butterfly argmin over the wave keeps the least value and, among
// equal values, the lowest index k = i*G + j -- the candidate the reference's
// strict '<' scan settles on.  The level's winner replaces the best point only
// when it is strictly below the minimum of the previous levels (:536-541).
// Every one of the 709 candidates is evaluated (no screening), in the same
// expression as the screened search's exact evaluations.
__device__ __forceinline__ void fermat_wave(const DemDev& d, double x1, double y1, double z1, double x2, double y2,
                                            double z2, double n1, double n2, double& qx, double& qy, double& qz)
{
    double bestX = (x1 + x2) / 2.0;                                  // :478-479
    double bestY = (y1 + y2) / 2.0;
    const double ex = x2 - x1, ey = y2 - y1;
    double R = std_max(__builtin_sqrt(ex * ex + ey * ey), 500.0);    // :483-484
    double minOpt = 1e9;                                             // :497
    const Leg L = make_leg(d, x1, y1, z1, x2, y2, z2, n1, n2);
    const bool empty = (d.flags & DEM_EMPTY) != 0;
    const bool xsorted = (d.flags & DEM_X_SORTED) != 0;
    const bool ysorted = (d.flags & DEM_Y_SORTED) != 0;
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll 1
    for (int level = 0; level < 5; ++level) {
        const int G = level == 0 ? 15 : 11;                          // :508
        const double minX = bestX - R, maxX = bestX + R;             // :509-512
        const double minY = bestY - R, maxY = bestY + R;
        const double spanX = maxX - minX, spanY = maxY - minY;
        double v = __builtin_inf();
        int kb = 0x7fffffff;
        for (int k = lane; k < G * G; k += 64) {
            const int i = k / G, j = k - (k / G) * G;
            const double px = minX + i * spanX / (G - 1);            // :525-526
            const double py = minY + j * spanY / (G - 1);
            int i0 = 0, di = 0;
            double wx = 0.0;
            if (!empty) axis_lookup(d.X, d.nx, d.x0, d.ihx, xsorted, px, i0, di, wx);
            const double dx1 = px - L.x1, dx2 = L.x2 - px;
            const double o = exact_opt(d, L, empty, ysorted, i0, di, wx, dx1 * dx1, dx2 * dx2, py);   // :529-534
            if (o < v) { v = o; kb = k; }                            // NaN never wins
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off);
            const int ok = __shfl_xor(kb, off);
            if (ov < v || (ov == v && ok < kb)) { v = ov; kb = ok; }
        }
        if (v < minOpt) {                                            // :536-541
            minOpt = v;
            const int bi = kb / G, bj = kb - (kb / G) * G;
            bestX = minX + bi * spanX / (G - 1);
            bestY = minY + bj * spanY / (G - 1);
        }
        R /= (level == 0) ? 5.0 : 3.0;                               // :551
    }
    qx = bestX;                                                      // :555-557
    qy = bestY;
    qz = dem_interp(d, bestX, bestY);
}

// One wave per leg (legs [m][6] -> out [m][3]), fermat_wave.
template <int kMode>
__global__ __launch_bounds__(256) void k_refraction_wave(DemDev dem, const double* __restrict__ legs, int64_t m,
                                                         double n1, double n2, double* __restrict__ out)
{
    extern __shared__ double lds[];
    const DemDev d = dem_stage<kMode>(dem, lds);
    const int64_t k = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;   // wave-uniform
    if (k >= m) return;
    const double* s = legs + 6 * k;
    double qx, qy, qz;
    fermat_wave(d, s[0], s[1], s[2], s[3], s[4], s[5], n1, n2, qx, qy, qz);
    if ((threadIdx.x & 63) == 0) {
        out[3 * k] = qx;
        out[3 * k + 1] = qy;
        out[3 * k + 2] = qz;
    }
}

// Thread per (point v, pulse i): both legs of one calculaSlantRangeFermat
// iteration (raw_data_processing.cpp:166-181).  Points run fastest across
// threads, so a wavefront searches around one antenna position for
// neighbouring points (coherent DEM accesses and screening outcomes); the
// outputs are [V][N][3] as calculaSlantRangeFermat returns them per target.
template <int kMode>
__global__ __launch_bounds__(dem_block(kMode, 3), dem_waves(kMode, 3)) void k_slant_range(
    DemDev dem, Traj tr, const double* __restrict__ pts, int64_t V, double n1, double n2, double* __restrict__ ida,
    double* __restrict__ volta)
{
    extern __shared__ double lds[];
    const DemDev d = dem_stage<kMode>(dem, lds);
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= V * tr.n) return;
    const int i = (int)(k / V);
    const int64_t v = k - (int64_t)i * V;
    const int64_t o = 3 * (v * tr.n + i);
    const double Px = pts[3 * v], Py = pts[3 * v + 1], Pz = pts[3 * v + 2];
    double qx, qy, qz;
    fermat_solve<false, kMode == DEM_LDS>(d, tr.tx[i], tr.ty[i], tr.tz[i], Px, Py, Pz, n1, n2, qx, qy, qz);
    ida[o] = qx; ida[o + 1] = qy; ida[o + 2] = qz;
    fermat_solve<false, kMode == DEM_LDS>(d, Px, Py, Pz, tr.rx[i], tr.ry[i], tr.rz[i], n2, n1, qx, qy, qz);
    volta[o] = qx; volta[o + 1] = qy; volta[o + 2] = qz;
}

// splineInterpolation, interpolation.cpp:4-51: one thread per output
// sample i of a pulse (grid x), pulses along grid y (and a stride loop when
// there are more than gridDim.y of them).  in: [N][S] complex (complex128,
// or complex64 widened exactly to complex128 on load), out: [N][S*f]
// complex128; out32, when not null, receives the same samples rounded to
// complex64 in the same pass (the mixed form's copy).  The per-sample terms
// (position, index, Hermite weights) depend on i alone and are computed once
// per thread.
__device__ __forceinline__ double2 widen(double2 a) { return a; }
__device__ __forceinline__ double2 widen(float2 a) { return make_double2((double)a.x, (double)a.y); }

template <typename In>
__global__ __launch_bounds__(256) void k_spline(const In* __restrict__ in, int S, int N, int f,
                                                double2* __restrict__ out, float2* __restrict__ out32)
{
    const int S4 = S * f;   // < 2^30 (launch_spline)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S4) return;
    double pos = (double)i / f;                                       // :18
    int idx = (int)pos;                                               // :19
    double frac = pos - (double)idx;                                  // :20
    if (idx >= S - 1) { idx = S - 2; frac = 1.0; }                    // :22-25
    const double t = frac, t2 = t * t, t3 = t2 * t;
    const double h1 = 2 * t3 - 3 * t2 + 1;                            // :33-36
    const double h2 = -2 * t3 + 3 * t2;
    const double h3 = t3 - 2 * t2 + t;
    const double h4 = t3 - t2;
    const bool hasp = idx > 0, hasn = idx < S - 2;
    for (int p = blockIdx.y; p < N; p += gridDim.y) {
        const In* d = in + (int64_t)p * S;
        const double2 a = widen(d[idx]), b = widen(d[idx + 1]);
        double2 dp = make_double2(0.0, 0.0), dn = make_double2(0.0, 0.0);
        if (hasp) { const double2 c = widen(d[idx - 1]); dp = make_double2(a.x - c.x, a.y - c.y); }   // :39,41
        if (hasn) { const double2 c = widen(d[idx + 2]); dn = make_double2(c.x - b.x, c.y - b.y); }   // :40,42
        double2 r;
        r.x = h1 * a.x + h2 * b.x + h3 * dp.x + h4 * dn.x;            // :45
        r.y = h1 * a.y + h2 * b.y + h3 * dp.y + h4 * dn.y;            // :46
        // streaming (non-temporal) stores: the upsampled data is far larger
        // than L2 and is read back only by later kernels (3 % faster)
        typedef double ND __attribute__((ext_vector_type(2)));
        typedef float NF __attribute__((ext_vector_type(2)));
        const int64_t k = (int64_t)p * S4 + i;
        __builtin_nontemporal_store(ND{r.x, r.y}, reinterpret_cast<ND*>(out + k));
        if (out32) __builtin_nontemporal_store(NF{(float)r.x, (float)r.y}, reinterpret_cast<NF*>(out32 + k));
    }
}

// Plain FP64 back-projection.  Thread = (pulse group g, voxel v); it carries
// the running sum of its voxel across launches (acc), adding the pulses of
// this launch's window in increasing order -- with one group that is exactly
// the reference's sequential accumulation (proc_test_script.cpp:589-641).
template <int kMode>
__global__ __launch_bounds__(dem_block(kMode, 4), dem_waves(kMode, 4)) void k_bp_fp64(BpArgs a)
{
    extern __shared__ double lds[];
    const DemDev d = dem_stage<kMode>(a.dem, lds);
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    // nvox_pad is a multiple of the wavefront size: g is wave-uniform.
    const int g = __builtin_amdgcn_readfirstlane((int)(t / a.nvox_pad));
    const int64_t vl = t - (int64_t)g * a.nvox_pad;
    if (g >= a.groups || vl >= a.nvox) return;
    const int gbeg = g * a.glen;
    const int gend = min(gbeg + a.glen, a.traj.n);
    const int p0 = gbeg + a.woff;
    const int p1 = min(p0 + a.wlen, gend);
    if (p0 >= p1) return;
    bp_fp64_thread<kMode == DEM_LDS, kMode == DEM_LDS>(d, a, g, a.vox0 + vl, p0, p1);
}

// out[v] = sum over groups in order; out is [nvox] complex (x-fastest voxels).
// out may be acc itself (group 0): each thread reads its voxel's groups before writing.
__global__ __launch_bounds__(256) void k_group_sum(const double2* acc, int groups, int64_t stride, int64_t nvox,
                                                   double2* out)
{
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    double r = 0.0, m = 0.0;
    for (int g = 0; g < groups; ++g) {
        const double2 x = acc[(int64_t)g * stride + v];
        r += x.x;
        m += x.y;
    }
    out[v] = make_double2(r, m);
}

// ---- host launchers -------------------------------------------------------

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }


hipError_t launch_dem_interp(const DemDev& dem, const double* xy, int64_t m, double* z, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dem_interp, dim3(blocks_for(m, 256)), dim3(256), 0, s, dem, xy, m, z);
    return hipGetLastError();
}

hipError_t launch_refraction(const DemDev& dem, const double* legs, int64_t m, double n1, double n2,
                             double* out, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    const DemPlan p = plan_dem(dem);
    BSAR_LAUNCH_DEM(k_refraction, 3, p, m, s, dem, legs, m, n1, n2, out);
    return hipGetLastError();
}

hipError_t launch_refraction_wave(const DemDev& dem, const double* legs, int64_t m, double n1, double n2,
                                  double* out, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    const DemPlan p = plan_dem(dem);
    const int64_t threads = m * 64;
    // every mode stages at most what the screened kernels stage; blocks of 256
    // threads hold four legs
    if (p.mode == DEM_LDS)
        hipLaunchKernelGGL(k_refraction_wave<DEM_LDS>, dim3(blocks_for(threads, 256)), dim3(256), p.bytes, s, dem,
                           legs, m, n1, n2, out);
    else if (p.mode == DEM_SPLIT) {
        (void)hipFuncSetAttribute((const void*)k_refraction_wave<DEM_SPLIT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.bytes);
        hipLaunchKernelGGL(k_refraction_wave<DEM_SPLIT>, dim3(blocks_for(threads, 256)), dim3(256), p.bytes, s, dem,
                           legs, m, n1, n2, out);
    } else
        hipLaunchKernelGGL(k_refraction_wave<DEM_GLOBAL>, dim3(blocks_for(threads, 256)), dim3(256), 0, s, dem, legs,
                           m, n1, n2, out);
    return hipGetLastError();
}

hipError_t launch_slant_range(const DemDev& dem, const Traj& tr, const double* pts, int64_t V, double n1,
                              double n2, double* ida, double* volta, hipStream_t s)
{
    const int64_t m = V * tr.n;
    if (m <= 0) return hipSuccess;
    const DemPlan p = plan_dem(dem);
    BSAR_LAUNCH_DEM(k_slant_range, 3, p, m, s, dem, tr, pts, V, n1, n2, ida, volta);
    return hipGetLastError();
}

template <typename In>
static hipError_t launch_spline_t(const In* in, int S, int N, int f, double2* out, float2* out32, hipStream_t s)
{
    const int64_t S4 = (int64_t)S * f;
    if (S < 2 || f < 1 || S4 >= (int64_t)1 << 30) return hipErrorInvalidValue;
    if (N <= 0) return hipSuccess;
    const unsigned gy = (unsigned)std::min(N, 65535);
    hipLaunchKernelGGL(k_spline<In>, dim3(blocks_for(S4, 256), gy), dim3(256), 0, s, in, S, N, f, out, out32);
    return hipGetLastError();
}

hipError_t launch_spline(const double2* in, int S, int N, int f, double2* out, float2* out32, hipStream_t s)
{
    return launch_spline_t(in, S, N, f, out, out32, s);
}

hipError_t launch_spline(const float2* in, int S, int N, int f, double2* out, float2* out32, hipStream_t s)
{
    return launch_spline_t(in, S, N, f, out, out32, s);
}

hipError_t launch_bp_fp64(const BpArgs& a, hipStream_t s)
{
    const int64_t threads = (int64_t)a.groups * a.nvox_pad;
    if (threads <= 0) return hipSuccess;
    const DemPlan p = plan_dem(a.dem);
    BSAR_LAUNCH_DEM(k_bp_fp64, 4, p, threads, s, a);
    return hipGetLastError();
}

hipError_t launch_group_sum(const double2* acc, int groups, int64_t stride, int64_t nvox, double2* out,
                            hipStream_t s)
{
    if (nvox <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_group_sum, dim3(blocks_for(nvox, 256)), dim3(256), 0, s, acc, groups, stride, nvox, out);
    return hipGetLastError();
}

}  // namespace bsar

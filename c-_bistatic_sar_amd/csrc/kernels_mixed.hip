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

#include "bsar_device.h"
#include "bsar_kernels.h"
#include "bsar_mixed.h"

namespace bsar {

// Mean distance table: value at lattice node (i, j, k) = (x0 + i h, y0 + j h,
// z0 + k h) is (1/N) sum_j |A_j - node| over the antenna set A.
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
    const double x = tab.x0 + i * tab.h, y = tab.y0 + j * tab.h, z = tab.z0 + k * tab.h;
    double s = 0.0;
    for (int m = 0; m < n; ++m) {
        const double a = ax[m] - x, b = ay[m] - y, c = az[m] - z;
        s += __builtin_sqrt(a * a + b * b + c * c);
    }
    out[t] = s / (double)n;
}

__device__ __forceinline__ void lagrange4(double t, double w[4])
{
    const double tm1 = t - 1.0, tm2 = t - 2.0, tp1 = t + 1.0;
    w[0] = -t * tm1 * tm2 * (1.0 / 6.0);
    w[1] = tp1 * tm1 * tm2 * 0.5;
    w[2] = -tp1 * t * tm2 * 0.5;
    w[3] = tp1 * t * tm1 * (1.0 / 6.0);
}

// Tricubic Lagrange lookup; false when the 4x4x4 stencil leaves the table.
__device__ __forceinline__ bool r1_lookup(const R1Table& tb, const double* __restrict__ F, double qx, double qy,
                                          double qz, double& val)
{
    const double ux = (qx - tb.x0) * tb.ih, uy = (qy - tb.y0) * tb.ih, uz = (qz - tb.z0) * tb.ih;
    if (!(ux >= 1.0 && ux < tb.nx - 2.0 && uy >= 1.0 && uy < tb.ny - 2.0 && uz >= 1.0 && uz < tb.nz - 2.0))
        return false;
    const double fx = floor(ux), fy = floor(uy), fz = floor(uz);
    const int ix = (int)fx - 1, iy = (int)fy - 1, iz = (int)fz - 1;
    double wx[4], wy[4], wz[4];
    lagrange4(ux - fx, wx);
    lagrange4(uy - fy, wy);
    lagrange4(uz - fz, wz);
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double accy = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double* row = F + ((int64_t)(iz + c) * tb.ny + (iy + b)) * tb.nx + ix;
            const double lx = wx[0] * row[0] + wx[1] * row[1] + wx[2] * row[2] + wx[3] * row[3];
            accy += wy[b] * lx;
        }
        acc += wz[c] * accy;
    }
    val = acc;
    return true;
}

__device__ __forceinline__ double r1_exact(const double* __restrict__ ax, const double* __restrict__ ay,
                                           const double* __restrict__ az, int n, double qx, double qy, double qz)
{
    double s = 0.0;
    for (int j = 0; j < n; ++j) {
        const double f0 = ax[j] - qx, f1 = ay[j] - qy, f2 = az[j] - qz;
        s += __builtin_sqrt(f0 * f0 + f1 * f1 + f2 * f2);
    }
    return s / (double)n;
}

// Kahan step: (s, c) += x
__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <bool kLds>
__global__ __launch_bounds__(256, 2) void k_bp_mixed(BpArgs a, MixedArgs m)
{
    extern __shared__ double lds[];
    DemDev d = a.dem;
    if (kLds) {
        d = dem_to_lds(a.dem, lds);
        __syncthreads();
    }
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int g = __builtin_amdgcn_readfirstlane((int)(t / a.nvox_pad));
    const int64_t vl = t - (int64_t)g * a.nvox_pad;
    if (g >= a.groups || vl >= a.nvox) return;
    const int gbeg = g * a.glen;
    const int gend = min(gbeg + a.glen, a.traj.n);
    const int p0 = gbeg + a.woff;
    const int p1 = min(p0 + a.wlen, gend);
    if (p0 >= p1) return;
    const int64_t v = a.vox0 + vl;
    const int xi = (int)(v % a.nx);
    const int yi = (int)((v / a.nx) % a.ny);
    const int zi = (int)(v / ((int64_t)a.nx * a.ny));
    const double Px = a.gx[xi], Py = a.gy[yi], Pz = a.gz[zi];
    const double n1 = a.n1, n2 = a.n2;
    const int n = a.traj.n;
    float4* st = m.acc + (int64_t)g * a.acc_stride + v;
    float4 s = *st;   // (sum.re, sum.im, comp.re, comp.im)
    unsigned exact = 0;

    for (int i = p0; i < p1; ++i) {
        double qax, qay, qaz, qbx, qby, qbz;
        fermat_solve(d, a.traj.tx[i], a.traj.ty[i], a.traj.tz[i], Px, Py, Pz, n1, n2, qax, qay, qaz);
        fermat_solve(d, Px, Py, Pz, a.traj.rx[i], a.traj.ry[i], a.traj.rz[i], n2, n1, qbx, qby, qbz);
        {
#pragma clang fp contract(fast)
            double e0 = Px - qax, e1 = Py - qay, e2 = Pz - qaz;
            const double R2t = __builtin_sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            e0 = Px - qbx; e1 = Py - qby; e2 = Pz - qbz;
            const double R2r = __builtin_sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            double R1t, R1r;
            if (!r1_lookup(m.tab, m.ftx, qax, qay, qaz, R1t)) {
                R1t = r1_exact(a.traj.tx, a.traj.ty, a.traj.tz, n, qax, qay, qaz);
                ++exact;
            }
            if (!r1_lookup(m.tab, m.frx, qbx, qby, qbz, R1r)) {
                R1r = r1_exact(a.traj.rx, a.traj.ry, a.traj.rz, n, qbx, qby, qbz);
                ++exact;
            }
            const double path = n1 * R1t + n2 * R2t + n2 * R2r + n1 * R1r;
            const double r = (1.0 + (a.inv_c * path) * a.fs4) - 1.0;
            if (!(r >= 0.0 && r <= a.s4m1)) continue;
            const int iL = (int)floor(r);
            const int iN = (int)ceil(r);
            const float q = (float)(r - (double)iL);
            // phase: FP64 argument reduced to [-pi, pi] in FP64, then FP32
            const double phi = a.kphi * path;
            const double red = phi - m.two_pi * rint(phi * m.inv_two_pi);
            float sn, cs;
            sincosf((float)red, &sn, &cs);
            const int az = min(i, a.ndata - 1);
            const float2* row = a.rc32 + (int64_t)az * a.S4;
            const float2 L = row[iL];
            float vr, vi;
            if (iL == iN) {
                vr = L.x; vi = L.y;
            } else {
                const float2 N = row[iN];
                vr = (1.0f - q) * L.x + q * N.x;
                vi = (1.0f - q) * L.y + q * N.y;
            }
            kahan_add(s.x, s.z, vr * cs - vi * sn);
            kahan_add(s.y, s.w, vr * sn + vi * cs);
        }
    }
    *st = s;
    if (exact && m.exact_count) atomicAdd(m.exact_count, (unsigned long long)exact);
}

// Ordered FP64 sum of the per-group compensated FP32 sums.
__global__ __launch_bounds__(256) void k_group_sum_mixed(const float4* __restrict__ acc, int groups, int64_t stride,
                                                         int64_t nvox, double2* __restrict__ out)
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
    const int cnt = a.dem.nx + a.dem.ny + a.dem.ny * a.dem.zs;
    if (cnt <= kLdsDemMax)
        hipLaunchKernelGGL(k_bp_mixed<true>, dim3(nb(threads, 256)), dim3(256), (size_t)cnt * sizeof(double), s, a, m);
    else
        hipLaunchKernelGGL(k_bp_mixed<false>, dim3(nb(threads, 256)), dim3(256), 0, s, a, m);
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

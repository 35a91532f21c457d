// bsar_bp.h -- per-voxel back-projection bodies shared by the one-image
// kernels (kernels_fp64.hip, kernels_mixed.hip) and the persistent serving
// kernel (kernels_serve.hip).  One call adds pulses [p0, p1) of pulse group g
// to the running sum of the voxel at thread position v (band_voxel), in
// increasing pulse order.
#pragma once
#include <math.h>

#include "bsar_device.h"
#include "bsar_kernels.h"

namespace bsar {

// Square root within 0.51 ulp of the exact root for a in [2^-1000, DBL_MAX]:
// v_rsq_f64 (relative error e below 2^-22) refined by one Goldschmidt step
// (error 1.5 e^2) and one Newton step on the square (error ~ e * 1.5 e^2,
// plus the final rounding).  A zero or subnormal a gives NaN, +inf gives NaN,
// NaN stays NaN.
__device__ __forceinline__ double sqrt_faithful(double a)
{
    const double y = __builtin_amdgcn_rsq(a);
    const double s0 = a * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, s0, 0.5);
    const double s1 = __builtin_fma(s0, r, s0);
    const double d = __builtin_fma(-s1, s1, a);
    return __builtin_fma(d, h, s1);
}

// Conditioning limit of the expanded squared distances in r1_mean_fp64.
constexpr double kR1Cond = 32.0;

// The antenna arrays are read-only for a whole launch: reading them through
// the constant address space lets the wave-uniform loads of the R1 sums be
// scalar loads also where the arrays' addresses come from memory (the
// serving kernel's job table) rather than from kernel arguments.
template <class T>
using const_as_ptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ const_as_ptr<T> as_const(const T* p)
{
    return (const_as_ptr<T>)p;
}

// Mean distance from the n antennas (ax, ay, az) to q (proc:518-548),
// summed in antenna order with faithful square roots.
//
// Squared distances.  When q is well separated from the antennas relative
// to their spread -- bsq + |q - o|^2 <= kR1Cond gap^2, gap = q's vertical
// distance from the antennas' height range, so every squared distance a is
// at least gap^2 -- each one is expanded about the origin o of g (R1Geo):
// a = c - 2 b.(q - o) + |q - o|^2 as one add and three fmas (12 instructions
// per antenna with the root and the sum).  Every intermediate is at most
// 2 (bsq + |q - o|^2), so the roundings (these four, |q - o|^2, c and the
// rounded b and q - o) stay below 11 eps (bsq + |q - o|^2) <= 352 eps a:
// 4e-14 relative on a, 2e-14 on the distance.  Otherwise (antennas spread
// wide against their height above q, q among them, non-finite values) the
// direct form: f0^2 + f1^2 + f2^2 with the last two terms fused (within 1 ulp
// of the reference's value; 14 instructions per antenna).
// When a squared distance is zero or subnormal, sqrt_faithful's NaN makes
// the sum NaN and it is redone with correctly rounded roots, which also
// gives the reference's NaN / infinite mean for NaN / infinite distances
// (the pulse then fails the range-window test).  Squared distances in
// [2^-1022, 2^-1000) -- distances below 1e-150 m -- are outside
// sqrt_faithful's range and may be off by a few ulp.
__device__ __forceinline__ double r1_mean_fp64(const double* __restrict__ ax, const double* __restrict__ ay,
                                               const double* __restrict__ az, const R1Geo& g,
                                               const double* __restrict__ gb, int n, double qx, double qy, double qz)
{
    double s = 0.0;
    const double ux = qx - g.ox, uy = qy - g.oy, uz = qz - g.oz;
    const double Q = __builtin_fma(uz, uz, __builtin_fma(uy, uy, ux * ux));
    const double gap = fmax(fmax(g.zlo - qz, qz - g.zhi), 0.0);
    if (g.bsq + Q <= kR1Cond * (gap * gap)) {
        const const_as_ptr<double> bx = as_const(gb);
        const const_as_ptr<double> by = as_const(gb + g.stride);
        const const_as_ptr<double> bz = as_const(gb + 2 * (int64_t)g.stride);
        const const_as_ptr<double> bc = as_const(gb + 3 * (int64_t)g.stride);
        const double mx = -2.0 * ux, my = -2.0 * uy, mz = -2.0 * uz;
#pragma unroll 4
        for (int j = 0; j < n; ++j) {
            const double a = __builtin_fma(bx[j], mx, __builtin_fma(by[j], my, __builtin_fma(bz[j], mz, bc[j] + Q)));
            s += sqrt_faithful(a);
        }
    } else {
        const const_as_ptr<double> cx = as_const(ax), cy = as_const(ay), cz = as_const(az);
#pragma unroll 4
        for (int j = 0; j < n; ++j) {
            const double f0 = cx[j] - qx, f1 = cy[j] - qy, f2 = cz[j] - qz;
            const double a = __builtin_fma(f2, f2, __builtin_fma(f1, f1, f0 * f0));
            s += sqrt_faithful(a);
        }
    }
    if (!(s == s)) {
        s = 0.0;
        for (int j = 0; j < n; ++j) {
            const double f0 = ax[j] - qx, f1 = ay[j] - qy, f2 = az[j] - qz;
            s += __builtin_sqrt(f0 * f0 + f1 * f1 + f2 * f2);
        }
    }
    return s / (double)n;
}

// Voxel of thread position p.  Positions run through bands of kBandY rows x
// kBandZ planes, x-major inside a band with the band's (y, z) cells
// innermost, so the 64 lanes of a wavefront hold one x column of an 8 x 8
// (y, z) block.  Neighbouring voxels take nearly the same path through the
// refraction search (the same kept rows, the same survivors), so the
// wavefront's union of its lanes' branches stays short: 6.5 % faster than
// 64 consecutive x voxels on the large image's 201 x 201 x 64 slab (on a
// 201 x 201 x 24 slab, 8 x 8 (x, y) tiles gained 4.4 % and 4 x 4 x 4 blocks
// 5.4 %).  Bijective over the grid (bands at
// the ends keep the rows and planes that remain); returns the voxel index
// (x fastest) and its coordinates.  Each voxel's arithmetic is unchanged, so
// the image does not depend on the mapping.
constexpr int kBandY = 8, kBandZ = 8;

__device__ __forceinline__ int64_t band_voxel(int64_t p, int nx, int ny, int nz, int& xi, int& yi, int& zi)
{
    const int64_t plane = (int64_t)nx * ny;
    const int c = (int)(p / (kBandZ * plane));
    const int bd = min(kBandZ, nz - kBandZ * c);               // planes in this band
    const int64_t r = p - (int64_t)c * kBandZ * plane;
    const int b = (int)(r / ((int64_t)kBandY * bd * nx));
    const int bh = min(kBandY, ny - kBandY * b);               // rows in this band
    const int w = (int)(r - (int64_t)b * kBandY * bd * nx);
    const int cell = bh * bd;
    xi = w / cell;
    const int s = w - xi * cell;
    const int zz = s / bh;
    yi = kBandY * b + (s - zz * bh);
    zi = kBandZ * c + zz;
    return (int64_t)zi * plane + (int64_t)yi * nx + xi;
}

// Plain FP64 body, proc_test_script.cpp:484-668 in the reference's operation
// order (built with -ffp-contract=off); the R1 means take faithful square
// roots (r1_mean_fp64).  Both bodies run the search with the side stop when
// the DEM is in LDS (kSplit); this one without the probe reuse (kReuse) and
// with the 15-wide first level in two column parts (the mixed body runs it in
// one: with its packed column terms that measured 1.5 % faster, while this
// body measured 3 % slower without the split).
template <bool kSplit = false, bool kPairs = false>
__device__ __forceinline__ void bp_fp64_thread(const DemDev& d, const BpArgs& a, int g, int64_t v, int p0, int p1)
{
    int xi, yi, zi;
    v = band_voxel(v, a.nx, a.ny, a.nz, xi, yi, zi);
    const double Px = a.gx[xi], Py = a.gy[yi], Pz = a.gz[zi];
    const double n1 = a.n1, n2 = a.n2;
    const int n = a.traj.n;
    double2* accp = a.acc + (int64_t)g * a.acc_stride + v;
    double accr = accp->x, acci = accp->y;

    for (int i = p0; i < p1; ++i) {
        double qax, qay, qaz, qbx, qby, qbz;
        fermat_solve<kSplit, kPairs, kSplit, false>(d, a.traj.tx[i], a.traj.ty[i], a.traj.tz[i], Px, Py, Pz, n1, n2, qax, qay, qaz);   // rdp:166-172
        fermat_solve<kSplit, kPairs, kSplit, false>(d, Px, Py, Pz, a.traj.rx[i], a.traj.ry[i], a.traj.rz[i], n2, n1, qbx, qby, qbz);   // rdp:175-181
        double e0 = Px - qax, e1 = Py - qay, e2 = Pz - qaz;
        const double R2t = __builtin_sqrt(e0 * e0 + e1 * e1 + e2 * e2);                                   // proc:496-505
        e0 = Px - qbx; e1 = Py - qby; e2 = Pz - qbz;
        const double R2r = __builtin_sqrt(e0 * e0 + e1 * e1 + e2 * e2);                                   // proc:507-516
        const double R1t = r1_mean_fp64(a.traj.tx, a.traj.ty, a.traj.tz, a.r1, a.r1.t, n, qax, qay, qaz); // proc:518-532
        const double R1r = r1_mean_fp64(a.traj.rx, a.traj.ry, a.traj.rz, a.r1, a.r1.r, n, qbx, qby, qbz); // proc:534-548
        const double path = n1 * R1t + n2 * R2t + n2 * R2r + n1 * R1r;
        const double tv = a.inv_c * path;                                                                // proc:567
        const double rngBin = 1.0 + (tv * a.fs4);                                                        // proc:571
        const double phi = a.kphi * path;                                                                // proc:575
        const double r = rngBin - 1.0;                                                                   // proc:593
        if (!(r >= 0.0 && r <= a.s4m1)) continue;                                                        // proc:595-599
        const int iL = (int)floor(r);
        const int iN = (int)ceil(r);
        const double q = r - (double)iL;                                                                 // proc:602
        const int az = min(i, a.ndata - 1);                                                              // proc:607
        const double2* row = a.rc + (int64_t)az * a.S4;
        const double2 L = row[iL];
        double vr, vi;
        if (iL == iN) {                                                                                  // proc:626-628
            vr = L.x; vi = L.y;
        } else {                                                                                         // proc:632-634
            const double2 N = row[iN];
            vr = (1.0 - q) * L.x + q * N.x;
            vi = (1.0 - q) * L.y + q * N.y;
        }
        double sn, cs;
        sincos(phi, &sn, &cs);                                                                           // proc:638
        accr += vr * cs - vi * sn;                                                                       // proc:641
        acci += vr * sn + vi * cs;
    }
    *accp = make_double2(accr, acci);
}

__device__ __forceinline__ void lagrange4(double t, double w[4])
{
    const double tm1 = t - 1.0, tm2 = t - 2.0, tp1 = t + 1.0;
    w[0] = -t * tm1 * tm2 * (1.0 / 6.0);
    w[1] = tp1 * tm1 * tm2 * 0.5;
    w[2] = -tp1 * t * tm2 * 0.5;
    w[3] = tp1 * t * tm1 * (1.0 / 6.0);
}

// Tricubic Lagrange lookup; false outside the table's lookup box.  The box
// keeps the 4x4x4 stencil inside the table ([lo, hi) within [a + 1, a + n - 2)).
__device__ __forceinline__ bool r1_lookup(const R1Table& tb, const double* __restrict__ F, double qx, double qy,
                                          double qz, double& val)
{
    const double ux = qx * tb.ih, uy = qy * tb.ih, uz = qz * tb.ih;   // exact: ih is a power of two
    if (!(ux >= tb.lo[0] && ux < tb.hi[0] && uy >= tb.lo[1] && uy < tb.hi[1] && uz >= tb.lo[2] && uz < tb.hi[2]))
        return false;
    const double fx = floor(ux), fy = floor(uy), fz = floor(uz);
    const int ix = (int)fx - 1 - tb.ax, iy = (int)fy - 1 - tb.ay, iz = (int)fz - 1 - tb.az;
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

// Mixed-precision body: exact FP64 refraction points, FP64 delay and phase
// argument, tabulated R1 means, FP32 interpolation and Kahan accumulation.
template <bool kSplit = false, bool kPairs = false, bool kSplit15 = false>
__device__ __forceinline__ void bp_mixed_thread(const DemDev& d, const BpArgs& a, const MixedArgs& m, int g,
                                                int64_t v, int p0, int p1, unsigned& exact)
{
    int xi, yi, zi;
    v = band_voxel(v, a.nx, a.ny, a.nz, xi, yi, zi);
    const double Px = a.gx[xi], Py = a.gy[yi], Pz = a.gz[zi];
    const double n1 = a.n1, n2 = a.n2;
    const int n = a.traj.n;
    float4* st = m.acc + (int64_t)g * a.acc_stride + v;
    float4 s = *st;   // (sum.re, sum.im, comp.re, comp.im)

    for (int i = p0; i < p1; ++i) {
        double qax, qay, qaz, qbx, qby, qbz;
        fermat_solve<kSplit15, kPairs, kSplit>(d, a.traj.tx[i], a.traj.ty[i], a.traj.tz[i], Px, Py, Pz, n1, n2, qax, qay, qaz);
        fermat_solve<kSplit15, kPairs, kSplit>(d, Px, Py, Pz, a.traj.rx[i], a.traj.ry[i], a.traj.rz[i], n2, n1, qbx, qby, qbz);
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
}

}  // namespace bsar

// bsar_device.h -- device-side building blocks shared by the HIP kernels:
// the DEM view, the first-match interval lookup of interpolateDEMElevation
// and the exact 5-level Fermat grid search of calculateRefractionPoint.
//
// Exactness contract: compiled with -ffp-contract=off, every expression here
// evaluates the same IEEE operations in the same order as the reference's
// x86-64 build (raw_data_processing.cpp:455-572, 615-680), so the selected
// refraction point is bit-identical.  Only + - * / sqrt enter the argmin, and
// all five are correctly rounded on gfx950 (hipcc's f64 div/sqrt expansions).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsar {

enum : int { DEM_EMPTY = 1, DEM_X_SORTED = 2, DEM_Y_SORTED = 4 };

// Device-resident DEM: X[nx], Y[ny], Z[ny][zs] (zs = nx + 1, the last
// column zero so pair loads at i0+1 stay in bounds).  Missing Z entries of a
// smaller or ragged source matrix are stored as 0.0.
struct DemDev {
    const double* X;
    const double* Y;
    const double* Z;
    int nx, ny, zs, flags;
    double x0, ihx, y0, ihy;   // first value and inverse mean spacing (index guess only)
};

// std::min / std::max as libstdc++ defines them (NaN-propagation included).
__device__ __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }

// First i in [0, n-2] with A[i] <= x && x <= A[i+1] (raw_data_processing.cpp:
// 625-640); i0 = 0, di = 0 when there is none.  w is the clamped weight of
// :667-671 (dx := 1 when the interval collapses, :664-665).
__device__ __forceinline__ void axis_lookup(const double* A, int n, double a0, double ih, bool sorted,
                                            double x, int& i0, int& di, double& w)
{
    int found = -1;
    if (sorted) {
        // A strictly increasing: with l = #{A[k] < x}, the only candidate is
        // max(l-1, 0); test it with the reference predicate itself.
        double gd = (x - a0) * ih + 1.0;
        int g;
        if (!(gd > 0.0)) g = 0;
        else if (gd >= (double)n) g = n;
        else g = (int)gd;
        while (g > 0 && A[g - 1] >= x) --g;
        while (g < n && A[g] < x) ++g;
        int c = g > 0 ? g - 1 : 0;
        if (c <= n - 2 && A[c] <= x && x <= A[c + 1]) found = c;
    } else {
        for (int i = 0; i + 1 < n; ++i) {
            if (x >= A[i] && x <= A[i + 1]) { found = i; break; }
        }
    }
    double ww;
    if (found >= 0) {
        i0 = found; di = 1;
        double d = A[found + 1] - A[found];
        if (d == 0) d = 1.0;
        ww = (x - A[found]) / d;
    } else {
        i0 = 0; di = 0;
        ww = x - A[0];           // (x - A[0]) / 1.0
    }
    w = std_max(0.0, std_min(1.0, ww));
}

// interpolateDEMElevation, raw_data_processing.cpp:615-680.
__device__ __forceinline__ double dem_interp(const DemDev& d, double x, double y)
{
    if (d.flags & DEM_EMPTY) return 0.0;
    int i0, di, j0, dj;
    double wx, wy;
    axis_lookup(d.X, d.nx, d.x0, d.ihx, (d.flags & DEM_X_SORTED) != 0, x, i0, di, wx);
    axis_lookup(d.Y, d.ny, d.y0, d.ihy, (d.flags & DEM_Y_SORTED) != 0, y, j0, dj, wy);
    const double* r0 = d.Z + (size_t)j0 * d.zs;
    const double* r1 = d.Z + (size_t)(j0 + dj) * d.zs;
    double z00 = r0[i0], z01 = r0[i0 + di], z10 = r1[i0], z11 = r1[i0 + di];
    if (!di && !dj) return z00;
    return z00 * (1 - wx) * (1 - wy) +
           z01 * wx * (1 - wy) +
           z10 * (1 - wx) * wy +
           z11 * wx * wy;
}

// One refinement level of calculateRefractionPoint (raw_data_processing.cpp:
// 508-551) on a G x G grid.  The scan keeps the reference's strict '<'
// against the minimum carried across levels; bestX/bestY are rebuilt from
// the winning (i, j) with the same expression that produced them.
template <int G>
__device__ __forceinline__ void fermat_level(const DemDev& d, double x1, double y1, double z1,
                                             double x2, double y2, double z2, double n1, double n2,
                                             double R, double& bestX, double& bestY, double& minOpt)
{
    const double minX = bestX - R, maxX = bestX + R;
    const double minY = bestY - R, maxY = bestY + R;
    const double spanX = maxX - minX, spanY = maxY - minY;
    const bool empty = (d.flags & DEM_EMPTY) != 0;
    const bool ysorted = (d.flags & DEM_Y_SORTED) != 0;
    const bool xsorted = (d.flags & DEM_X_SORTED) != 0;

    // Per-j terms (depend on py only): rows, weights, squared y offsets.
    int jr0[G], jr1[G];
    double wy[G], dya[G], dyb[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
        double py = minY + j * spanY / (G - 1);
        int j0 = 0, dj = 0;
        double w = 0.0;
        if (!empty) axis_lookup(d.Y, d.ny, d.y0, d.ihy, ysorted, py, j0, dj, w);
        jr0[j] = j0 * d.zs;
        jr1[j] = (j0 + dj) * d.zs;
        wy[j] = w;
        double a = py - y1, b = y2 - py;
        dya[j] = a * a;
        dyb[j] = b * b;
    }
    int bi = -1, bj = 0;
    for (int i = 0; i < G; ++i) {
        double px = minX + i * spanX / (G - 1);
        int i0 = 0, di = 0;
        double wx = 0.0;
        if (!empty) axis_lookup(d.X, d.nx, d.x0, d.ihx, xsorted, px, i0, di, wx);
        const double omwx = 1 - wx;
        const double a = px - x1, b = x2 - px;
        const double dxa = a * a, dxb = b * b;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            double pz;
            if (empty) {
                pz = 0.0;
            } else {
                const double* r0 = d.Z + jr0[j] + i0;
                const double* r1 = d.Z + jr1[j] + i0;
                double z00 = r0[0], z0b = r0[1], z10 = r1[0], z1b = r1[1];
                double z01 = di ? z0b : z00;
                double z11 = di ? z1b : z10;
                if (!di && jr0[j] == jr1[j] && jr0[j] == 0) {
                    pz = z00;
                } else {
                    const double omwy = 1 - wy[j];
                    pz = z00 * omwx * omwy + z01 * wx * omwy + z10 * omwx * wy[j] + z11 * wx * wy[j];
                }
            }
            const double c1 = pz - z1;
            const double d1 = __builtin_sqrt(dxa + dya[j] + c1 * c1);
            const double c2 = z2 - pz;
            const double d2 = __builtin_sqrt(dxb + dyb[j] + c2 * c2);
            const double opt = n1 * d1 + n2 * d2;
            if (opt < minOpt) { minOpt = opt; bi = i; bj = j; }
        }
    }
    if (bi >= 0) {
        bestX = minX + bi * spanX / (G - 1);
        bestY = minY + bj * spanY / (G - 1);
    }
}

// calculateRefractionPoint, raw_data_processing.cpp:455-572.
__device__ __forceinline__ void fermat_solve(const DemDev& d, double x1, double y1, double z1,
                                             double x2, double y2, double z2, double n1, double n2,
                                             double& qx, double& qy, double& qz)
{
    double bestX = (x1 + x2) / 2.0;                                  // :478-479
    double bestY = (y1 + y2) / 2.0;
    const double ex = x2 - x1, ey = y2 - y1;
    double R = std_max(__builtin_sqrt(ex * ex + ey * ey), 500.0);    // :483-484
    double minOpt = 1e9;                                             // :497
    fermat_level<15>(d, x1, y1, z1, x2, y2, z2, n1, n2, R, bestX, bestY, minOpt);
    R /= 5.0;                                                        // :551
#pragma unroll 1
    for (int level = 1; level < 5; ++level) {
        fermat_level<11>(d, x1, y1, z1, x2, y2, z2, n1, n2, R, bestX, bestY, minOpt);
        R /= 3.0;
    }
    qx = bestX;                                                      // :555-557
    qy = bestY;
    qz = dem_interp(d, bestX, bestY);
}

// Cooperative copy of the DEM into LDS; returns a view onto the copy.
__device__ __forceinline__ DemDev dem_to_lds(const DemDev& g, double* lds)
{
    DemDev d = g;
    const int nz = g.ny * g.zs;
    const int total = g.nx + g.ny + nz;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        double v;
        if (k < g.nx) v = g.X[k];
        else if (k < g.nx + g.ny) v = g.Y[k - g.nx];
        else v = g.Z[k - g.nx - g.ny];
        lds[k] = v;
    }
    d.X = lds;
    d.Y = lds + g.nx;
    d.Z = lds + g.nx + g.ny;
    return d;
}

}  // namespace bsar

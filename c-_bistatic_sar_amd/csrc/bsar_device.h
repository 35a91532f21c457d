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

#include <type_traits>

namespace bsar {

enum : int { DEM_EMPTY = 1, DEM_X_SORTED = 2, DEM_Y_SORTED = 4 };

// Device-resident DEM: X[nx], Y[ny], Z[ny][zs] (zs = nx + 1, the last
// column zero so pair loads at i0+1 stay in bounds).  Missing Z entries of a
// smaller or ragged source matrix are stored as 0.0.
struct DemDev {
    const double* X;
    const double* Y;
    const double* Z;
    const float* Zf;           // Z rounded to FP32, same layout (screening estimates only)
    const float2* Zp;          // (Z, Z[.][i+1] - Z[.][i]) in FP32 per node, same layout (screening, DEM in LDS)
    int nx, ny, zs, flags;
    double x0, ihx, y0, ihy;   // first value and inverse mean spacing (index guess only)
    double xn, yn;             // last values X[nx-1], Y[ny-1]
    double zlo, zhi;           // height range of the stored Z (0, 0 when empty)
    double zabs;               // max |Z|
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
    double lo = 0.0, hi = 0.0;   // A[found], A[found + 1]
    if (sorted) {
        // A strictly increasing: with l = #{A[k] < x}, the only candidate is
        // max(l-1, 0); test it with the reference predicate itself.
        double gd = (x - a0) * ih + 1.0;
        int g;
        if (!(gd > 0.0)) g = 0;
        else if (gd >= (double)n) g = n;
        else g = (int)gd;
        // Inside the axis the uniform-spacing guess is nearly always l: both
        // nodes around it load in one round trip, and A[g-1] < x <= A[g]
        // confirms l = g and the interval g - 1 at once.
        const int gl = g > 0 ? g - 1 : 0, gh = g < n ? g : n - 1;
        lo = A[gl];
        hi = A[gh];
        if (g > 0 && g < n && lo < x && x <= hi) {
            found = g - 1;
        } else {
            while (g > 0 && A[g - 1] >= x) --g;
            while (g < n && A[g] < x) ++g;
            int c = g > 0 ? g - 1 : 0;
            if (c <= n - 2 && A[c] <= x && x <= A[c + 1]) {
                found = c;
                lo = A[c];
                hi = A[c + 1];
            }
        }
    } else {
        for (int i = 0; i + 1 < n; ++i) {
            if (x >= A[i] && x <= A[i + 1]) {
                found = i;
                lo = A[i];
                hi = A[i + 1];
                break;
            }
        }
    }
    double ww;
    if (found >= 0) {
        i0 = found; di = 1;
        double d = hi - lo;
        if (d == 0) d = 1.0;
        ww = (x - lo) / d;
    } else {
        i0 = 0; di = 0;
        ww = x - A[0];           // (x - A[0]) / 1.0
    }
    w = std_max(0.0, std_min(1.0, ww));
}

// The interval of axis_lookup for an approximate position x on a strictly
// increasing axis, with an FP32 weight for the screening estimates only:
// (x - A[i0]) * rcp(A[i0+1] - A[i0]) from FP32 operands (four roundings and
// v_rcp_f32's 1 ulp), clamped to [0, 1].  Returns false -- the caller then
// looks up the exact position -- when x lies within eps of the interval's
// ends: the exact position, within eps of x, could then fall in the
// neighbouring interval or outside the axis (where the lookup's
// out-of-range rule jumps to row 0), and with non-finite heights the two
// intervals need not agree on whether the estimate is NaN.
__device__ __forceinline__ bool axis_cell_inside(const double* A, int n, double a0, double an, double ih, double x,
                                                 double eps, int& i0, float& wf)
{
    if (!(x - a0 > eps && an - x > eps)) return false;   // a0 = A[0], an = A[n-1]
    double gd = (x - a0) * ih + 1.0;
    int g = !(gd > 0.0) ? 0 : (gd >= (double)n ? n : (int)gd);
    const int gl = g > 0 ? g - 1 : 0, gh = g < n ? g : n - 1;
    double lo = A[gl], hi = A[gh];
    if (!(g > 0 && g < n && lo < x && x <= hi)) {
        while (g > 0 && A[g - 1] >= x) --g;
        while (g < n && A[g] < x) ++g;
        g = max(min(g, n - 1), 1);   // A[0] < x <= A[n-1]: g - 1 is the interval
        lo = A[g - 1];
        hi = A[g];
    }
    if (!(x - lo > eps && hi - x > eps)) return false;
    i0 = g - 1;
    const float w = (float)(x - lo) * __builtin_amdgcn_rcpf((float)(hi - lo));
    wf = __builtin_fminf(__builtin_fmaxf(w, 0.0f), 1.0f);
    return true;
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

// ---- screened exact grid search ----------------------------------------
//
// The reference scans every candidate of a level in (i, j) order with a
// strict '<' against the minimum carried across levels.  The same argmin is
// found here by evaluating exactly (reference operation order, correctly
// rounded sqrt) only the candidates that can still be the minimum:
//
//   * rows i / columns j whose Euclidean lower bound
//       n1*sqrt(dx1^2 + g1^2) + n2*sqrt(dx2^2 + g2^2)
//     (horizontal offset along one axis, vertical gap g to the DEM height
//     range) exceeds the best value known so far are skipped;
//   * every other candidate gets a cheap FP32 estimate (bilinear height
//     from the FP32 copy of Z, raw v_sqrt_f32) and is evaluated exactly only
//     when the estimate, widened by its error bound, could reach the best
//     value;
//   * candidates are visited centre-out (the previous level's best sits at
//     the centre), and ties between exact values resolve to the lowest
//     reference index i*G + j -- the first one the sequential scan meets.
//
// Error bounds.  Lower bounds (lower_bound32): FP32 with v_rsq_f32 (about 1
// ulp), at most 4.8e-7 relative in all; kDefl deflates them by 8e-6, more
// than an order of magnitude beyond that.  Estimates: with u = 2^-24, the FP32 inputs (Z, the
// bilinear weights, z1, z2, the squared horizontal offsets, n1, n2: one
// rounding each; with the (height, x difference) pairs of an LDS-staged DEM
// the difference is rounded once from FP64 instead of formed from two
// rounded heights), the bilinear blend (~30 u |Z|max absolute), the sums and
// v_sqrt_f32 (1 ulp) give
//   |est - exact| <= 13 u est + 3 (n1 + n2) 31 u (|Z|max + |z1| + |z2|),
// i.e. below kRho32 * est + D32 with kRho32 = 4e-6 (5x) and
// D32 = (n1 + n2) 1e-5 (|Z|max + |z1| + |z2| + 1) (1.6x).  The column terms
// come from a position within epsy (~1e-11 m) of the candidate's and an FP32
// column weight (axis_cell_inside: 5 u relative, i.e. at most 10 u |Z|max more
// on a height, (n1 + n2) 10 u |Z|max on the estimate), which the 1.6x leaves
// room for (75 u of the 1e-5).  The y offsets are FP32: dy1 = py - y1 (one
// rounding) and dy2 = (y2 - y1) - dy1 (three), so dy2 is off by at most
// 2u (|y2 - y1| + |dy2|): the relative part is within the 5x of kRho32, the
// absolute part adds (n1 + n2) 2^-22 |y2 - y1| to D32 (make_leg).  A candidate is
// dropped only when est > M (1 + 2 kRho32) + 2 D32, which puts its exact
// value above M; M itself only takes exact values or est upper bounds
// (est + D32)(1 + 3 kRho32).  Any candidate dropped by a bound or an estimate
// has an exact value strictly above an exact value already reached (or above
// the minimum of the previous levels), so it can neither be the minimum nor
// tie with it.
constexpr double kRho32 = 4e-6;
constexpr double kDefl = 1.0 - 8e-6;

__device__ __forceinline__ double sqrt_est(double a) { return a * __builtin_amdgcn_rsq(a); }
__device__ __forceinline__ float sqrt_f32(float a) { return __builtin_amdgcn_sqrtf(a); }

// Deflated lower bound n1 sqrt(a^2 + g1sq) + n2 sqrt(b^2 + g2sq) in FP32 for
// a, b >= 0: the roundings of a, b, g1sq, g2sq, n1, n2 to FP32, the fma and
// v_rsq_f32 (1 ulp) keep the relative error below 8 * 2^-24 = 4.8e-7, far
// inside the 8e-6 of kDefl, so the result stays below the exact bound.  A
// zero argument gives 0 * inf = NaN, which never prunes.
__device__ __forceinline__ double lower_bound32(double a, double b, float g1f, float g2f, float n1f, float n2f)
{
    const float af = (float)a, bf = (float)b;
    const float s1 = __builtin_fmaf(af, af, g1f), s2 = __builtin_fmaf(bf, bf, g2f);
    const float lb = __builtin_fmaf(n2f, s2 * __builtin_amdgcn_rsqf(s2), n1f * (s1 * __builtin_amdgcn_rsqf(s1)));
    return (double)(lb * (float)kDefl);
}

// Slope of the one-axis bound l(q) = n1 sqrt((q - c1)^2 + g1sq) + n2 sqrt((q - c2)^2 + g2sq)
// from the signed offsets u1 = q - c1, u2 = q - c2, in FP32.  Each term
// n u / sqrt(u^2 + g) lies in [-n, n]; the roundings of u, g, n, the fma,
// v_rsq_f32 and the products keep its error below 8 * 2^-24 * n, so the sum
// is within 5e-7 (n1 + n2) of the exact slope -- 20x inside kSlope.  u = g = 0
// gives 0 * inf = NaN, which certifies nothing.
constexpr float kSlope = 1e-5f;

__device__ __forceinline__ float slope32(double u1, double u2, float g1f, float g2f, float n1f, float n2f)
{
    const float a = (float)u1, b = (float)u2;
    const float s1 = __builtin_fmaf(a, a, g1f), s2 = __builtin_fmaf(b, b, g2f);
    return __builtin_fmaf(n2f * b, __builtin_amdgcn_rsqf(s2), (n1f * a) * __builtin_amdgcn_rsqf(s1));
}

struct Leg {
    double x1, y1, z1, x2, y2, z2, n1, n2;
    double g1sq, g2sq;   // squared vertical gaps from z1 / z2 to the DEM height range
    double D;            // absolute error bound of the FP32 estimate (D32)
    float z1f, z2f, n1f, n2f;
    bool screen;         // bounds valid (n1, n2 >= 0, finite, moderate); else every candidate is exact
};

__device__ __forceinline__ Leg make_leg(const DemDev& d, double x1, double y1, double z1, double x2, double y2,
                                        double z2, double n1, double n2)
{
    Leg L{};
    L.x1 = x1; L.y1 = y1; L.z1 = z1; L.x2 = x2; L.y2 = y2; L.z2 = z2; L.n1 = n1; L.n2 = n2;
    const double g1 = fmax(fmax(d.zlo - z1, z1 - d.zhi), 0.0);
    const double g2 = fmax(fmax(d.zlo - z2, z2 - d.zhi), 0.0);
    L.g1sq = g1 * g1;
    L.g2sq = g2 * g2;
    L.D = (fabs(n1) + fabs(n2)) * (1e-5 * (d.zabs + fabs(z1) + fabs(z2) + 1.0) + 0x1p-22 * fabs(y2 - y1));
    L.z1f = (float)z1; L.z2f = (float)z2; L.n1f = (float)n1; L.n2f = (float)n2;
    // FP32 estimates need every magnitude well inside FP32 range
    const double big = 1e15;
    L.screen = n1 >= 0.0 && n2 >= 0.0 && n1 + n2 < 1e6 && fabs(x1) < big && fabs(y1) < big && fabs(z1) < big &&
               fabs(x2) < big && fabs(y2) < big && fabs(z2) < big && d.zabs < big;
    return L;
}

__device__ __forceinline__ int center_out(int k, int c) { return (k & 1) ? c - ((k + 1) >> 1) : c + (k >> 1); }

// Largest i with A[i] <= x (0 when none) and smallest i with A[i] >= x
// (n - 1 when none) on a strictly increasing axis, walked from any start i.
__device__ __forceinline__ int floor_walk(const double* A, int n, double x, int i)
{
    while (i > 0 && A[i] > x) --i;
    while (i + 1 < n && A[i + 1] <= x) ++i;
    return i;
}
__device__ __forceinline__ int ceil_walk(const double* A, int n, double x, int i)
{
    while (i + 1 < n && A[i] < x) ++i;
    while (i > 0 && A[i - 1] >= x) --i;
    return i;
}

// floor_walk of xa, ya and ceil_walk of xb, yb from the uniform-spacing
// guesses (floor and ceiling of the fractional index).  The eight nodes that
// confirm the four guesses load together, so when every guess is right the
// window costs one LDS round trip; a wrong guess walks.
__device__ __forceinline__ void window_nodes(const DemDev& d, double xa, double xb, double ya, double yb,
                                             int& i0, int& i1, int& j0, int& j1)
{
    auto gfloor = [](int n, double a0, double ih, double x) {
        const double gd = (x - a0) * ih;
        return !(gd > 0.0) ? 0 : (gd >= (double)(n - 1) ? n - 1 : (int)gd);
    };
    auto gceil = [](int n, double a0, double ih, double x) {
        const double gd = (x - a0) * ih;
        return !(gd > 0.0) ? 0 : (gd >= (double)(n - 1) ? n - 1 : min((int)gd + 1, n - 1));
    };
    int a = gfloor(d.nx, d.x0, d.ihx, xa), b = gceil(d.nx, d.x0, d.ihx, xb);
    int c = gfloor(d.ny, d.y0, d.ihy, ya), e = gceil(d.ny, d.y0, d.ihy, yb);
    const double pa = d.X[a], qa = d.X[min(a + 1, d.nx - 1)];
    const double pb = d.X[b], qb = d.X[max(b - 1, 0)];
    const double pc = d.Y[c], qc = d.Y[min(c + 1, d.ny - 1)];
    const double pe = d.Y[e], qe = d.Y[max(e - 1, 0)];
    if (!(pa <= xa && (a + 1 >= d.nx || qa > xa))) a = floor_walk(d.X, d.nx, xa, a);
    if (!(pb >= xb && (b == 0 || qb < xb))) b = ceil_walk(d.X, d.nx, xb, b);
    if (!(pc <= ya && (c + 1 >= d.ny || qc > ya))) c = floor_walk(d.Y, d.ny, ya, c);
    if (!(pe >= yb && (e == 0 || qe < yb))) e = ceil_walk(d.Y, d.ny, yb, e);
    i0 = a; i1 = b; j0 = c; j1 = e;
}

// Squared vertical gaps from z1 / z2 to the heights the DEM lookup can return
// inside the window [xa, xb] x [ya, yb]: convex combinations of the nodes of
// the intervals the window meets, plus column / row 0 where it leaves the
// axis range (raw_data_processing.cpp:621-671).  Left unchanged (the global
// range) for unsorted axes or windows over more than 16 nodes.
__device__ __forceinline__ void local_gaps(const DemDev& d, double z1, double z2, double xa, double xb, double ya,
                                           double yb, double& g1sq, double& g2sq)
{
    int i0, i1, j0, j1;
    window_nodes(d, xa, xb, ya, yb, i0, i1, j0, j1);
    if (i1 < i0 || j1 < j0 || (i1 - i0 + 1) * (j1 - j0 + 1) > 16) return;
    const bool wx0 = xa < d.x0 || xb > d.xn;
    const bool wy0 = ya < d.y0 || yb > d.yn;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    if (i1 - i0 < 4 && j1 - j0 < 4) {
        // up to 4 x 4 nodes in one round trip: the clamped indices repeat a
        // node of the window, which leaves the minimum and maximum unchanged
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double* r = d.Z + min(j0 + b, j1) * d.zs;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double v = r[min(i0 + a, i1)];
                lo = fmin(lo, v);
                hi = fmax(hi, v);
            }
        }
    } else {
        for (int j = j0; j <= j1; ++j) {
            const double* r = d.Z + j * d.zs;
            for (int i = i0; i <= i1; ++i) { lo = fmin(lo, r[i]); hi = fmax(hi, r[i]); }
        }
    }
    if (wx0) {
        for (int j = j0; j <= j1; ++j) { lo = fmin(lo, d.Z[j * d.zs]); hi = fmax(hi, d.Z[j * d.zs]); }
    }
    if (wy0) {
        for (int i = i0; i <= i1; ++i) { lo = fmin(lo, d.Z[i]); hi = fmax(hi, d.Z[i]); }
        lo = fmin(lo, d.Z[0]);
        hi = fmax(hi, d.Z[0]);
    }
    if (!(hi >= lo)) return;   // every node NaN: keep the global range
    const double g1 = fmax(fmax(lo - z1, z1 - hi), 0.0);
    const double g2 = fmax(fmax(lo - z2, z2 - hi), 0.0);
    g1sq = g1 * g1;
    g2sq = g2 * g2;
}

// Relative margin of window_gaps' FP32 height range (|Z|max units): the
// heights are formed from FP32 node heights and differences (one rounding
// each), an FP32 weight (the rounded FP64 offset times v_rcp_f32 of the
// rounded interval width: about 3 ulp) and two fmas, at most ~8 ulp of
// |Z|max (8 * 2^-24 = 4.8e-7) from the exact surface; the margin is 8x that.
constexpr double kGapMargin = 4e-6;

// The squared vertical gaps of local_gaps from the heights at the window
// itself instead of whole DEM cells.  On a strictly increasing axis the
// lookup returns the bilinear surface of the cells the window meets (for
// positions strictly inside the axis range), which is linear along each
// axis inside a cell; over [xa, xb] x [ya, yb] its extremes therefore lie
// at the corners of the window's pieces in each cell: the x positions xa,
// the DEM columns strictly inside, xb, against the y positions ya, the rows
// strictly inside, yb.  Those (up to 4 x 4) heights are formed in FP32 from
// the (height, x difference) pairs (kPairs) or the FP32 heights, and the
// range is widened by kGapMargin |Z|max.  Windows that reach the axis ends
// or span more than three intervals per axis keep local_gaps' node range.
// NaN heights are skipped by fmin/fmax: a candidate whose lookup meets one
// gets a NaN path, which never wins.
template <bool kPairs>
__device__ __forceinline__ void window_gaps(const DemDev& d, double z1, double z2, double xa, double xb,
                                            double ya, double yb, double& g1sq, double& g2sq)
{
    int i0, i1, j0, j1;
    window_nodes(d, xa, xb, ya, yb, i0, i1, j0, j1);
    // window_nodes: X[i0] <= xa, xb <= X[i1], Y[j0] <= ya, yb <= Y[j1]; strictly
    // inside the axis range i0 < i1 and j0 < j1 whenever xa < xb and ya < yb
    const bool tight = xa > d.x0 && xb < d.xn && ya > d.y0 && yb < d.yn && i0 < i1 && j0 < j1 &&
                       i1 - i0 <= 3 && j1 - j0 <= 3;
    if (!tight) {
        local_gaps(d, z1, z2, xa, xb, ya, yb, g1sq, g2sq);
        return;
    }
    auto pair = [&](int j, int i) -> float2 {
        if (kPairs) return d.Zp[j * d.zs + i];
        const float a = d.Zf[j * d.zs + i];
        return make_float2(a, d.Zf[j * d.zs + i + 1] - a);
    };
    auto weight = [](double x, double lo, double hi) {
        const float w = (float)(x - lo) * __builtin_amdgcn_rcpf((float)(hi - lo));
        return __builtin_fminf(__builtin_fmaxf(w, 0.0f), 1.0f);
    };
    const float wa = weight(xa, d.X[i0], d.X[i0 + 1]);
    const float wb = weight(xb, d.X[i1 - 1], d.X[i1]);
    const float wya = weight(ya, d.Y[j0], d.Y[j0 + 1]);
    const float wyb = weight(yb, d.Y[j1 - 1], d.Y[j1]);
    const int lx = i1 - i0, ly = j1 - j0;   // 1..3 intervals per axis
    // P[b][m]: node row j0 + b (clamped to j1) at the x positions xa, the
    // inner columns i0 + 1 .. i1 - 1, xb (position m clamped to lx)
    float P[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int j = j0 + min(b, ly);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int mm = min(m, lx);
            if (mm == 0) {
                const float2 p = pair(j, i0);
                P[b][m] = __builtin_fmaf(p.y, wa, p.x);
            } else if (mm == lx) {
                const float2 p = pair(j, i1 - 1);
                P[b][m] = __builtin_fmaf(p.y, wb, p.x);
            } else {
                P[b][m] = pair(j, i0 + mm).x;
            }
        }
    }
    float lo = __builtin_inff(), hi = -__builtin_inff();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        // y positions: ya (rows j0, j0 + 1), the inner rows, yb (rows j1 - 1, j1)
        const float ha = __builtin_fmaf(P[1][m] - P[0][m], wya, P[0][m]);
        const float pb0 = ly == 1 ? P[0][m] : (ly == 2 ? P[1][m] : P[2][m]);
        const float pb1 = ly == 1 ? P[1][m] : (ly == 2 ? P[2][m] : P[3][m]);
        const float hb = __builtin_fmaf(pb1 - pb0, wyb, pb0);
        lo = __builtin_fminf(lo, __builtin_fminf(ha, hb));
        hi = __builtin_fmaxf(hi, __builtin_fmaxf(ha, hb));
        if (ly >= 2) {
            lo = __builtin_fminf(lo, P[1][m]);
            hi = __builtin_fmaxf(hi, P[1][m]);
        }
        if (ly == 3) {
            lo = __builtin_fminf(lo, P[2][m]);
            hi = __builtin_fmaxf(hi, P[2][m]);
        }
    }
    if (!(hi >= lo)) return;   // every height NaN: keep the global range
    const double mg = kGapMargin * d.zabs;
    const double dlo = (double)lo - mg, dhi = (double)hi + mg;
    const double g1 = fmax(fmax(dlo - z1, z1 - dhi), 0.0);
    const double g2 = fmax(fmax(dlo - z2, z2 - dhi), 0.0);
    g1sq = g1 * g1;
    g2sq = g2 * g2;
}

// Exact optical path of candidate (px, py) in the reference's operation order
// (raw_data_processing.cpp:525-534 with interpolateDEMElevation :615-680);
// i0, di, wx are px's x-interval terms, dxa = (px - x1)^2, dxb = (x2 - px)^2.
__device__ __forceinline__ double exact_opt(const DemDev& d, const Leg& L, bool empty, bool ysorted, int i0, int di,
                                            double wx, double dxa, double dxb, double py)
{
    double pz = 0.0;
    if (!empty) {
        int j0 = 0, dj = 0;
        double w = 0.0;
        axis_lookup(d.Y, d.ny, d.y0, d.ihy, ysorted, py, j0, dj, w);
        const double* r0 = d.Z + j0 * d.zs + i0;
        const double* r1 = d.Z + (j0 + dj) * d.zs + i0;
        const double z00 = r0[0], z10 = r1[0];
        const double z01 = di ? r0[1] : z00;
        const double z11 = di ? r1[1] : z10;
        if (!di && !dj) {
            pz = z00;
        } else {
            const double omwx = 1 - wx, omwy = 1 - w;
            pz = z00 * omwx * omwy + z01 * wx * omwy + z10 * omwx * w + z11 * wx * w;
        }
    }
    const double dy1 = py - L.y1, dy2 = L.y2 - py;
    const double c1 = pz - L.z1;
    const double d1 = __builtin_sqrt(dxa + dy1 * dy1 + c1 * c1);
    const double c2 = L.z2 - pz;
    const double d2 = __builtin_sqrt(dxb + dy2 * dy2 + c2 * c2);
    return L.n1 * d1 + L.n2 * d2;
}

// Lower bound of the path through any point of a row (fixed x) of a window:
//   h(y) = n1 sqrt(a1 + (y - c1)^2) + n2 sqrt(a2 + (y - c2)^2),  y in [0, w],
// with a1 = dx1^2 + g1^2, a2 = dx2^2 + g2^2 (squared x offsets of the row from
// the end points plus their squared vertical gaps to the window's height
// range) and c1, c2 the end points' y in the window's frame.  h is convex, so
// its tangent at any y~ in [0, w] lies below it there:
//   min h >= h(y~) + min(h'(y~) (0 - y~), h'(y~) (w - y~)),
// which is tight when y~ is near the minimiser: y~ is one Newton step on h'
// from the straight line's crossing of the row (clamped to [0, w]).  FP32
// with v_rsq_f32: h within ~6 ulp relative, h' within ~8 ulp of n1 + n2
// absolute; the result is deflated by kDefl and by `slack` (row_slack:
// 1e-5 (n1 + n2) (w + 1) for those and for candidates up to 1e-5 outside
// [0, w], plus 2^-22 (n1 |c1| + n2 |c2|) for c1, c2 rounded to FP32).
// Non-finite or degenerate input gives NaN (no pruning).
__device__ __forceinline__ float row_slack(float c1, float c2, float w, float n1, float n2)
{
    return 1e-5f * (n1 + n2) * (w + 1.0f) + 0x1p-22f * (n1 * __builtin_fabsf(c1) + n2 * __builtin_fabsf(c2));
}

__device__ __forceinline__ float row_bound2d(float a1, float a2, float c1, float c2, float w, float n1, float n2,
                                             float slack, float& upper)
{
    // start: the straight line between the end points, weighted by their
    // distances from the row
    const float r1 = __builtin_amdgcn_rsqf(a1), r2 = __builtin_amdgcn_rsqf(a2);
    const float f = r2 * __builtin_amdgcn_rcpf(r1 + r2);   // sqrt(a1) / (sqrt(a1) + sqrt(a2))
    float y = __builtin_fmaf(c2 - c1, f, c1);
    y = __builtin_fminf(__builtin_fmaxf(y, 0.0f), w);
    {
        const float t1 = y - c1, t2 = y - c2;
        const float q1 = __builtin_fmaf(t1, t1, a1), q2 = __builtin_fmaf(t2, t2, a2);
        const float s1 = __builtin_amdgcn_rsqf(q1), s2 = __builtin_amdgcn_rsqf(q2);
        const float u1 = n1 * s1, u2 = n2 * s2;
        const float d1 = __builtin_fmaf(u1, t1, u2 * t2);                              // h'
        const float d2 = __builtin_fmaf(u1 * a1, s1 * s1, (u2 * a2) * (s2 * s2));      // h''
        y = __builtin_fminf(__builtin_fmaxf(y - d1 * __builtin_amdgcn_rcpf(d2), 0.0f), w);
    }
    const float t1 = y - c1, t2 = y - c2;
    const float q1 = __builtin_fmaf(t1, t1, a1), q2 = __builtin_fmaf(t2, t2, a2);
    const float s1 = __builtin_amdgcn_rsqf(q1), s2 = __builtin_amdgcn_rsqf(q2);
    const float h = __builtin_fmaf(n1 * q1, s1, (n2 * q2) * s2);
    const float d = __builtin_fmaf(n1 * t1, s1, (n2 * t2) * s2);
    const float drop = __builtin_fminf(d * (0.0f - y), d * (w - y));
    upper = __builtin_fmaf(h, 1.0f + 2e-5f, slack);   // >= h(y~) >= the row's least path
    return __builtin_fmaf(h, (float)kDefl, drop) - slack;
}

template <int G, bool kSplit, bool kPairs, bool kSide, bool kReuse>
__device__ __forceinline__ void fermat_level(const DemDev& d, const Leg& L, double R, double& bestX, double& bestY,
                                             double& minOpt, int lev)
{
    constexpr int C = (G - 1) / 2;
    const double minX = bestX - R, maxX = bestX + R;
    const double minY = bestY - R, maxY = bestY + R;
    const double spanX = maxX - minX, spanY = maxY - minY;
    const bool empty = (d.flags & DEM_EMPTY) != 0;
    const bool ysorted = (d.flags & DEM_Y_SORTED) != 0;
    const bool xsorted = (d.flags & DEM_X_SORTED) != 0;
    const double stepX = spanX / (G - 1), stepY = spanY / (G - 1);
    const double epsx = 1e-12 * (fabs(minX) + fabs(spanX) + 1.0);
    const double epsy = 1e-12 * (fabs(minY) + fabs(spanY) + 1.0);
    // vertical gaps for the bounds: the DEM's global height range, or the
    // range inside this level's window once the window is small
    double g1sq = L.g1sq, g2sq = L.g2sq;
    if (G == 11 && !empty && xsorted && ysorted)
        window_gaps<kPairs>(d, L.z1, L.z2, minX - epsx, maxX + epsx, minY - epsy, maxY + epsy, g1sq, g2sq);
    const float g1f = (float)g1sq, g2f = (float)g2sq;

    // Probe: the centre candidate (the previous level's best) evaluated
    // exactly first, so the bounds below start from a tight value.  When its
    // grid coordinates round to exactly the previous best's (about half the
    // time), its value is already known: minOpt was computed at those very
    // coordinates by the same expression (minOpt < 1e9: a previous level set
    // it; the first level's centre is the unevaluated midpoint).  A tie with
    // minOpt never updates the result (strict '<' across levels), so starting
    // the level from it is the same as evaluating it.  kReuse: on except in
    // the plain FP64 back-projection kernel, which measured faster without it.
    double lvMin, M;
    int lvK = C * G + C;
    {
        const double px = minX + C * spanX / (G - 1), py = minY + C * spanY / (G - 1);
        if (kReuse && px == bestX && py == bestY && minOpt < 1e9) {
            lvMin = minOpt;
        } else {
            int i0 = 0, di = 0;
            double wx = 0.0;
            if (!empty) axis_lookup(d.X, d.nx, d.x0, d.ihx, xsorted, px, i0, di, wx);
            const double dx1 = px - L.x1, dx2 = L.x2 - px;
            lvMin = exact_opt(d, L, empty, ysorted, i0, di, wx, dx1 * dx1, dx2 * dx2, py);
            if (!(lvMin == lvMin)) {          // a NaN path never wins (strict '<' in the reference)
                lvMin = __builtin_inf();
                lvK = 0x7fffffff;
            }
        }
        M = L.screen ? fmin(minOpt, lvMin) : __builtin_inf();
    }

    // The columns jj = JB .. JB + NJ - 1 (centre-out order) against every row.
    // Columns whose lower bound can reach M (bit t <-> column
    // center_out(JB + t), fixed for the part: M only decreases, so later rows
    // may estimate a column the bound would now drop) and their FP32 terms: Z
    // element offsets of the rows j0 and j0 + dj, weight, y offset.
    // A dropped column keeps dyf = +inf, which makes its estimate +inf (or
    // NaN), so the screening loop needs no per-candidate mask test.
    //
    // Side stop.  The one-axis bound is convex in the column (row) position,
    // so once it prunes a column whose bound still rises in the direction of
    // travel -- its slope, taken epsy before the column's true position,
    // exceeds kSlope (n1 + n2) there -- every column further out on that side
    // has an exact bound at least as large, above M now and later: those
    // columns are dropped without evaluating their bounds (bit 0 of cdone the
    // + side, even k; bit 1 the - side, odd k).  Rows do the same.  What is
    // dropped this way could not win or tie, so the argmin is unchanged.
    // kSide: on in the back-projection kernels with the DEM in LDS, where the
    // saved bound evaluations pay; the search-only kernels, at 3 waves per
    // SIMD, would spill the extra state.  kSide also ends a row's screening
    // loop early (below), which measured slower in the search-only kernels.
    const float slope = kSlope * (L.n1f + L.n2f);
    const float dyL = (float)(L.y2 - L.y1);
    unsigned cdone = 0u;
    auto part = [&](auto jb, auto nj) {
    constexpr int JB = decltype(jb)::value, NJ = decltype(nj)::value;
    // kPairs (DEMs staged in LDS, fewer than 2^16 pair slots): both row
    // offsets in one word, o0 | o1 << 16; otherwise o0 and o1.  The squared y
    // offsets come from dyf = py - y1 and dyL = y2 - y1 in FP32.
    int o0[NJ];
    int o1[kPairs ? 1 : NJ];
    float wyf[NJ], dyf[NJ];
    unsigned colmask = 0u;
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
        const int kk = JB + t;
        const int j = center_out(kk, C);
        o0[t] = 0;
        if (!kPairs) o1[t] = 0;
        wyf[t] = 0.0f;
        dyf[t] = __builtin_inff();
        const unsigned side = (kk & 1) ? 2u : 1u;
        if (kSide && kk > 0 && (cdone & side)) continue;
        const double pya = __builtin_fma((double)j, stepY, minY);
        const double a = fmax(fabs(pya - L.y1) - epsy, 0.0), b = fmax(fabs(pya - L.y2) - epsy, 0.0);
        // (on the last level, whose 7.4 m-wide window keeps nearly every
        // column, the back-projection kernels (kSide) skip the bound: it
        // cost them more than it saved)
        if ((!kSide || lev < 4) && lower_bound32(a, b, g1f, g2f, L.n1f, L.n2f) > M) {
            if (kSide && kk > 0) {
                const double q = (kk & 1) ? pya + epsy : pya - epsy;
                const float sl = slope32(q - L.y1, q - L.y2, g1f, g2f, L.n1f, L.n2f);
                if ((kk & 1) ? sl < -slope : sl > slope) cdone |= side;
            }
            continue;
        }
        colmask |= 1u << t;
        // Terms of the FP32 estimates only (the exact evaluations look the
        // column up again at its exact position): pya is within epsy of it,
        // and away from the axis ends (where the out-of-range rule of the
        // lookup jumps to row 0) its interval and an FP32 weight serve; near
        // them, or on an unsorted axis, the exact position and lookup.
        int j0 = 0, dj = 0;
        float wf = 0.0f;
        if (!empty) {
            if (ysorted && axis_cell_inside(d.Y, d.ny, d.y0, d.yn, d.ihy, pya, epsy, j0, wf)) {
                dj = 1;
            } else {
                double w = 0.0;
                axis_lookup(d.Y, d.ny, d.y0, d.ihy, ysorted, minY + j * spanY / (G - 1), j0, dj, w);
                wf = (float)w;
            }
        }
        if (kPairs) {
            o0[t] = j0 * d.zs | ((j0 + dj) * d.zs) << 16;
        } else {
            o0[t] = j0 * d.zs;
            o1[t] = (j0 + dj) * d.zs;
        }
        wyf[t] = wf;
        dyf[t] = (float)(pya - L.y1);
    }
    // FP32 screening constants (rounded to FP32: well inside the 2 kRho32 and
    // D32 slack of the bounds)
    const float k3 = (float)(1.0 + 3.0 * kRho32), k2 = (float)(1.0 + 2.0 * kRho32);
    const float Dk3 = (float)(L.D * (1.0 + 3.0 * kRho32)), D2 = (float)(2.0 * L.D);

    unsigned rdone = 0u;   // side stop of the rows (bit 0: + side, bit 1: - side)
    float upP = __builtin_inff(), upM = __builtin_inff();   // row_bound2d's upper value, last row per side
#pragma unroll 1
    for (int k = 0; k < G; ++k) {
        const int i = center_out(k, C);
        // row lower bound from an approximate px (|error| < epsx)
        {
            const unsigned side = (k & 1) ? 2u : 1u;
            if (kSide && k > 0 && (rdone & side)) {
                if (rdone == 3u) break;
                continue;
            }
            const double pxa = __builtin_fma((double)i, stepX, minX);
            const double a1 = fmax(fabs(pxa - L.x1) - epsx, 0.0);
            const double a2 = fmax(fabs(pxa - L.x2) - epsx, 0.0);
            if (lower_bound32(a1, a2, g1f, g2f, L.n1f, L.n2f) > M) {
                if (kSide && k > 0) {
                    const double q = (k & 1) ? pxa + epsx : pxa - epsx;
                    const float sl = slope32(q - L.x1, q - L.x2, g1f, g2f, L.n1f, L.n2f);
                    if ((k & 1) ? sl < -slope : sl > slope) rdone |= side;
                }
                continue;
            }
        }
        // the probe (i, j) = (C, C) is bit 0 of row C in the first part
        const unsigned colok = (JB == 0 && i == C) ? (colmask & ~1u) : colmask;
        if (!colok) continue;
        const double px = minX + i * spanX / (G - 1);
        const double dx1 = px - L.x1, dx2 = L.x2 - px;
        const double dxa = dx1 * dx1, dxb = dx2 * dx2;
        const float dxaf = (float)dxa, dxbf = (float)dxb;
        float Mf = (float)(M * (1.0 + 0x1p-22));   // >= M: an upper bound of the minimum
        // Row bound over both axes (the last two levels of the back-projection
        // kernels' search, once the one-axis bound kept the row): the least
        // path through the row's x at the window's height range, row_bound2d.
        // A row it prunes needs nothing more, not even its x lookup.  On the
        // wider windows of the earlier levels it pruned almost no row the
        // one-axis bound kept, and the search-only kernels (3 waves per SIMD,
        // no kSide) measured 3 % slower with it.
        if (kSide && G == 11 && L.screen && lev >= 3) {
            // end points' y relative to the window's low edge (recomputed per
            // row rather than kept live through the row loop)
            const float wy1 = (float)(L.y1 - minY), wy2 = (float)(L.y2 - minY), wspan = (float)spanY;
            float up;
            const float b = row_bound2d(dxaf + g1f, dxbf + g2f, wy1, wy2, wspan, L.n1f, L.n2f,
                                        row_slack(wy1, wy2, wspan, L.n1f, L.n2f), up);
            // Side stop: the row's least path m(x) is convex in x (a minimum
            // over y of a jointly convex path), so when this pruned row's
            // bound exceeds the upper value of the previous row on its side,
            // m rises outward from here and every further row on that side is
            // above Mf too (Mf only decreases).
            const bool minus = (k & 1) != 0;
            if (b > Mf) {
                if (kSide && k > 0 && b > (minus ? upM : upP)) rdone |= minus ? 2u : 1u;
                if (minus) upM = up; else upP = up;
                continue;
            }
            if (k == 0) upM = up;
            if (minus) upM = up; else upP = up;
        }
        int i0 = 0, di = 0;
        double wx = 0.0;
        if (!empty) axis_lookup(d.X, d.nx, d.x0, d.ihx, xsorted, px, i0, di, wx);
        // Screening pass, branch-free: every candidate of the row is
        // estimated from the node pairs (i0, i0 + 1) of rows o0 / o1 (in
        // bounds: i0 + 1 <= nx - 1, or the zero padding column), so the LDS
        // loads of a row issue together.  When the x interval was not found
        // (di = 0) the second node gets weight 0: a non-finite neighbour then
        // makes the estimate NaN, which only sends the candidate to the exact
        // evaluation.  Bit t of pass: candidate (i, center_out(JB + t)) may
        // still be the minimum.
        unsigned pass = colok;
        if (L.screen) {
            pass = 0u;
            const float* zr = d.Zf + i0;
            const float2* zp = d.Zp + i0;
            const float wxf = di ? (float)wx : 0.0f;
            // Columns past the row's last kept one (centre-out order) have
            // +inf (or NaN) estimates that neither lower Mf nor pass colok, so
            // with kSide the loop stops at a chunk of four columns once no
            // lane of the wave keeps one in it; loads within a chunk still
            // issue together.
            const int nact = 32 - __builtin_clz(colok);   // colok != 0
#pragma unroll
            for (int t = 0; t < NJ; ++t) {
                if (kSide && t > 0 && (t & 3) == 0 && nact <= t) break;
                float za, zb;
                if (kPairs) {
                    // (z, dz/dx) pairs: one 8-byte LDS load per row of the cell
                    const float2 a0 = zp[o0[t] & 0xffff], a1 = zp[(unsigned)o0[t] >> 16];
                    za = __builtin_fmaf(a0.y, wxf, a0.x);
                    zb = __builtin_fmaf(a1.y, wxf, a1.x);
                } else {
                    const float* p0 = zr + o0[t];
                    const float* p1 = zr + o1[t];
                    const float z00 = p0[0], z01 = p0[1];
                    const float z10 = p1[0], z11 = p1[1];
                    za = __builtin_fmaf(z01 - z00, wxf, z00);
                    zb = __builtin_fmaf(z11 - z10, wxf, z10);
                }
                const float ze = __builtin_fmaf(zb - za, wyf[t], za);
                const float e1 = ze - L.z1f, e2 = L.z2f - ze;
                const float dy2 = dyL - dyf[t];
                const float A = __builtin_fmaf(e1, e1, __builtin_fmaf(dyf[t], dyf[t], dxaf));
                const float B = __builtin_fmaf(e2, e2, __builtin_fmaf(dy2, dy2, dxbf));
                const float os = __builtin_fmaf(L.n2f, sqrt_f32(B), L.n1f * sqrt_f32(A));
                // Mf = min(Mf, (est + D32)(1 + 3 kRho32)); NaN and +inf (dropped
                // column) estimates leave it; the probe's estimate bounds its
                // exact value, which M already includes
                Mf = __builtin_fminf(__builtin_fmaf(os, k3, Dk3), Mf);
                const bool keep = !(os > __builtin_fmaf(Mf, k2, D2));
                pass |= keep ? (1u << t) : 0u;
            }
            pass &= colok;
            M = fmin(M, (double)Mf);
        }
        // exact evaluation of the survivors
        while (pass) {
            const int t = __builtin_ctz(pass);
            pass &= pass - 1u;
            const int j = center_out(JB + t, C);
            const double opt = exact_opt(d, L, empty, ysorted, i0, di, wx, dxa, dxb, minY + j * spanY / (G - 1));
            const int kr = i * G + j;
            if (opt < lvMin || (opt == lvMin && kr < lvK)) { lvMin = opt; lvK = kr; }
            if (L.screen) M = fmin(M, lvMin);
        }
    }
    };
    // kSplit: a 15-wide level runs its columns in two parts (8 + 7), which
    // keeps the per-column arrays at the size of the 11-wide levels' (for
    // kernels whose register budget would otherwise spill them)
    if (kSplit && G == 15) {
        part(std::integral_constant<int, 0>{}, std::integral_constant<int, 8>{});
        part(std::integral_constant<int, 8>{}, std::integral_constant<int, G - 8>{});
    } else {
        part(std::integral_constant<int, 0>{}, std::integral_constant<int, G>{});
    }
    if (lvMin < minOpt) {
        minOpt = lvMin;
        const int bi = lvK / G, bj = lvK - (lvK / G) * G;
        bestX = minX + bi * spanX / (G - 1);
        bestY = minY + bj * spanY / (G - 1);
    }
}

// calculateRefractionPoint, raw_data_processing.cpp:455-572.
template <bool kSplit = false, bool kPairs = false, bool kSide = false, bool kReuse = true>
__device__ __forceinline__ void fermat_solve(const DemDev& d, double x1, double y1, double z1,
                                             double x2, double y2, double z2, double n1, double n2,
                                             double& qx, double& qy, double& qz)
{
    double bestX = (x1 + x2) / 2.0;                                  // :478-479
    double bestY = (y1 + y2) / 2.0;
    const double ex = x2 - x1, ey = y2 - y1;
    double R = std_max(__builtin_sqrt(ex * ex + ey * ey), 500.0);    // :483-484
    double minOpt = 1e9;                                             // :497
    const Leg L = make_leg(d, x1, y1, z1, x2, y2, z2, n1, n2);
    fermat_level<15, kSplit, kPairs, kSide, kReuse>(d, L, R, bestX, bestY, minOpt, 0);
    R /= 5.0;                                                        // :551
#pragma unroll 1
    for (int level = 1; level < 5; ++level) {
        fermat_level<11, kSplit, kPairs, kSide, kReuse>(d, L, R, bestX, bestY, minOpt, level);
        R /= 3.0;
    }
    qx = bestX;                                                      // :555-557
    qy = bestY;
    qz = dem_interp(d, bestX, bestY);
}

// 8-byte LDS slots of a staged DEM: X, Y, Z (FP64), then Zp (FP32 pairs).
__host__ __device__ __forceinline__ int dem_lds_slots(const DemDev& g)
{
    return g.nx + g.ny + 2 * g.ny * g.zs;
}

// View of a DEM staged at lds (no copy).
__device__ __forceinline__ DemDev dem_lds_view(const DemDev& g, double* lds)
{
    DemDev d = g;
    d.X = lds;
    d.Y = lds + g.nx;
    d.Z = lds + g.nx + g.ny;
    d.Zp = reinterpret_cast<const float2*>(lds + g.nx + g.ny + g.ny * g.zs);
    return d;
}

// Cooperative copy of the search's hot part of a DEM into LDS: X, Y (FP64)
// and the FP32 heights; the FP64 heights (read by exact evaluations only)
// stay in global memory.  Slots: nx + ny + ceil(ny * zs / 2).
__host__ __device__ __forceinline__ int dem_split_slots(const DemDev& g)
{
    return g.nx + g.ny + (g.ny * g.zs + 1) / 2;
}

__device__ __forceinline__ DemDev dem_to_lds_split(const DemDev& g, double* lds)
{
    const int nxy = g.nx + g.ny;
    for (int k = threadIdx.x; k < nxy; k += blockDim.x) lds[k] = (k < g.nx) ? g.X[k] : g.Y[k - g.nx];
    float* zf = reinterpret_cast<float*>(lds + nxy);
    const int nz = g.ny * g.zs;
    for (int k = threadIdx.x; k < nz; k += blockDim.x) zf[k] = g.Zf[k];
    DemDev d = g;
    d.X = lds;
    d.Y = lds + g.nx;
    d.Zf = zf;
    return d;
}

// Cooperative copy of the DEM into LDS; returns a view onto the copy.
__device__ __forceinline__ DemDev dem_to_lds(const DemDev& g, double* lds)
{
    const int nz = g.ny * g.zs;
    const int total = g.nx + g.ny + nz;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        double v;
        if (k < g.nx) v = g.X[k];
        else if (k < g.nx + g.ny) v = g.Y[k - g.nx];
        else v = g.Z[k - g.nx - g.ny];
        lds[k] = v;
    }
    float2* zp = reinterpret_cast<float2*>(lds + total);
    for (int k = threadIdx.x; k < nz; k += blockDim.x) zp[k] = g.Zp[k];
    return dem_lds_view(g, lds);
}

}  // namespace bsar

namespace bsar {

// Where a kernel finds the DEM (template argument of the search kernels).
enum DemMode : int { DEM_GLOBAL = 0, DEM_LDS = 1, DEM_SPLIT = 2 };

// Stages the DEM for mode kMode (all threads of the block call it).
template <int kMode>
__device__ __forceinline__ DemDev dem_stage(const DemDev& g, double* lds)
{
    if (kMode == DEM_LDS) {
        const DemDev d = dem_to_lds(g, lds);
        __syncthreads();
        return d;
    }
    if (kMode == DEM_SPLIT) {
        const DemDev d = dem_to_lds_split(g, lds);
        __syncthreads();
        return d;
    }
    return g;
}

}  // namespace bsar

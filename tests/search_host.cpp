// The refraction search of c-_bistatic_sar_amd/csrc/bsar_device.h
// (fermat_solve) compiled for the CPU through tests/host_shim, run over legs
// read from a file: the CPU test of the search's pruning bounds against the
// oracle's exhaustive 709-candidate search (tests/test_search_host.py).
//
//   usage: search_host IN OUT CONFIG [PLANT]
//     IN:  int32 nx, ny; int64 m; X[nx], Y[ny], Z[ny][nx] (float64);
//          m legs of 8 float64 (x1 y1 z1 x2 y2 z2 n1 n2)
//     OUT: m x 3 float64 refraction points
//     CONFIG: the template arguments of a kernel's search (see the switch)
//     PLANT: a factor on the FP32 reciprocal root of the search's lower
//            bounds (the test's own sanity check: above 1 the bounds overshoot
//            and must drop true minima), default 1
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bsar_device.h"

using namespace bsar;

float bsar_host_rsqf_scale = 1.0f;

static bool increasing(const std::vector<double>& a)
{
    for (double v : a)
        if (!std::isfinite(v)) return false;
    for (size_t i = 0; i + 1 < a.size(); ++i)
        if (!(a[i] < a[i + 1])) return false;
    return true;
}

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int nx = 0, ny = 0;
    long m = 0;
    if (fread(&nx, 4, 1, f) != 1 || fread(&ny, 4, 1, f) != 1 || fread(&m, 8, 1, f) != 1) return 2;
    std::vector<double> X(nx), Y(ny), Zr((size_t)nx * ny), legs(8 * (size_t)m);
    if (fread(X.data(), 8, nx, f) != (size_t)nx || fread(Y.data(), 8, ny, f) != (size_t)ny ||
        fread(Zr.data(), 8, Zr.size(), f) != Zr.size() || fread(legs.data(), 8, legs.size(), f) != legs.size())
        return 2;
    fclose(f);
    const int cfg = atoi(argv[3]);
    if (argc > 4) bsar_host_rsqf_scale = (float)atof(argv[4]);
    // the device DEM exactly as the library uploads it (api.cpp upload_dem)
    const int zs = nx + 1;
    std::vector<double> Z((size_t)ny * zs, 0.0);
    double lo = INFINITY, hi = -INFINITY;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double v = Zr[(size_t)j * nx + i];
            Z[(size_t)j * zs + i] = v;
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
    std::vector<float> Zf(Z.size());
    for (size_t k = 0; k < Z.size(); ++k) Zf[k] = (float)Z[k];
    std::vector<float2> Zp(Z.size());
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < zs; ++i) {
            const size_t k = (size_t)j * zs + i;
            Zp[k] = {(float)Z[k], i + 1 < zs ? (float)(Z[k + 1] - Z[k]) : 0.0f};
        }
    DemDev d{};
    d.X = X.data(); d.Y = Y.data(); d.Z = Z.data(); d.Zf = Zf.data(); d.Zp = Zp.data();
    d.nx = nx; d.ny = ny; d.zs = zs; d.flags = 0;
    if (increasing(X)) d.flags |= DEM_X_SORTED;
    if (increasing(Y)) d.flags |= DEM_Y_SORTED;
    d.x0 = X[0]; d.y0 = Y[0]; d.xn = X[nx - 1]; d.yn = Y[ny - 1];
    d.ihx = (nx > 1 && (d.flags & DEM_X_SORTED)) ? (nx - 1) / (X[nx - 1] - X[0]) : 0.0;
    d.ihy = (ny > 1 && (d.flags & DEM_Y_SORTED)) ? (ny - 1) / (Y[ny - 1] - Y[0]) : 0.0;
    d.zlo = lo; d.zhi = hi;
    d.zabs = std::max(std::fabs(lo), std::fabs(hi));
    std::vector<double> out(3 * (size_t)m);
    for (long k = 0; k < m; ++k) {
        const double* L = &legs[8 * (size_t)k];
        double* q = &out[3 * (size_t)k];
        switch (cfg) {
        // mixed back-projection, DEM in LDS (bp_mixed_thread)
        case 0: fermat_solve<false, true, true, true>(d, L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], q[0], q[1], q[2]); break;
        // plain FP64 back-projection, DEM in LDS (bp_fp64_thread)
        case 1: fermat_solve<true, true, true, false>(d, L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], q[0], q[1], q[2]); break;
        // refraction / slant-range / forward-model kernels, DEM in LDS
        case 2: fermat_solve<false, true, false, true>(d, L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], q[0], q[1], q[2]); break;
        // every search with the DEM in global memory or split (FP32 heights)
        case 3: fermat_solve<false, false, false, true>(d, L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], q[0], q[1], q[2]); break;
        // plain FP64 back-projection with the DEM outside LDS
        case 4: fermat_solve<false, false, false, false>(d, L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], q[0], q[1], q[2]); break;
        default: return 2;
        }
    }
    f = fopen(argv[2], "wb");
    if (!f || fwrite(out.data(), 8, out.size(), f) != out.size()) return 2;
    fclose(f);
    return 0;
}

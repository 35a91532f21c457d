// api_internal.h -- host-side state behind the C ABI (bsar.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_kernels.h"

struct bsar_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Last DEM of bsar_dem_interp / bsar_refraction_points, resident on the
    // device with a host copy of its arrays: a call with the same DEM contents
    // reuses it (the facade calls these once per voxel).
    struct {
        bool valid = false;
        int nx = 0, ny = 0, zrows = 0, zcols = 0;
        std::vector<double> x, y, z;
        double *dX = nullptr, *dY = nullptr, *dZ = nullptr;
        bsar::DemDev dem{};
    } dem_cache;
    // grow-only point / result buffers of the same calls
    void* in_buf = nullptr;
    size_t in_cap = 0;
    void* out_buf = nullptr;
    size_t out_cap = 0;
};

struct bsar_dataset {
    bsar_ctx* owner = nullptr;
    int device = 0;
    bsar_params p{};
    int n = 0;                       // trajectory pulses
    // DEM on device
    double* dX = nullptr;
    double* dY = nullptr;
    double* dZ = nullptr;
    bsar::DemDev dem{};
    double zmin = 0.0, zmax = 0.0;   // DEM height range (0 when empty)
    // antennas (SoA) on device + host copy
    double* dtraj = nullptr;         // 14 * n doubles: positions, then R1Geo's terms
    std::vector<double> htx, hrx;    // [n][3]
    bsar::Traj traj{};
    bsar::R1Geo r1{};                // expansion terms of the plain form's R1 sums (in dtraj)
    std::vector<double> r1_bsq, r1_zlo, r1_zhi;   // their antenna bounds over pulses 0..i
    // upsampled range-compressed data
    double2* rc = nullptr;
    float2* rc32 = nullptr;
    int ndata = 0, S4 = 0;
    int64_t bytes = 0;
    // R1-mean tables (mixed form), built lazily per (dataset, pulse count)
    void* r1tab = nullptr;
    std::mutex tab_mu;               // guards r1tab (the serving thread shares datasets)
};

extern "C" void bsar_r1tab_free(bsar_dataset* ds);

// Upsampled range data fully loaded (both copies, at least one pulse): a load
// that failed part-way leaves the dataset without, so nothing images it.
inline bool rcomp_ready(const bsar_dataset* ds) { return ds->rc && ds->rc32 && ds->ndata > 0 && ds->S4 > 0; }

#define BSAR_CHECK_HIP(ctx, call)                                     \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return ::bsar::hip_fail((ctx), e_, #call); \
    } while (0)

namespace bsar {

int set_error(bsar_ctx* ctx, int status, const std::string& msg);

// Looks up, once per device and with that device current, the device address
// of the no-allocation entry point's file-scope scratch array (rc.cpp).
int scratch_init(int device);
int hip_fail(bsar_ctx* ctx, hipError_t e, const char* what);

// Builds the device DEM description from the caller's arrays (validation,
// padding of missing Z entries with 0.0, sortedness flags, z range).
int upload_dem(bsar_ctx* ctx, const bsar_dem* dem, double** dX, double** dY, double** dZ, DemDev* out,
               double* zmin, double* zmax);

// Work-splitting of one back-projection: pulse groups (partial images summed
// in group order at the end) and the pulse window of one launch.  The group
// split depends only on (nvox, n_pulses), so every caller that plans the
// same image sums each voxel's pulses in the same order.
struct BpPlan {
    int groups = 1, glen = 0, wlen = 0;
    int64_t nvox = 0, nvox_pad = 0;
};
BpPlan plan_bp(int64_t nvox, int n_pulses, double cost_per_vp);

// Lane-instruction estimates per voxel-pulse (launch sizing and scheduling).
inline double bp_cost_fp64(int n) { return 1418.0 * 60.0 + 2.0 * n * 26.0 + 200.0; }
inline double bp_cost_mixed() { return 1418.0 * 60.0 + 600.0; }

// Argument block of one image over the first K pulses of ds (device axes
// dg = gx | gy | gz) with plan pl; accumulators are set by the caller.
BpArgs make_bp_args(const bsar_dataset* ds, int K, const double* dg, int nx, int ny, int nz, const BpPlan& pl);

hipError_t fill_zero(double2* p, int64_t n, hipStream_t s);

// Final image from the group accumulators: out[v] = ordered sum over groups.
// out may be acc itself (16 bytes per voxel in both forms).
hipError_t finish_fp64(const double2* acc, int groups, int64_t nvox, double2* out, hipStream_t s);
hipError_t finish_mixed(const float4* acc, int groups, int64_t nvox, double2* out, hipStream_t s);

// Mixed-precision back-projection (kernels_mixed.hip / mixed.cpp).
int backproject_mixed(bsar_ctx* ctx, const bsar_dataset* ds, const double* gx, int nx, const double* gy, int ny,
                      const double* gz, int nz, double2* d_out, bsar_report* rep);

// R1 tables for the first K pulses of ds, with the lookup box of this grid's
// refraction points (built on ctx's stream when missing; ds->tab_mu held).
// The device arrays stay valid until the dataset is destroyed.
int r1_tables(bsar_ctx* ctx, bsar_dataset* ds, int K, const double* gx, int nx, const double* gy, int ny,
              const double* gz, int nz, R1Table* tab, const double** ftx, const double** frx);

}  // namespace bsar

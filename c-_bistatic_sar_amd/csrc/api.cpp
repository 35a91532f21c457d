// api.cpp -- C ABI of the imaging hot path (declarations and reference
// citations in include/bsar.h).
#include "api_internal.h"
#include "bsar_fft.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>

namespace bsar {

int set_error(bsar_ctx* ctx, int status, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return status;
}

int hip_fail(bsar_ctx* ctx, hipError_t e, const char* what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, e == hipErrorOutOfMemory ? BSAR_ERR_NOMEM : BSAR_ERR_DEVICE, m);
}

#define BSAR_HIP(ctx, call) BSAR_CHECK_HIP(ctx, call)

static bool strictly_increasing(const double* a, int n)
{
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    for (int i = 0; i + 1 < n; ++i)
        if (!(a[i] < a[i + 1])) return false;
    return true;
}

int upload_dem(bsar_ctx* ctx, const bsar_dem* dem, double** dX, double** dY, double** dZ, DemDev* out,
               double* zmin, double* zmax)
{
    if (!dem) return set_error(ctx, BSAR_ERR_INVALID, "dem is null");
    if (dem->nx < 0 || dem->ny < 0 || dem->zrows < 0 || dem->zcols < 0)
        return set_error(ctx, BSAR_ERR_INVALID, "negative DEM size");
    if ((dem->nx && !dem->x) || (dem->ny && !dem->y) || (dem->zrows && dem->zcols && !dem->z))
        return set_error(ctx, BSAR_ERR_INVALID, "DEM array pointer is null");
    const bool empty = dem->nx == 0 || dem->ny == 0 || dem->zrows == 0 || dem->zcols == 0;
    const int nx = std::max(dem->nx, 1), ny = std::max(dem->ny, 1);
    const int zs = nx + 1;
    std::vector<double> X(nx, 0.0), Y(ny, 0.0), Z((size_t)ny * zs, 0.0);
    if (dem->nx) std::copy(dem->x, dem->x + dem->nx, X.begin());
    if (dem->ny) std::copy(dem->y, dem->y + dem->ny, Y.begin());
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    if (!empty) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double v = (j < dem->zrows && i < dem->zcols) ? dem->z[(size_t)j * dem->zcols + i] : 0.0;
                Z[(size_t)j * zs + i] = v;
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
    } else {
        lo = hi = 0.0;
    }
    *zmin = lo;
    *zmax = hi;
    // Z in FP64, followed in the same allocation by its FP32 rounding (the
    // refraction search's screening estimates read only the FP32 copy)
    std::vector<float> Zf(Z.size());
    for (size_t k = 0; k < Z.size(); ++k) Zf[k] = (float)Z[k];
    // ... and by (height, x difference to the next node) pairs in FP32, one
    // rounding each, for the screening of DEMs staged whole in LDS
    std::vector<float> Zp(2 * Z.size());
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < zs; ++i) {
            const size_t k = (size_t)j * zs + i;
            Zp[2 * k] = (float)Z[k];
            Zp[2 * k + 1] = i + 1 < zs ? (float)(Z[k + 1] - Z[k]) : 0.0f;
        }
    size_t bx = X.size() * sizeof(double), by = Y.size() * sizeof(double), bz = Z.size() * sizeof(double);
    size_t bzf = (Zf.size() * sizeof(float) + 7) / 8 * 8, bzp = Zp.size() * sizeof(float);
    BSAR_HIP(ctx, hipMalloc((void**)dX, bx));
    BSAR_HIP(ctx, hipMalloc((void**)dY, by));
    BSAR_HIP(ctx, hipMalloc((void**)dZ, bz + bzf + bzp));
    BSAR_HIP(ctx, hipMemcpyAsync(*dX, X.data(), bx, hipMemcpyHostToDevice, ctx->stream));
    BSAR_HIP(ctx, hipMemcpyAsync(*dY, Y.data(), by, hipMemcpyHostToDevice, ctx->stream));
    BSAR_HIP(ctx, hipMemcpyAsync(*dZ, Z.data(), bz, hipMemcpyHostToDevice, ctx->stream));
    BSAR_HIP(ctx, hipMemcpyAsync((char*)*dZ + bz, Zf.data(), Zf.size() * sizeof(float), hipMemcpyHostToDevice,
                                 ctx->stream));
    BSAR_HIP(ctx, hipMemcpyAsync((char*)*dZ + bz + bzf, Zp.data(), bzp, hipMemcpyHostToDevice, ctx->stream));
    BSAR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DemDev d{};
    d.X = *dX;
    d.Y = *dY;
    d.Z = *dZ;
    d.Zf = reinterpret_cast<const float*>((char*)*dZ + bz);
    d.Zp = reinterpret_cast<const float2*>((char*)*dZ + bz + bzf);
    d.nx = dem->nx;
    d.ny = dem->ny;
    d.zs = zs;
    d.flags = 0;
    if (empty) {
        d.flags |= DEM_EMPTY;
        d.nx = nx;
        d.ny = ny;
    }
    if (!empty && strictly_increasing(X.data(), dem->nx)) d.flags |= DEM_X_SORTED;
    if (!empty && strictly_increasing(Y.data(), dem->ny)) d.flags |= DEM_Y_SORTED;
    d.x0 = X[0];
    d.y0 = Y[0];
    d.xn = X[nx - 1];
    d.yn = Y[ny - 1];
    d.ihx = (dem->nx > 1 && (d.flags & DEM_X_SORTED)) ? (dem->nx - 1) / (X[dem->nx - 1] - X[0]) : 0.0;
    d.ihy = (dem->ny > 1 && (d.flags & DEM_Y_SORTED)) ? (dem->ny - 1) / (Y[dem->ny - 1] - Y[0]) : 0.0;
    d.zlo = lo;
    d.zhi = hi;
    d.zabs = std::max(std::fabs(lo), std::fabs(hi));
    *out = d;
    return BSAR_OK;
}

BpPlan plan_bp(int64_t nvox, int n_pulses, double cost_per_vp)
{
    BpPlan p;
    p.nvox = nvox;
    p.nvox_pad = (nvox + 63) / 64 * 64;
    if (p.nvox_pad <= 0) {            // empty grid: nothing is launched
        p.nvox_pad = 64;
        p.glen = std::max(n_pulses, 1);
        p.wlen = p.glen;
        return p;
    }
    const int64_t target_threads = int64_t(1) << 19;
    int64_t groups = (target_threads + p.nvox_pad - 1) / p.nvox_pad;
    groups = std::max<int64_t>(1, std::min<int64_t>(groups, std::max(n_pulses, 1)));
    p.glen = (int)((n_pulses + groups - 1) / groups);
    if (p.glen < 1) p.glen = 1;
    p.groups = (int)((n_pulses + p.glen - 1) / p.glen);
    if (p.groups < 1) p.groups = 1;
    // Launch length: about 0.25 s of device work at a conservative 1.5e13
    // lane-instructions per second.
    const double work_per_pulse = (double)p.groups * (double)p.nvox_pad * cost_per_vp;
    double w = 0.25 * 1.5e13 / std::max(work_per_pulse, 1.0);
    p.wlen = (int)std::max(1.0, std::min((double)p.glen, std::floor(w)));
    return p;
}

// Per-launch fill (no hipMemsetAsync: callers may capture launches in graphs).
__global__ void k_fill2(double2* p, int64_t n, double2 v)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

hipError_t fill_zero(double2* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, make_double2(0.0, 0.0));
    return hipGetLastError();
}

// Device scratch that frees itself.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

static int check_axes(bsar_ctx* ctx, const double* gx, int nx, const double* gy, int ny, const double* gz, int nz)
{
    if (nx < 0 || ny < 0 || nz < 0) return set_error(ctx, BSAR_ERR_INVALID, "negative grid size");
    if ((nx && !gx) || (ny && !gy) || (nz && !gz)) return set_error(ctx, BSAR_ERR_INVALID, "grid axis is null");
    return BSAR_OK;
}

BpArgs make_bp_args(const bsar_dataset* ds, int K, const double* dg, int nx, int ny, int nz, const BpPlan& pl)
{
    BpArgs a{};
    a.dem = ds->dem;
    a.traj = ds->traj;
    a.traj.n = K;                                  // first K pulses; R1 means over those
    a.r1 = ds->r1;
    if (K > 0 && K <= ds->n) {                     // antenna bounds of the first K pulses
        a.r1.bsq = ds->r1_bsq[K - 1];
        a.r1.zlo = ds->r1_zlo[K - 1];
        a.r1.zhi = ds->r1_zhi[K - 1];
    }
    a.rc = ds->rc;
    a.rc32 = ds->rc32;
    a.ndata = ds->ndata;
    a.S4 = ds->S4;
    a.s4m1 = (double)(ds->S4 - 1);
    a.n1 = ds->p.n1;
    a.n2 = ds->p.n2;
    a.inv_c = 1.0 / ds->p.c;                       // proc_test_script.cpp:567
    a.fs4 = ds->p.fs * ds->p.upsample;             // proc_test_script.cpp:223
    const double lambda = ds->p.c / ds->p.fc;      // proc_test_script.cpp:213
    a.kphi = 2.0 * M_PI / lambda;                  // proc_test_script.cpp:575
    a.gx = dg;
    a.gy = dg + nx;
    a.gz = dg + nx + ny;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.vox0 = 0;
    a.nvox = pl.nvox;
    a.nvox_pad = pl.nvox_pad;
    a.groups = pl.groups;
    a.glen = pl.glen;
    a.woff = 0;
    a.wlen = pl.wlen;
    a.acc = nullptr;
    a.acc_stride = pl.nvox;
    return a;
}

hipError_t finish_fp64(const double2* acc, int groups, int64_t nvox, double2* out, hipStream_t s)
{
    if (nvox <= 0) return hipSuccess;
    if (groups > 1) return launch_group_sum(acc, groups, nvox, nvox, out, s);
    if (acc == out) return hipSuccess;
    return hipMemcpyAsync(out, acc, (size_t)nvox * sizeof(double2), hipMemcpyDeviceToDevice, s);
}

static int backproject_fp64(bsar_ctx* ctx, const bsar_dataset* ds, const double* gx, int nx, const double* gy,
                            int ny, const double* gz, int nz, double2* d_out, bsar_report* rep)
{
    const int64_t V = (int64_t)nx * ny * nz;
    const int n = ds->n;
    BpPlan pl = plan_bp(V, n, bp_cost_fp64(n));
    DevBuf axes, accb;
    BSAR_HIP(ctx, hipMalloc(&axes.p, (size_t)(nx + ny + nz + 1) * sizeof(double)));
    double* dgx = (double*)axes.p;
    BSAR_HIP(ctx, hipMemcpyAsync(dgx, gx, nx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    BSAR_HIP(ctx, hipMemcpyAsync(dgx + nx, gy, ny * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    BSAR_HIP(ctx, hipMemcpyAsync(dgx + nx + ny, gz, nz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double2* acc = d_out;
    if (pl.groups > 1) {
        BSAR_HIP(ctx, hipMalloc(&accb.p, (size_t)pl.groups * V * sizeof(double2)));
        acc = (double2*)accb.p;
    }
    BSAR_HIP(ctx, fill_zero(acc, (int64_t)pl.groups * V, ctx->stream));
    BpArgs a = make_bp_args(ds, n, dgx, nx, ny, nz, pl);
    a.acc = acc;

    int launches = 0;
    BSAR_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (V > 0) {
        for (int woff = 0; woff < pl.glen; woff += pl.wlen) {
            a.woff = woff;
            BSAR_HIP(ctx, launch_bp_fp64(a, ctx->stream));
            ++launches;
        }
        if (pl.groups > 1) {
            BSAR_HIP(ctx, finish_fp64(acc, pl.groups, V, d_out, ctx->stream));
            ++launches;
        }
    }
    BSAR_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    BSAR_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (rep) {
        float ms = 0.f;
        BSAR_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        rep->seconds = ms * 1e-3;
        rep->voxel_pulses = (double)V * n;
        rep->voxels = V;
        rep->pulses = n;
        rep->launches = launches;
        rep->r1_exact_frac = 1.0;
    }
    return BSAR_OK;
}

}  // namespace bsar

using namespace bsar;

extern "C" {

int bsar_abi_version(void) { return BSAR_ABI_VERSION; }

const char* bsar_status_string(int s)
{
    switch (s) {
    case BSAR_OK: return "ok";
    case BSAR_ERR_INVALID: return "invalid argument";
    case BSAR_ERR_DEVICE: return "device error";
    case BSAR_ERR_NOMEM: return "out of memory";
    case BSAR_ERR_IO: return "i/o error";
    case BSAR_ERR_FORMAT: return "format error";
    case BSAR_ERR_CAPACITY: return "capacity exceeded";
    case BSAR_ERR_STATE: return "invalid state";
    default: return "unknown status";
    }
}

int bsar_ctx_create(int device, bsar_ctx** out)
{
    if (!out) return BSAR_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BSAR_ERR_DEVICE;
    if (device < 0 || device >= count) return BSAR_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BSAR_ERR_DEVICE;
    bsar_ctx* c = new (std::nothrow) bsar_ctx();
    if (!c) return BSAR_ERR_NOMEM;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        bsar_ctx_destroy(c);
        return BSAR_ERR_DEVICE;
    }
    // device address of this GPU's no-allocation scratch (once per device);
    // a failed lookup only disables bsar_range_compress_device_noalloc
    (void)bsar::scratch_init(device);
    *out = c;
    return BSAR_OK;
}

void bsar_ctx_destroy(bsar_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : {(void*)c->dem_cache.dX, (void*)c->dem_cache.dY, (void*)c->dem_cache.dZ, c->in_buf, c->out_buf})
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* bsar_last_error(const bsar_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bsar_ctx_synchronize(bsar_ctx* c)
{
    if (!c) return BSAR_ERR_INVALID;
    BSAR_HIP(c, hipSetDevice(c->device));
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

double bsar_flops_per_voxel_pulse(int n) { return 65406.0 + 20.0 * (double)n; }

// The context's device copy of `dem`: reused when the arrays hold the same
// sizes and bytes as the last call's, uploaded (and the old copy freed)
// otherwise.
static int ctx_dem(bsar_ctx* c, const bsar_dem* dem, DemDev* out)
{
    auto& k = c->dem_cache;
    if (dem && k.valid && dem->nx == k.nx && dem->ny == k.ny && dem->zrows == k.zrows && dem->zcols == k.zcols) {
        const size_t nz = (size_t)std::max(dem->zrows, 0) * (size_t)std::max(dem->zcols, 0);
        const bool same = (!dem->nx || (dem->x && !std::memcmp(dem->x, k.x.data(), k.x.size() * 8))) &&
                          (!dem->ny || (dem->y && !std::memcmp(dem->y, k.y.data(), k.y.size() * 8))) &&
                          (!nz || (dem->z && !std::memcmp(dem->z, k.z.data(), nz * 8)));
        if (same) {
            *out = k.dem;
            return BSAR_OK;
        }
    }
    DevBuf bx, by, bz;
    DemDev d;
    double zmin, zmax;
    int st = upload_dem(c, dem, (double**)&bx.p, (double**)&by.p, (double**)&bz.p, &d, &zmin, &zmax);
    if (st) return st;
    for (double* p : {k.dX, k.dY, k.dZ})
        if (p) (void)hipFree(p);
    k.dX = (double*)bx.p;
    k.dY = (double*)by.p;
    k.dZ = (double*)bz.p;
    bx.p = by.p = bz.p = nullptr;
    const size_t nz = (size_t)dem->zrows * (size_t)dem->zcols;
    k.x.assign(dem->x, dem->x + (dem->x ? dem->nx : 0));
    k.y.assign(dem->y, dem->y + (dem->y ? dem->ny : 0));
    k.z.assign(dem->z, dem->z + (dem->z ? nz : 0));
    k.nx = dem->nx;
    k.ny = dem->ny;
    k.zrows = dem->zrows;
    k.zcols = dem->zcols;
    k.dem = d;
    k.valid = true;
    *out = d;
    return BSAR_OK;
}

// A context buffer of at least `need` bytes (grow-only).
static int ctx_buffer(bsar_ctx* c, void** p, size_t* cap, size_t need)
{
    if (need <= *cap) return BSAR_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    BSAR_HIP(c, hipMalloc(p, need));
    *cap = need;
    return BSAR_OK;
}

int bsar_dem_interp(bsar_ctx* c, const bsar_dem* dem, const double* xy, int64_t m, double* z)
{
    if (!c) return BSAR_ERR_INVALID;
    if (m < 0 || (m && (!xy || !z))) return set_error(c, BSAR_ERR_INVALID, "bad point buffer");
    BSAR_HIP(c, hipSetDevice(c->device));
    DemDev d;
    int st = ctx_dem(c, dem, &d);
    if (st) return st;
    if (m == 0) return BSAR_OK;
    if ((st = ctx_buffer(c, &c->in_buf, &c->in_cap, (size_t)m * 2 * sizeof(double)))) return st;
    if ((st = ctx_buffer(c, &c->out_buf, &c->out_cap, (size_t)m * sizeof(double)))) return st;
    BSAR_HIP(c, hipMemcpyAsync(c->in_buf, xy, (size_t)m * 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    BSAR_HIP(c, launch_dem_interp(d, (const double*)c->in_buf, m, (double*)c->out_buf, c->stream));
    BSAR_HIP(c, hipMemcpyAsync(z, c->out_buf, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

int bsar_refraction_points(bsar_ctx* c, const bsar_dem* dem, const double* legs, int64_t m, double n1, double n2,
                           double* out)
{
    if (!c) return BSAR_ERR_INVALID;
    if (m < 0 || (m && (!legs || !out))) return set_error(c, BSAR_ERR_INVALID, "bad leg buffer");
    BSAR_HIP(c, hipSetDevice(c->device));
    DemDev d;
    int st = ctx_dem(c, dem, &d);
    if (st) return st;
    if (m == 0) return BSAR_OK;
    if ((st = ctx_buffer(c, &c->in_buf, &c->in_cap, (size_t)m * 6 * sizeof(double)))) return st;
    if ((st = ctx_buffer(c, &c->out_buf, &c->out_cap, (size_t)m * 3 * sizeof(double)))) return st;
    BSAR_HIP(c, hipMemcpyAsync(c->in_buf, legs, (size_t)m * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    BSAR_HIP(c, launch_refraction(d, (const double*)c->in_buf, m, n1, n2, (double*)c->out_buf, c->stream));
    BSAR_HIP(c, hipMemcpyAsync(out, c->out_buf, (size_t)m * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

int bsar_refraction_points_wave(bsar_ctx* c, const bsar_dem* dem, const double* legs, int64_t m, double n1,
                                double n2, double* out, double* seconds)
{
    if (!c) return BSAR_ERR_INVALID;
    if (seconds) *seconds = 0.0;
    if (m < 0 || m >= (int64_t)1 << 33 || (m && (!legs || !out))) return set_error(c, BSAR_ERR_INVALID, "bad leg buffer");
    BSAR_HIP(c, hipSetDevice(c->device));
    DemDev d;
    int st = ctx_dem(c, dem, &d);
    if (st) return st;
    if (m == 0) return BSAR_OK;
    if ((st = ctx_buffer(c, &c->in_buf, &c->in_cap, (size_t)m * 6 * sizeof(double)))) return st;
    if ((st = ctx_buffer(c, &c->out_buf, &c->out_cap, (size_t)m * 3 * sizeof(double)))) return st;
    BSAR_HIP(c, hipMemcpyAsync(c->in_buf, legs, (size_t)m * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    BSAR_HIP(c, hipEventRecord(c->ev0, c->stream));
    BSAR_HIP(c, launch_refraction_wave(d, (const double*)c->in_buf, m, n1, n2, (double*)c->out_buf, c->stream));
    BSAR_HIP(c, hipEventRecord(c->ev1, c->stream));
    BSAR_HIP(c, hipMemcpyAsync(out, c->out_buf, (size_t)m * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    if (seconds) {
        float ms = 0.0f;
        BSAR_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *seconds = ms * 1e-3;
    }
    return BSAR_OK;
}

int bsar_slant_range_fermat(bsar_ctx* c, const bsar_dataset* ds, const double* pts, int64_t V, double* ida,
                            double* volta)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    if (V < 0 || (V && (!pts || !ida || !volta))) return set_error(c, BSAR_ERR_INVALID, "bad point buffer");
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (V == 0 || ds->n == 0) return BSAR_OK;
    const size_t m = (size_t)V * ds->n;
    DevBuf bp, ba, bb;
    BSAR_HIP(c, hipMalloc(&bp.p, (size_t)V * 3 * sizeof(double)));
    BSAR_HIP(c, hipMalloc(&ba.p, m * 3 * sizeof(double)));
    BSAR_HIP(c, hipMalloc(&bb.p, m * 3 * sizeof(double)));
    BSAR_HIP(c, hipMemcpyAsync(bp.p, pts, (size_t)V * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    BSAR_HIP(c, launch_slant_range(ds->dem, ds->traj, (const double*)bp.p, V, ds->p.n1, ds->p.n2, (double*)ba.p,
                                   (double*)bb.p, c->stream));
    BSAR_HIP(c, hipMemcpyAsync(ida, ba.p, m * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BSAR_HIP(c, hipMemcpyAsync(volta, bb.p, m * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

int bsar_slant_range_fermat_device(bsar_ctx* c, const bsar_dataset* ds, const double* d_pts, int64_t V,
                                   double* d_ida, double* d_volta, bsar_report* rep)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    if (V < 0 || (V && (!d_pts || !d_ida || !d_volta))) return set_error(c, BSAR_ERR_INVALID, "bad device buffer");
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (rep) std::memset(rep, 0, sizeof *rep);
    BSAR_HIP(c, hipEventRecord(c->ev0, c->stream));
    if (V > 0 && ds->n > 0)
        BSAR_HIP(c, launch_slant_range(ds->dem, ds->traj, d_pts, V, ds->p.n1, ds->p.n2, d_ida, d_volta, c->stream));
    BSAR_HIP(c, hipEventRecord(c->ev1, c->stream));
    BSAR_HIP(c, hipEventSynchronize(c->ev1));
    if (rep) {
        float ms = 0.f;
        BSAR_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        rep->seconds = ms * 1e-3;
        rep->voxel_pulses = (double)V * ds->n;
        rep->voxels = V;
        rep->pulses = ds->n;
        rep->launches = (V > 0 && ds->n > 0) ? 1 : 0;
    }
    return BSAR_OK;
}

int bsar_upsample(bsar_ctx* c, const double* in, int S, int N, int factor, double* out)
{
    if (!c) return BSAR_ERR_INVALID;
    if (S < 2 || N < 0 || factor < 1 || (N && (!in || !out)))
        return set_error(c, BSAR_ERR_INVALID, "upsample needs S >= 2, N >= 0, factor >= 1");
    if ((int64_t)S * factor >= (int64_t)1 << 30) return set_error(c, BSAR_ERR_INVALID, "S * factor too large");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (N == 0) return BSAR_OK;
    const size_t nin = (size_t)S * N, nout = nin * factor;
    DevBuf bi, bo;
    BSAR_HIP(c, hipMalloc(&bi.p, nin * sizeof(double2)));
    BSAR_HIP(c, hipMalloc(&bo.p, nout * sizeof(double2)));
    BSAR_HIP(c, hipMemcpyAsync(bi.p, in, nin * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    BSAR_HIP(c, launch_spline((const double2*)bi.p, S, N, factor, (double2*)bo.p, nullptr, c->stream));
    BSAR_HIP(c, hipMemcpyAsync(out, bo.p, nout * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

int bsar_chirp(double f0, double f1, double T, double fs, double* out)
{
    // createCHIRP, raw_data_processing.cpp:267-282
    const double L_d = T * fs;
    if (!(L_d >= 0.0) || L_d > 1e9) return BSAR_ERR_INVALID;
    const double k = (f1 - f0) / T;
    const int L = (int)(T * fs) + 1;
    if (out) {
        for (int n = 0; n < L; n++) {
            const double t = n / fs;
            const double phase = 2.0 * M_PI * (f0 * t + (k / 2.0) * t * t);
            out[2 * n] = std::cos(phase);
            out[2 * n + 1] = std::sin(phase);
        }
    }
    return L;
}

int bsar_dataset_create(bsar_ctx* c, const bsar_params* p, const bsar_dem* dem, const double* tx, const double* rx,
                        int n, bsar_dataset** out)
{
    if (!c) return BSAR_ERR_INVALID;
    if (!out || !p) return set_error(c, BSAR_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n < 0 || (n && (!tx || !rx))) return set_error(c, BSAR_ERR_INVALID, "bad trajectory buffers");
    if (!(p->c > 0) || !(p->fc > 0) || !(p->fs > 0) || p->upsample < 1)
        return set_error(c, BSAR_ERR_INVALID, "params need c, fc, fs > 0 and upsample >= 1");
    BSAR_HIP(c, hipSetDevice(c->device));
    bsar_dataset* ds = new (std::nothrow) bsar_dataset();
    if (!ds) return set_error(c, BSAR_ERR_NOMEM, "host allocation");
    ds->owner = c;
    ds->device = c->device;
    ds->p = *p;
    ds->n = n;
    int st = upload_dem(c, dem, &ds->dX, &ds->dY, &ds->dZ, &ds->dem, &ds->zmin, &ds->zmax);
    if (st) { bsar_dataset_destroy(ds); return st; }
    ds->htx.assign(tx, tx + 3 * (size_t)n);
    ds->hrx.assign(rx, rx + 3 * (size_t)n);
    // [tx ty tz rx ry rz] followed by the R1 sums' expansion terms (R1Geo):
    // Tx and Rx each as b = A - o per axis and c = |b|^2.  The origin o (the
    // first pulse's Tx/Rx midpoint) and the bounds kept per prefix of pulses
    // make an image over the first K pulses sum exactly the terms a dataset
    // holding only those K pulses sums.
    std::vector<double> soa(14 * (size_t)std::max(n, 1), 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            soa[(size_t)k * n + i] = tx[3 * i + k];
            soa[(size_t)(3 + k) * n + i] = rx[3 * i + k];
        }
    double o[3] = {0.0, 0.0, 0.0};
    if (n > 0)
        for (int k = 0; k < 3; ++k) o[k] = (tx[k] + rx[k]) / 2.0;
    try {
        ds->r1_bsq.resize(n);
        ds->r1_zlo.resize(n);
        ds->r1_zhi.resize(n);
    } catch (const std::bad_alloc&) {
        bsar_dataset_destroy(ds);
        return set_error(c, BSAR_ERR_NOMEM, "host allocation");
    }
    double bsq = 0.0, zlo = INFINITY, zhi = -INFINITY;
    for (int i = 0; i < n; ++i) {
        for (int set = 0; set < 2; ++set) {
            const double* A = set ? rx : tx;
            double* B = soa.data() + (size_t)(6 + 4 * set) * n;
            double cc = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double b = A[3 * i + k] - o[k];
                B[(size_t)k * n + i] = b;
                cc += b * b;
            }
            B[(size_t)3 * n + i] = cc;
            bsq = std::max(bsq, cc);   // a NaN term leaves bsq; then Q is NaN too and the sums go direct
            zlo = std::min(zlo, A[3 * i + 2]);
            zhi = std::max(zhi, A[3 * i + 2]);
        }
        ds->r1_bsq[i] = bsq * (1.0 + 1e-12);   // above every stored c despite their roundings
        ds->r1_zlo[i] = zlo;
        ds->r1_zhi[i] = zhi;
    }
    hipError_t e = hipMalloc((void**)&ds->dtraj, soa.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(ds->dtraj, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) { bsar_dataset_destroy(ds); return hip_fail(c, e, "trajectory upload"); }
    ds->traj.tx = ds->dtraj;
    ds->traj.ty = ds->dtraj + n;
    ds->traj.tz = ds->dtraj + 2 * (size_t)n;
    ds->traj.rx = ds->dtraj + 3 * (size_t)n;
    ds->traj.ry = ds->dtraj + 4 * (size_t)n;
    ds->traj.rz = ds->dtraj + 5 * (size_t)n;
    ds->traj.n = n;
    ds->r1.t = ds->dtraj + 6 * (size_t)n;
    ds->r1.r = ds->dtraj + 10 * (size_t)n;
    ds->r1.stride = n;
    ds->r1.ox = o[0];
    ds->r1.oy = o[1];
    ds->r1.oz = o[2];
    ds->bytes = (int64_t)soa.size() * 8 + (int64_t)(ds->dem.nx + ds->dem.ny + ds->dem.ny * ds->dem.zs) * 8 +
                (int64_t)ds->dem.ny * ds->dem.zs * 4;
    *out = ds;
    return BSAR_OK;
}

void bsar_dataset_destroy(bsar_dataset* ds)
{
    if (!ds) return;
    // The dataset may outlive the context that created it: wait on the device,
    // never on the creator's stream.
    (void)hipSetDevice(ds->device);
    (void)hipDeviceSynchronize();
    bsar_r1tab_free(ds);
    if (ds->dX) (void)hipFree(ds->dX);
    if (ds->dY) (void)hipFree(ds->dY);
    if (ds->dZ) (void)hipFree(ds->dZ);
    if (ds->dtraj) (void)hipFree(ds->dtraj);
    if (ds->rc) (void)hipFree(ds->rc);
    if (ds->rc32) (void)hipFree(ds->rc32);
    delete ds;
}

int64_t bsar_dataset_device_bytes(const bsar_dataset* ds) { return ds ? ds->bytes : 0; }

__global__ void k_to_f32(const double2* __restrict__ a, float2* __restrict__ b, int64_t n)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) b[k] = make_float2((float)a[k].x, (float)a[k].y);
}

// FP32 copy of the upsampled samples for the mixed form (when the spline did
// not write it in the same pass), then the dataset's byte count.
static int finish_rcomp(bsar_ctx* c, bsar_dataset* ds)
{
    const int64_t n = (int64_t)ds->ndata * ds->S4;
    if (!ds->rc32) {
        BSAR_HIP(c, hipMalloc((void**)&ds->rc32, (size_t)n * sizeof(float2)));
        hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, ds->rc, ds->rc32, n);
        BSAR_HIP(c, hipGetLastError());
    }
    BSAR_HIP(c, hipStreamSynchronize(c->stream));
    ds->bytes += n * (int64_t)(sizeof(double2) + sizeof(float2));
    return BSAR_OK;
}

static void drop_rcomp(bsar_dataset* ds)
{
    if (ds->rc) { (void)hipFree(ds->rc); ds->rc = nullptr; }
    if (ds->rc32) { (void)hipFree(ds->rc32); ds->rc32 = nullptr; }
    ds->bytes -= (int64_t)ds->ndata * ds->S4 * (int64_t)(sizeof(double2) + sizeof(float2));
    ds->ndata = 0;
    ds->S4 = 0;
}

int bsar_dataset_load_rootdata(bsar_ctx* c, bsar_dataset* ds, const double* root, int S, int n_data)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    if (S < 2 || n_data < 1 || !root) return set_error(c, BSAR_ERR_INVALID, "rootData needs S >= 2, n_data >= 1");
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    const int64_t S4 = (int64_t)S * ds->p.upsample;
    if (S4 > std::numeric_limits<int>::max() / 2) return set_error(c, BSAR_ERR_INVALID, "S * upsample too large");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (ds->rc) drop_rcomp(ds);
    const size_t nin = (size_t)S * n_data;
    DevBuf bi;
    BSAR_HIP(c, hipMalloc(&bi.p, nin * sizeof(double2)));
    BSAR_HIP(c, hipMemcpyAsync(bi.p, root, nin * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    BSAR_HIP(c, hipMalloc((void**)&ds->rc, nin * ds->p.upsample * sizeof(double2)));
    BSAR_HIP(c, hipMalloc((void**)&ds->rc32, nin * ds->p.upsample * sizeof(float2)));
    BSAR_HIP(c, launch_spline((const double2*)bi.p, S, n_data, ds->p.upsample, ds->rc, ds->rc32, c->stream));
    ds->ndata = n_data;
    ds->S4 = (int)S4;
    return finish_rcomp(c, ds);
}

int bsar_dataset_load_rootdata_device(bsar_ctx* c, bsar_dataset* ds, const void* d_root, int S, int n_data,
                                      int precision)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    if (S < 2 || n_data < 1 || !d_root) return set_error(c, BSAR_ERR_INVALID, "rootData needs S >= 2, n_data >= 1");
    if (precision != BSAR_FP64 && precision != BSAR_MIXED) return set_error(c, BSAR_ERR_INVALID, "bad precision");
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    const int64_t S4 = (int64_t)S * ds->p.upsample;
    if (S4 > std::numeric_limits<int>::max() / 2) return set_error(c, BSAR_ERR_INVALID, "S * upsample too large");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (ds->rc) drop_rcomp(ds);
    const size_t nin = (size_t)S * n_data;
    BSAR_HIP(c, hipMalloc((void**)&ds->rc, nin * ds->p.upsample * sizeof(double2)));
    BSAR_HIP(c, hipMalloc((void**)&ds->rc32, nin * ds->p.upsample * sizeof(float2)));
    // complex64 samples (BSAR_MIXED) are widened exactly inside the spline
    if (precision == BSAR_MIXED)
        BSAR_HIP(c, launch_spline((const float2*)d_root, S, n_data, ds->p.upsample, ds->rc, ds->rc32, c->stream));
    else
        BSAR_HIP(c, launch_spline((const double2*)d_root, S, n_data, ds->p.upsample, ds->rc, ds->rc32, c->stream));
    ds->ndata = n_data;
    ds->S4 = (int)S4;
    return finish_rcomp(c, ds);
}

int bsar_dataset_load_upsampled(bsar_ctx* c, bsar_dataset* ds, const double* rcomp, int S4, int n_data)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    if (S4 < 1 || n_data < 1 || !rcomp) return set_error(c, BSAR_ERR_INVALID, "bad upsampled block");
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (ds->rc) drop_rcomp(ds);
    const size_t n = (size_t)S4 * n_data;
    BSAR_HIP(c, hipMalloc((void**)&ds->rc, n * sizeof(double2)));
    BSAR_HIP(c, hipMemcpyAsync(ds->rc, rcomp, n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    ds->ndata = n_data;
    ds->S4 = S4;
    return finish_rcomp(c, ds);
}

int bsar_backproject_device(bsar_ctx* c, const bsar_dataset* ds, const double* gx, int nx, const double* gy, int ny,
                            const double* gz, int nz, int precision, void* d_out, bsar_report* rep)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    int st = check_axes(c, gx, nx, gy, ny, gz, nz);
    if (st) return st;
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    if (!rcomp_ready(ds)) return set_error(c, BSAR_ERR_STATE, "no range-compressed data loaded");
    if (!d_out && (int64_t)nx * ny * nz > 0) return set_error(c, BSAR_ERR_INVALID, "d_out is null");
    BSAR_HIP(c, hipSetDevice(c->device));
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (precision == BSAR_FP64) return backproject_fp64(c, ds, gx, nx, gy, ny, gz, nz, (double2*)d_out, rep);
    if (precision == BSAR_MIXED) return backproject_mixed(c, ds, gx, nx, gy, ny, gz, nz, (double2*)d_out, rep);
    return set_error(c, BSAR_ERR_INVALID, "unknown precision");
}

int bsar_backproject(bsar_ctx* c, const bsar_dataset* ds, const double* gx, int nx, const double* gy, int ny,
                     const double* gz, int nz, int precision, double* out, bsar_report* rep)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    const int64_t V = (int64_t)nx * ny * nz;
    if (V > 0 && !out) return set_error(c, BSAR_ERR_INVALID, "out is null");
    BSAR_HIP(c, hipSetDevice(c->device));
    DevBuf img;
    if (V > 0) BSAR_HIP(c, hipMalloc(&img.p, (size_t)V * sizeof(double2)));
    int st = bsar_backproject_device(c, ds, gx, nx, gy, ny, gz, nz, precision, img.p, rep);
    if (st) return st;
    if (V > 0) {
        BSAR_HIP(c, hipMemcpyAsync(out, img.p, (size_t)V * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        BSAR_HIP(c, hipStreamSynchronize(c->stream));
    }
    return BSAR_OK;
}

}  // extern "C"

// scenario.cpp -- the reference's scenario generators and DEM loader as C ABI
// calls (host code, no device work):
//
//   bsar_spiral      funcao_espiral (funcao_espiral.cpp:5-49) and the
//                    Tx/Rx wrappers (trajectory_generator.cpp:4-33)
//   bsar_grid_axis   one axis of crearGrid (crear_grid.cpp:6-73)
//   bsar_dem_load    loadDEM of the build without GDAL
//                    (load_dem_simple.cpp:9-224), or a GeoTIFF decoded to
//                    local metres (geotiff.cpp)
//
// The generators keep the reference's operation order (time and axis values
// are running sums, not i * step) and its math library calls so that the
// positions are the same bits.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_files.h"

namespace bsar {

thread_local std::string g_file_error;

int file_fail(int status, const std::string& msg)
{
    g_file_error = msg;
    return status;
}

int dem_grid_fill(bsar_dem_grid* out, const std::vector<double>& x, const std::vector<double>& y,
                  const std::vector<double>& z, int zrows, int zcols)
{
    auto dup = [](const std::vector<double>& v) -> double* {
        double* p = (double*)std::malloc((v.empty() ? 1 : v.size()) * sizeof(double));
        if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(double));
        return p;
    };
    double* px = dup(x);
    double* py = dup(y);
    double* pz = dup(z);
    if (!px || !py || !pz) {
        std::free(px);
        std::free(py);
        std::free(pz);
        return file_fail(BSAR_ERR_NOMEM, "host allocation failed");
    }
    out->dem = bsar_dem{px, (int)x.size(), py, (int)y.size(), pz, zrows, zcols};
    return BSAR_OK;
}

namespace {

// The DEM loadDEM returns for a .tif name or a file it cannot open
// (load_dem_simple.cpp:16-72 and :78-134): 32 x 32 nodes from -500 m to
// +500 m on both axes, all heights zero (every terrain term there carries a
// 0.0 factor).
void default_dem(std::vector<double>& x, std::vector<double>& y, std::vector<double>& z, int& rows, int& cols)
{
    const double lim = 500.0;
    const int M = 32;
    const double cell = (lim - -lim) / (M - 1);
    x.resize(M);
    y.resize(M);
    for (int i = 0; i < M; ++i) x[i] = -lim + i * cell;
    for (int i = 0; i < M; ++i) y[i] = -lim + i * cell;
    z.assign((size_t)M * M, 0.0);
    rows = cols = M;
}

// ESRI ASCII grid as load_dem_simple.cpp:136-214 reads it: header lines
// "key value" until the first line whose first token is not a header key;
// that line and the following ones are rows (values that do not parse read
// as NODATA_value); X_vec[i] = xllcorner + i cellsize, Y_vec[j] = yllcorner
// + j cellsize with row j of the file at Y_vec[j].  Tokens are parsed with
// the same stream extraction the reference uses.
int read_ascii_grid(std::ifstream& file, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z,
                    int& rows_out, int& cols_out)
{
    std::string line;
    bool header_done = false;
    int row = 0, ncols = 0, nrows = 0;
    double xll = 0.0, yll = 0.0, cell = 1.0, nodata = -9999.0;
    auto parse_row = [&](std::istringstream& s) {
        for (int c = 0; c < ncols; ++c) {
            double v;
            z[(size_t)row * ncols + c] = (s >> v) ? v : nodata;
        }
        ++row;
    };
    while (std::getline(file, line)) {
        std::istringstream s(line);
        if (!header_done) {
            std::string key;
            s >> key;
            if (key == "ncols") s >> ncols;
            else if (key == "nrows") s >> nrows;
            else if (key == "xllcorner") s >> xll;
            else if (key == "yllcorner") s >> yll;
            else if (key == "cellsize") s >> cell;
            else if (key == "NODATA_value" || key == "nodata_value") s >> nodata;
            else {
                header_done = true;
                if (ncols < 0 || nrows < 0 || (int64_t)ncols * nrows > ((int64_t)1 << 31))
                    return file_fail(BSAR_ERR_FORMAT, "ESRI grid: bad ncols/nrows");
                z.assign((size_t)nrows * ncols, 0.0);
                x.resize(ncols);
                y.resize(nrows);
                for (int i = 0; i < ncols; ++i) x[i] = xll + i * cell;
                for (int i = 0; i < nrows; ++i) y[i] = yll + i * cell;
                std::istringstream d(line);
                if (row < nrows) parse_row(d);
                else ++row;
            }
        } else if (row < nrows) {
            parse_row(s);
        }
    }
    rows_out = nrows;
    cols_out = ncols;
    return BSAR_OK;
}

}  // namespace
}  // namespace bsar

using namespace bsar;

extern "C" {

BSAR_API const char* bsar_file_error(void) { return g_file_error.c_str(); }

BSAR_API int bsar_spiral(double turns, double r_minor, double r_major, double h_high, double h_low, double vt,
                         double prf, int mirror_xy, double* px, double* py, double* pz, int64_t cap,
                         int64_t* count)
{
    if (!count) return file_fail(BSAR_ERR_INVALID, "count is NULL");
    // the reference starts at the major radius and ends at the minor one
    const double r_start = r_major;
    const double r_end = r_minor;
    const double vr = vt * (std::log(r_start) - std::log(r_end)) / (turns * 2 * M_PI);
    const double T = (r_start - r_end) / vr;
    const double vd = (h_high - h_low) / T;
    const double dt = 1.0 / prf;
    // times 0, dt, 2 dt, ... as a running sum while <= T
    if (!(dt > 0.0) || !std::isfinite(dt) || (std::isinf(T) && T > 0))
        return file_fail(BSAR_ERR_INVALID, "spiral: the time loop does not end (PRF <= 0 or radial rate 0)");
    const int64_t kMax = (int64_t)1 << 31;
    int64_t n = 0;
    for (double t = 0.0; t <= T; t += dt) {
        if (n < cap && px && py && pz) {
            const double theta = (vt / vr) * (std::log(r_start) - std::log(r_start - vr * t));
            const double R = r_start - vr * t;
            // the reference's optimised build evaluates cos/sin of one angle
            // as one glibc sincos call, which differs from separate cos and
            // sin calls in the last bit for some angles
            double sn, cs;
            sincos(theta, &sn, &cs);
            const double x = R * cs;
            const double y = R * sn;
            px[n] = mirror_xy ? -x : x;
            py[n] = mirror_xy ? -y : y;
            pz[n] = h_high - vd * t;
        }
        if (++n > kMax) return file_fail(BSAR_ERR_INVALID, "spiral: more than 2^31 positions");
    }
    *count = n;
    if ((px || py || pz) && n > cap) return file_fail(BSAR_ERR_INVALID, "spiral: output capacity too small");
    return BSAR_OK;
}

BSAR_API int bsar_grid_axis(double center, double step, double half, double* out, int cap, int* count)
{
    if (!count) return file_fail(BSAR_ERR_INVALID, "count is NULL");
    if (!(step > 0.0) || !std::isfinite(step) || !std::isfinite(center) || !std::isfinite(half))
        return file_fail(BSAR_ERR_INVALID, "grid axis: step must be finite and > 0");
    int64_t n = 0;
    for (double v = center - half; v <= center + half + step / 2; v += step) {
        if (out && n < cap) out[n] = v;
        if (++n > ((int64_t)1 << 30)) return file_fail(BSAR_ERR_INVALID, "grid axis: more than 2^30 values");
    }
    *count = (int)n;
    if (out && n > cap) return file_fail(BSAR_ERR_INVALID, "grid axis: output capacity too small");
    return BSAR_OK;
}

static int dem_load_impl(const char* path, int mode, bsar_dem_grid* out);

BSAR_API int bsar_dem_load(const char* path, int mode, bsar_dem_grid* out)
{
    try {
        return dem_load_impl(path, mode, out);
    } catch (const std::bad_alloc&) {
        return file_fail(BSAR_ERR_NOMEM, "DEM file: host allocation failed");
    } catch (const std::exception& e) {
        return file_fail(BSAR_ERR_FORMAT, std::string("DEM file: ") + e.what());
    }
}

static int dem_load_impl(const char* path, int mode, bsar_dem_grid* out)
{
    if (!path || !out) return file_fail(BSAR_ERR_INVALID, "path or out is NULL");
    std::memset(out, 0, sizeof(*out));
    out->lon0 = out->lat0 = NAN;
    if (mode == BSAR_DEM_FILE_GEOTIFF) return geotiff_load(path, out);
    if (mode != BSAR_DEM_FILE_REFERENCE) return file_fail(BSAR_ERR_INVALID, "unknown DEM load mode");
    std::vector<double> x, y, z;
    int rows = 0, cols = 0;
    const std::string name(path);
    std::ifstream file;
    if (name.find(".tif") != std::string::npos || name.find(".TIF") != std::string::npos) {
        default_dem(x, y, z, rows, cols);
        out->source = BSAR_DEM_SOURCE_DEFAULT;
    } else {
        file.open(name);
        if (!file.is_open()) {
            default_dem(x, y, z, rows, cols);
            out->source = BSAR_DEM_SOURCE_DEFAULT;
        } else {
            int st = read_ascii_grid(file, x, y, z, rows, cols);
            if (st != BSAR_OK) return st;
            out->source = BSAR_DEM_SOURCE_ASCII_GRID;
        }
    }
    return dem_grid_fill(out, x, y, z, rows, cols);
}

BSAR_API void bsar_dem_grid_free(bsar_dem_grid* g)
{
    if (!g) return;
    std::free((void*)g->dem.x);
    std::free((void*)g->dem.y);
    std::free((void*)g->dem.z);
    g->dem = bsar_dem{};
}

}  // extern "C"

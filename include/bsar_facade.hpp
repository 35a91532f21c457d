// bsar_facade.hpp -- C++ facade with the reference's function names and
// signatures over the C ABI (bsar.h), so existing call sites switch by
// including this header and linking libbsar.so instead of compiling
// raw_data_processing.cpp / interpolation.cpp / fft_utils.cpp.
//
// The functions are templates over the caller's own structs: any type with
// the reference's member names works unchanged --
//   DEM  : X_vec, Y_vec (std::vector<double>), Z_DEM (vector<vector<double>>)   radar_params.h:75-79
//   TRAJ : Tx, Rx (vector<vector<double>> [N][3]), P (vector<double>(3))        raw_data_processing.h:11-15
//   ENV  : n1, n2                                                               radar_params.h:68-71
//   REFR : P_refrac_ida, P_refrac_volta (vector<vector<double>>)                raw_data_processing.h:25-28
//   REFL : P_reflection (vector<vector<double>>)                                raw_data_processing.h:31-34
// Errors are raised as std::runtime_error (the reference's I/O functions
// throw the same type, binary_reader.cpp:12-68).
//
// Each host thread gets its own context on device 0 (BSAR_DEVICE overrides),
// matching the reference's reentrant use from OpenMP threads
// (proc_test_script.cpp:451).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "bsar.h"

namespace bsar {
namespace facade {

inline void check(int st, bsar_ctx* ctx, const char* what)
{
    if (st != BSAR_OK) {
        std::string m = std::string(what) + ": " + bsar_status_string(st);
        if (ctx) m += std::string(" (") + bsar_last_error(ctx) + ")";
        throw std::runtime_error(m);
    }
}

// Per-thread context.
inline bsar_ctx* context()
{
    struct Holder {
        bsar_ctx* c = nullptr;
        ~Holder() { if (c) bsar_ctx_destroy(c); }
    };
    thread_local Holder h;
    if (!h.c) {
        const char* dev = std::getenv("BSAR_DEVICE");
        check(bsar_ctx_create(dev ? std::atoi(dev) : 0, &h.c), nullptr, "bsar_ctx_create");
    }
    return h.c;
}

// Flattened DEM (row-major Z, short rows padded with 0.0 as the reference's
// bounds checks read them, raw_data_processing.cpp:649-652).
struct FlatDem {
    std::vector<double> x, y, z;
    bsar_dem d{};
    template <class DEM>
    explicit FlatDem(const DEM& dem) : x(dem.X_vec), y(dem.Y_vec)
    {
        size_t cols = 0;
        for (const auto& r : dem.Z_DEM) cols = r.size() > cols ? r.size() : cols;
        z.assign(dem.Z_DEM.size() * cols, 0.0);
        for (size_t j = 0; j < dem.Z_DEM.size(); ++j)
            for (size_t i = 0; i < dem.Z_DEM[j].size(); ++i) z[j * cols + i] = dem.Z_DEM[j][i];
        d = bsar_dem{x.data(), (int)x.size(), y.data(), (int)y.size(), z.data(), (int)dem.Z_DEM.size(), (int)cols};
    }
};

// interpolateDEMElevation, raw_data_processing.cpp:615-680
template <class DEM>
double interpolateDEMElevation(double x, double y, const DEM& dem)
{
    FlatDem f(dem);
    const double xy[2] = {x, y};
    double z = 0.0;
    bsar_ctx* c = context();
    check(bsar_dem_interp(c, &f.d, xy, 1, &z), c, "interpolateDEMElevation");
    return z;
}

// calculateRefractionPoint, raw_data_processing.cpp:455-572
template <class DEM>
std::vector<double> calculateRefractionPoint(double x1, double y1, double z1, double x2, double y2, double z2,
                                             const DEM& dem, double n1, double n2, bool /*bPlotVerbose*/ = false)
{
    FlatDem f(dem);
    const double leg[6] = {x1, y1, z1, x2, y2, z2};
    std::vector<double> out(3);
    bsar_ctx* c = context();
    check(bsar_refraction_points(c, &f.d, leg, 1, n1, n2, out.data()), c, "calculateRefractionPoint");
    return out;
}

// calculateReflectionPoint, raw_data_processing.cpp:575-612 (host: the point
// is dead for the image): the Tx/Rx horizontal midpoint at the height of the
// DEM node nearest to it, nodes scanned x-major with the first minimum kept;
// nodes outside a short Z row are skipped, and no node leaves height 0.
template <class DEM>
std::vector<double> calculateReflectionPoint(double txX, double txY, double /*txZ*/, double rxX, double rxY,
                                             double /*rxZ*/, const DEM& dem)
{
    const double mx = (txX + rxX) / 2.0, my = (txY + rxY) / 2.0;
    double best = 1e9, bz = 0.0;
    if (!dem.X_vec.empty() && !dem.Y_vec.empty() && !dem.Z_DEM.empty())
        for (size_t a = 0; a < dem.X_vec.size(); ++a)
            for (size_t b = 0; b < dem.Y_vec.size() && b < dem.Z_DEM.size(); ++b) {
                if (a >= dem.Z_DEM[b].size()) continue;
                const double dx = dem.X_vec[a] - mx, dy = dem.Y_vec[b] - my;
                const double dd = std::sqrt(dx * dx + dy * dy);
                if (dd < best) { best = dd; bz = dem.Z_DEM[b][a]; }
            }
    return {mx, my, bz};
}

// RangeData, raw_data_processing.h:18-23.  Converts to any caller type with
// the same four members (the reference's own RangeData), so
// `RangeData ranges = calculaSlantRangeFermat(...)` (proc_test_script.cpp:
// 484-485) compiles with the caller's struct.
struct RangeData {
    std::vector<double> r1Tx, r2Tx, r1Rx, r2Rx;

    template <class T, class = decltype(std::declval<T&>().r1Tx = std::vector<double>(),
                                        std::declval<T&>().r2Tx = std::vector<double>(),
                                        std::declval<T&>().r1Rx = std::vector<double>(),
                                        std::declval<T&>().r2Rx = std::vector<double>())>
    operator T() const
    {
        T t{};
        t.r1Tx = r1Tx;
        t.r2Tx = r2Tx;
        t.r1Rx = r1Rx;
        t.r2Rx = r2Rx;
        return t;
    }
};

// calculaSlantRangeFermat, raw_data_processing.cpp:112-255: refraction points
// for every pulse of trajectories.Tx / Rx and the target trajectories.P, plus
// the four ranges.  The reflection point (dead for the image, :188-193) is
// the horizontal Tx/Rx midpoint at the height of the nearest DEM node.
template <class DEM, class TRAJ, class ENV, class REFL, class REFR>
RangeData calculaSlantRangeFermat(const DEM& dem, const TRAJ& traj, const ENV& env, REFL& refl, REFR& refr,
                                  bool /*bPlotVerbose*/ = false)
{
    const size_t n = traj.Tx.size();
    std::vector<double> legs(12 * n), out(6 * n);
    for (size_t i = 0; i < n; ++i) {
        const double* P = traj.P.data();
        double* a = &legs[6 * i];
        double* b = &legs[6 * (n + i)];
        for (int k = 0; k < 3; ++k) {
            a[k] = traj.Tx[i][k];
            a[3 + k] = P[k];
            b[k] = P[k];
            b[3 + k] = traj.Rx[i][k];
        }
    }
    FlatDem f(dem);
    bsar_ctx* c = context();
    RangeData r;
    r.r1Tx.resize(n); r.r2Tx.resize(n); r.r1Rx.resize(n); r.r2Rx.resize(n);
    refr.P_refrac_ida.assign(n, std::vector<double>(3));
    refr.P_refrac_volta.assign(n, std::vector<double>(3));
    refl.P_reflection.assign(n, std::vector<double>(3));
    if (n == 0) return r;
    // outbound legs use (n1, n2), return legs (n2, n1) (raw_data_processing.cpp:166-181)
    check(bsar_refraction_points(c, &f.d, legs.data(), (int64_t)n, env.n1, env.n2, out.data()), c,
          "calculaSlantRangeFermat");
    check(bsar_refraction_points(c, &f.d, legs.data() + 6 * n, (int64_t)n, env.n2, env.n1, out.data() + 3 * n), c,
          "calculaSlantRangeFermat");
    auto dist = [](const double* a, const double* b) {
        const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    for (size_t i = 0; i < n; ++i) {
        const double* qa = &out[3 * i];
        const double* qb = &out[3 * (n + i)];
        for (int k = 0; k < 3; ++k) { refr.P_refrac_ida[i][k] = qa[k]; refr.P_refrac_volta[i][k] = qb[k]; }
        const double tx[3] = {traj.Tx[i][0], traj.Tx[i][1], traj.Tx[i][2]};
        const double rx[3] = {traj.Rx[i][0], traj.Rx[i][1], traj.Rx[i][2]};
        r.r1Tx[i] = dist(tx, qa);
        r.r2Tx[i] = dist(traj.P.data(), qa);
        r.r1Rx[i] = dist(rx, qb);
        r.r2Rx[i] = dist(traj.P.data(), qb);
        refl.P_reflection[i] = calculateReflectionPoint(tx[0], tx[1], tx[2], rx[0], rx[1], rx[2], dem);
    }
    return r;
}

// splineInterpolation, interpolation.cpp:4-51
inline std::vector<std::complex<double>> splineInterpolation(const std::vector<std::complex<double>>& data, int factor)
{
    std::vector<std::complex<double>> out(data.size() * (size_t)factor);
    bsar_ctx* c = context();
    check(bsar_upsample(c, reinterpret_cast<const double*>(data.data()), (int)data.size(), 1, factor,
                        reinterpret_cast<double*>(out.data())), c, "splineInterpolation");
    return out;
}

// FFTProcessor, fft_utils.h:13-84 / fft_utils.cpp:11-195: the transforms run
// on the GPU (bsar_fft, bsar_range_compress); the helpers are host code with
// the reference's rules.
struct FFTProcessor {
    // fft / ifft, fft_utils.cpp:11-60: in place; a length that is not a power
    // of two leaves the forward transform a no-op with the reference's warning
    // on stderr (:17-20), and the inverse a division by the length.
    static void fft(std::vector<std::complex<double>>& input) { transform(input, 0); }
    static void ifft(std::vector<std::complex<double>>& input) { transform(input, 1); }

    static std::vector<std::complex<double>> fftCorrelation(const std::vector<std::complex<double>>& signal,
                                                            const std::vector<std::complex<double>>& reference)
    {
        return run(signal, reference, 1);
    }
    static std::vector<std::complex<double>> fftConvolution(const std::vector<std::complex<double>>& signal,
                                                            const std::vector<std::complex<double>>& reference)
    {
        return run(signal, reference, 0);
    }

    // multiply, fft_utils.cpp:140-150: element-wise over the shorter length
    static std::vector<std::complex<double>> multiply(const std::vector<std::complex<double>>& a,
                                                      const std::vector<std::complex<double>>& b)
    {
        const size_t n = a.size() < b.size() ? a.size() : b.size();
        std::vector<std::complex<double>> r(n);
        for (size_t i = 0; i < n; ++i) r[i] = a[i] * b[i];
        return r;
    }

    // padToPowerOfTwo, fft_utils.cpp:154-161: zero-padded to nextPowerOfTwo
    static std::vector<std::complex<double>> padToPowerOfTwo(const std::vector<std::complex<double>>& input)
    {
        std::vector<std::complex<double>> p(nextPowerOfTwo(input.size()), 0.0);
        for (size_t i = 0; i < input.size(); ++i) p[i] = input[i];
        return p;
    }

    // bitReverse, fft_utils.cpp:164-181: the in-place bit-reversal permutation
    // (j counts in bit-reversed order beside i, carrying from the top bit of
    // length / 2 down; pairs with i < j swap), which a length that is not a
    // power of two runs through the same way
    static void bitReverse(std::vector<std::complex<double>>& input)
    {
        const size_t n = input.size();
        size_t j = 0;
        for (size_t i = 1; i < n; ++i) {
            size_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(input[i], input[j]);
        }
    }

    // isPowerOfTwo / nextPowerOfTwo, fft_utils.cpp:183-195 (nextPowerOfTwo(0) = 1)
    static bool isPowerOfTwo(size_t n) { return n > 0 && (n & (n - 1)) == 0; }
    static size_t nextPowerOfTwo(size_t n)
    {
        size_t p = 1;
        while (p < n) p *= 2;
        return p;
    }

private:
    static void transform(std::vector<std::complex<double>>& x, int inverse)
    {
        const size_t n = x.size();
        if (n > 1 && !isPowerOfTwo(n))
            std::fprintf(stderr, "Warning: FFT input size %zu is not a power of 2. Consider padding.\n", n);
        if (n == 0) return;
        if (n > (size_t)1 << 26) throw std::runtime_error("fft: length above 2^26");
        bsar_ctx* c = context();
        check(bsar_fft(c, reinterpret_cast<double*>(x.data()), (int)n, 1, inverse), c, inverse ? "ifft" : "fft");
    }

    static std::vector<std::complex<double>> run(std::vector<std::complex<double>> s,
                                                 std::vector<std::complex<double>> r, int conj)
    {
        const size_t N = s.size() > r.size() ? s.size() : r.size();
        s.resize(N);
        r.resize(N);
        std::vector<std::complex<double>> out(N);
        if (N == 0) return out;
        bsar_ctx* c = context();
        check(bsar_range_compress(c, reinterpret_cast<const double*>(s.data()), (int)N, 1,
                                  reinterpret_cast<const double*>(r.data()), conj, BSAR_FP64,
                                  reinterpret_cast<double*>(out.data())), c, "fftCorrelation");
        return out;
    }
};

// createCHIRP, raw_data_processing.cpp:257-290 (RADAR: FreqMenor, FreqMaior,
// DuracaoPulso, fs as in radar_params.h:7-27)
template <class RADAR>
std::vector<std::complex<double>> createCHIRP(const RADAR& tx)
{
    const int L = bsar_chirp(tx.FreqMenor, tx.FreqMaior, tx.DuracaoPulso, tx.fs, nullptr);
    check(L < 0 ? L : BSAR_OK, nullptr, "createCHIRP");
    std::vector<std::complex<double>> out((size_t)L);
    bsar_chirp(tx.FreqMenor, tx.FreqMaior, tx.DuracaoPulso, tx.fs, reinterpret_cast<double*>(out.data()));
    return out;
}

// MatrixInfo + readBinary<double>, binary_reader.h:10-26 / binary_reader.cpp:9-99
struct MatrixInfo {
    std::vector<uint64_t> dimensions;
    bool isComplex = false;
    std::string dataType;
    size_t totalElements = 0;
};

template <typename T>
std::vector<std::complex<T>> readBinary(const std::string& filename, MatrixInfo& info)
{
    static_assert(std::is_same<T, double>::value, "the stage files hold float64");
    uint64_t dims[64];
    int nd = 0, cplx = 0;
    uint64_t count = 0;
    int st = bsar_read_matrix(filename.c_str(), dims, 64, &nd, &cplx, nullptr, &count);
    if (st != BSAR_OK) throw std::runtime_error("readBinary: " + filename + ": " + bsar_status_string(st));
    std::vector<std::complex<T>> out(count);
    st = bsar_read_matrix(filename.c_str(), dims, 64, &nd, &cplx, reinterpret_cast<double*>(out.data()), &count);
    if (st != BSAR_OK) throw std::runtime_error("readBinary: " + filename + ": " + bsar_status_string(st));
    info.dimensions.assign(dims, dims + (nd < 64 ? nd : 64));
    info.isComplex = cplx != 0;
    info.dataType = "double";
    info.totalElements = count;
    return out;
}

// readCSV, csv_reader.cpp:6-26 (std::stod on every non-empty comma token)
inline std::vector<double> readCSV(const std::string& filename)
{
    std::ifstream f(filename);
    if (!f.is_open()) throw std::runtime_error("readCSV: cannot open " + filename);
    std::vector<double> v;
    std::string line, tok;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        while (std::getline(ss, tok, ','))
            if (!tok.empty()) v.push_back(std::stod(tok));
    }
    return v;
}

// readCSVMatrix, matrix_reader.cpp:6-40 (unparseable tokens skipped, empty rows dropped)
inline std::vector<std::vector<double>> readCSVMatrix(const std::string& filename)
{
    std::ifstream f(filename);
    if (!f.is_open()) throw std::runtime_error("readCSVMatrix: cannot open " + filename);
    std::vector<std::vector<double>> m;
    std::string line, tok;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::vector<double> row;
        std::stringstream ss(line);
        while (std::getline(ss, tok, ',')) {
            if (tok.empty()) continue;
            try { row.push_back(std::stod(tok)); } catch (const std::exception&) {}
        }
        if (!row.empty()) m.push_back(std::move(row));
    }
    return m;
}

// export3DComplexMatrix, data_export.cpp:180-309 (output[x][y][z] in, x-fastest file out)
inline void export3DComplexMatrix(const std::string& filename,
                                  const std::vector<std::vector<std::vector<std::complex<double>>>>& out3d, size_t nx,
                                  size_t ny, size_t nz, const std::vector<double>& gx, const std::vector<double>& gy,
                                  const std::vector<double>& gz)
{
    if (out3d.size() != nx || (nx > 0 && out3d[0].size() != ny) || (nx > 0 && ny > 0 && out3d[0][0].size() != nz))
        return;   // the reference logs and returns on a dimension mismatch (:188-197)
    std::vector<std::complex<double>> img(nx * ny * nz);
    for (size_t z = 0; z < nz; ++z)
        for (size_t y = 0; y < ny; ++y)
            for (size_t x = 0; x < nx; ++x) img[(z * ny + y) * nx + x] = out3d[x][y][z];
    const bool axes = !gx.empty() && !gy.empty() && !gz.empty();
    int st = bsar_write_image(filename.c_str(), reinterpret_cast<const double*>(img.data()), (int)nx, (int)ny, (int)nz,
                              axes ? gx.data() : nullptr, axes ? gy.data() : nullptr, axes ? gz.data() : nullptr);
    if (st != BSAR_OK) throw std::runtime_error("export3DComplexMatrix: " + filename + ": " + bsar_status_string(st));
}

// The back-projection loop of proc_test_script.cpp:395-798 as one call: the
// upsampled samples rCompData ([Npulses][4S], range fastest), trajectories
// and DEM in, output[x][y][z] out.  c, fs4 (= fs * factor), lambda as the
// reference computes them (proc_test_script.cpp:213, 223).
template <class DEM, class TX>
std::vector<std::vector<std::vector<std::complex<double>>>> backProjection(
    const std::vector<std::complex<double>>& rCompData, size_t rangeSamples4, const DEM& dem, const TX& Tx,
    const TX& Rx, const std::vector<double>& gx, const std::vector<double>& gy, const std::vector<double>& gz,
    double n1, double n2, double c, double fs, int factor, double fc, int precision = BSAR_FP64)
{
    const size_t n = Tx.size();
    std::vector<double> tx(3 * n), rx(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) { tx[3 * i + k] = Tx[i][k]; rx[3 * i + k] = Rx[i][k]; }
    FlatDem f(dem);
    bsar_ctx* ctx = context();
    bsar_params p{fc, c, fs, n1, n2, factor, 0};
    bsar_dataset* ds = nullptr;
    check(bsar_dataset_create(ctx, &p, &f.d, tx.data(), rx.data(), (int)n, &ds), ctx, "backProjection");
    std::unique_ptr<bsar_dataset, void (*)(bsar_dataset*)> hold(ds, bsar_dataset_destroy);
    const size_t ndata = rangeSamples4 ? rCompData.size() / rangeSamples4 : 0;
    check(bsar_dataset_load_upsampled(ctx, ds, reinterpret_cast<const double*>(rCompData.data()), (int)rangeSamples4,
                                      (int)ndata), ctx, "backProjection");
    const size_t nx = gx.size(), ny = gy.size(), nz = gz.size();
    std::vector<std::complex<double>> img(nx * ny * nz);
    check(bsar_backproject(ctx, ds, gx.data(), (int)nx, gy.data(), (int)ny, gz.data(), (int)nz, precision,
                           reinterpret_cast<double*>(img.data()), nullptr), ctx, "backProjection");
    std::vector<std::vector<std::vector<std::complex<double>>>> out(
        nx, std::vector<std::vector<std::complex<double>>>(ny, std::vector<std::complex<double>>(nz)));
    for (size_t z = 0; z < nz; ++z)
        for (size_t y = 0; y < ny; ++y)
            for (size_t x = 0; x < nx; ++x) out[x][y][z] = img[(z * ny + y) * nx + x];
    return out;
}

// ---- scenario generators and DEM files (host only, no GPU context) --------
// The results convert to the caller's own structs (any type with the
// reference's member names), so `TrajectoryPoints t = funcao_espiral(...)`
// and `Grid3D g = crearGrid(...)` compile unchanged.

inline void check_file(int st, const char* what)
{
    if (st != BSAR_OK)
        throw std::runtime_error(std::string(what) + ": " + bsar_status_string(st) + " (" + bsar_file_error() + ")");
}

// TrajectoryPoints, funcao_espiral.h:7-11
struct Points {
    std::vector<double> Px, Py, Pz;
    template <class T>
    operator T() const
    {
        T t;
        t.Px = Px;
        t.Py = Py;
        t.Pz = Pz;
        return t;
    }
};

inline Points spiral_points(double N, double R_minor, double R_major, double H_high, double H_low, double Vt,
                            double PRF, bool mirror)
{
    int64_t n = 0;
    check_file(bsar_spiral(N, R_minor, R_major, H_high, H_low, Vt, PRF, mirror, nullptr, nullptr, nullptr, 0, &n),
               "funcao_espiral");
    Points p;
    p.Px.resize((size_t)n);
    p.Py.resize((size_t)n);
    p.Pz.resize((size_t)n);
    check_file(bsar_spiral(N, R_minor, R_major, H_high, H_low, Vt, PRF, mirror, p.Px.data(), p.Py.data(),
                           p.Pz.data(), n, &n), "funcao_espiral");
    return p;
}

// funcao_espiral, funcao_espiral.cpp:5-49
inline Points funcao_espiral(double N, double R_minor, double R_major, double H_high, double H_low, double Vt,
                             double PRF)
{
    return spiral_points(N, R_minor, R_major, H_high, H_low, Vt, PRF, false);
}

// generateTransmitterTrajectory, trajectory_generator.cpp:4-15 (RADAR:
// NumVoltasEsp, RaioMenorEsp, RaioMaiorEsp, AltMaiorEsp, AltMenorEsp, Vt, PRF)
template <class RADAR>
Points generateTransmitterTrajectory(const RADAR& tx)
{
    return spiral_points(tx.NumVoltasEsp, tx.RaioMenorEsp, tx.RaioMaiorEsp, tx.AltMaiorEsp, tx.AltMenorEsp, tx.Vt,
                         tx.PRF, false);
}

// generateReceiverTrajectory, trajectory_generator.cpp:17-33: the same spiral with x, y negated
template <class RADAR>
Points generateReceiverTrajectory(const RADAR& rx)
{
    return spiral_points(rx.NumVoltasEsp, rx.RaioMenorEsp, rx.RaioMaiorEsp, rx.AltMaiorEsp, rx.AltMenorEsp, rx.Vt,
                         rx.PRF, true);
}

// Grid3D, crear_grid.h:8-14: axes plus the [ny][nx][nz] meshgrids
struct GridAxes {
    std::vector<double> xAxis, yAxis, zAxis;
    template <class G>
    operator G() const
    {
        G g;
        g.gridToProc.xAxis = xAxis;
        g.gridToProc.yAxis = yAxis;
        g.gridToProc.zAxis = zAxis;
        g.nx = (int)xAxis.size();
        g.ny = (int)yAxis.size();
        g.nz = (int)zAxis.size();
        const std::vector<double> col(zAxis.size());
        g.X.assign(yAxis.size(), std::vector<std::vector<double>>(xAxis.size(), col));
        g.Y = g.X;
        g.Z = g.X;
        for (size_t j = 0; j < yAxis.size(); ++j)
            for (size_t i = 0; i < xAxis.size(); ++i)
                for (size_t k = 0; k < zAxis.size(); ++k) {
                    g.X[j][i][k] = xAxis[i];
                    g.Y[j][i][k] = yAxis[j];
                    g.Z[j][i][k] = zAxis[k];
                }
        return g;
    }
};

inline std::vector<double> grid_axis(double center, double step, double half)
{
    int n = 0;
    check_file(bsar_grid_axis(center, step, half, nullptr, 0, &n), "crearGrid");
    std::vector<double> v((size_t)n);
    check_file(bsar_grid_axis(center, step, half, v.data(), n, &n), "crearGrid");
    return v;
}

// crearGrid, crear_grid.cpp:6-73 (TARGET: pos[3])
template <class TARGET>
GridAxes crearGrid(const TARGET& target, double dxy = 0.04, double lxy = 0.8, double dz = 0.08, double lz = 0.8)
{
    GridAxes g;
    g.xAxis = grid_axis(target.pos[0], dxy, lxy);
    g.yAxis = grid_axis(target.pos[1], dxy, lxy);
    g.zAxis = grid_axis(target.pos[2], dz, lz);
    return g;
}

// ProcessedDEM as loadDEM returns it, radar_params.h:75-79
struct DemFile {
    std::vector<double> X_vec, Y_vec;
    std::vector<std::vector<double>> Z_DEM;
    template <class D>
    operator D() const
    {
        D d;
        d.X_vec = X_vec;
        d.Y_vec = Y_vec;
        d.Z_DEM = Z_DEM;
        return d;
    }
};

inline DemFile load_dem_file(const std::string& filename, int mode)
{
    bsar_dem_grid g;
    check_file(bsar_dem_load(filename.c_str(), mode, &g), "loadDEM");
    DemFile d;
    d.X_vec.assign(g.dem.x, g.dem.x + g.dem.nx);
    d.Y_vec.assign(g.dem.y, g.dem.y + g.dem.ny);
    d.Z_DEM.resize((size_t)g.dem.zrows);
    for (int j = 0; j < g.dem.zrows; ++j)
        d.Z_DEM[j].assign(g.dem.z + (size_t)j * g.dem.zcols, g.dem.z + (size_t)(j + 1) * g.dem.zcols);
    bsar_dem_grid_free(&g);
    return d;
}

// loadDEM of the build without GDAL, load_dem_simple.cpp:9-224
inline DemFile loadDEM(const std::string& filename) { return load_dem_file(filename, BSAR_DEM_FILE_REFERENCE); }

// A single-band GeoTIFF decoded to local metres about its centre (what the
// GDAL loader documents, load_dem.h:9-17)
inline DemFile loadDEMGeoTIFF(const std::string& filename) { return load_dem_file(filename, BSAR_DEM_FILE_GEOTIFF); }

}  // namespace facade
}  // namespace bsar

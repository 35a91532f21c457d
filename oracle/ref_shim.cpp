// ref_shim.cpp -- extern "C" entry points over the reference's own hot-path
// functions, for building oracle/_ref/libref.so from the unmodified reference
// sources (src/raw_data_processing.cpp, src/fft_utils.cpp,
// src/interpolation.cpp, src/project_paths.cpp, and the scenario generators
// src/funcao_espiral.cpp, src/trajectory_generator.cpp, src/radar_init.cpp,
// src/crear_grid.cpp, src/load_dem_simple.cpp).
//
// TEST INFRASTRUCTURE ONLY: used to pin the C restatement in
// oracle/bsar_oracle.c against the reference itself (tests/golden/) and, on
// the same inputs, as a CPU timing baseline.  Nothing here is reference
// source; it only marshals flat arrays into the reference's types.
#include "raw_data_processing.h"
#include "fft_utils.h"
#include "interpolation.h"
#include "funcao_espiral.h"
#include "trajectory_generator.h"
#include "radar_init.h"
#include "crear_grid.h"
#include "load_dem.h"

#include <cstring>
#include <vector>
#include <complex>

namespace {

ProcessedDEM make_dem(const double* x, int nx, const double* y, int ny,
                      const double* z, int zrows, int zcols)
{
    ProcessedDEM d;
    d.X_vec.assign(x, x + nx);
    d.Y_vec.assign(y, y + ny);
    d.Z_DEM.resize(zrows);
    for (int j = 0; j < zrows; j++) d.Z_DEM[j].assign(z + (size_t)j * zcols, z + (size_t)(j + 1) * zcols);
    return d;
}

std::vector<std::complex<double>> to_cvec(const double* a, size_t n)
{
    std::vector<std::complex<double>> v(n);
    for (size_t i = 0; i < n; i++) v[i] = std::complex<double>(a[2 * i], a[2 * i + 1]);
    return v;
}

void from_cvec(const std::vector<std::complex<double>>& v, double* a)
{
    for (size_t i = 0; i < v.size(); i++) { a[2 * i] = v[i].real(); a[2 * i + 1] = v[i].imag(); }
}

}  // namespace

extern "C" {

// interpolateDEMElevation over m points xy[m][2]
void ref_dem_interp(const double* x, int nx, const double* y, int ny, const double* z, int zrows, int zcols,
                    const double* xy, int m, double* out)
{
    ProcessedDEM d = make_dem(x, nx, y, ny, z, zrows, zcols);
    for (int k = 0; k < m; k++) out[k] = interpolateDEMElevation(xy[2 * k], xy[2 * k + 1], d);
}

// calculateRefractionPoint over m legs seg[m][6] = (x1,y1,z1,x2,y2,z2)
void ref_refraction(const double* x, int nx, const double* y, int ny, const double* z, int zrows, int zcols,
                    const double* seg, int m, double n1, double n2, double* out)
{
    ProcessedDEM d = make_dem(x, nx, y, ny, z, zrows, zcols);
    for (int k = 0; k < m; k++) {
        const double* s = seg + 6 * (size_t)k;
        std::vector<double> p = calculateRefractionPoint(s[0], s[1], s[2], s[3], s[4], s[5], d, n1, n2, false);
        out[3 * k] = p[0]; out[3 * k + 1] = p[1]; out[3 * k + 2] = p[2];
    }
}

// calculaSlantRangeFermat for one target: refraction points and ranges
void ref_slant_range(const double* x, int nx, const double* y, int ny, const double* z, int zrows, int zcols,
                     const double* tx, const double* rx, int n, const double* P, double n1, double n2,
                     double* ida, double* volta, double* ranges4)
{
    ProcessedDEM d = make_dem(x, nx, y, ny, z, zrows, zcols);
    TrajectoryMatrix t;
    t.Tx.resize(n, std::vector<double>(3));
    t.Rx.resize(n, std::vector<double>(3));
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) { t.Tx[i][c] = tx[3 * i + c]; t.Rx[i][c] = rx[3 * i + c]; }
    t.P.assign(P, P + 3);
    Environment env; env.n1 = n1; env.n2 = n2;
    ReflectionPoints refl; RefractionPoints refr;
    RangeData r = calculaSlantRangeFermat(d, t, env, refl, refr, false);
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) { ida[3 * i + c] = refr.P_refrac_ida[i][c]; volta[3 * i + c] = refr.P_refrac_volta[i][c]; }
        ranges4[i] = r.r1Tx[i]; ranges4[n + i] = r.r2Tx[i]; ranges4[2 * n + i] = r.r1Rx[i]; ranges4[3 * n + i] = r.r2Rx[i];
    }
}

// calculateReflectionPoint over m antenna pairs pr[m][6] = (tx, ty, tz, rx, ry, rz)
void ref_reflection(const double* x, int nx, const double* y, int ny, const double* z, int zrows, int zcols,
                    const double* pr, int m, double* out)
{
    ProcessedDEM d = make_dem(x, nx, y, ny, z, zrows, zcols);
    for (int k = 0; k < m; k++) {
        const double* s = pr + 6 * (size_t)k;
        std::vector<double> q = calculateReflectionPoint(s[0], s[1], s[2], s[3], s[4], s[5], d);
        out[3 * k] = q[0]; out[3 * k + 1] = q[1]; out[3 * k + 2] = q[2];
    }
}

// the reflection points calculaSlantRangeFermat stores for one target (refl[n][3])
void ref_slant_reflection(const double* x, int nx, const double* y, int ny, const double* z, int zrows, int zcols,
                          const double* tx, const double* rx, int n, const double* P, double n1, double n2,
                          double* refl_out)
{
    ProcessedDEM d = make_dem(x, nx, y, ny, z, zrows, zcols);
    TrajectoryMatrix t;
    t.Tx.resize(n, std::vector<double>(3));
    t.Rx.resize(n, std::vector<double>(3));
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) { t.Tx[i][c] = tx[3 * i + c]; t.Rx[i][c] = rx[3 * i + c]; }
    t.P.assign(P, P + 3);
    Environment env; env.n1 = n1; env.n2 = n2;
    ReflectionPoints refl; RefractionPoints refr;
    calculaSlantRangeFermat(d, t, env, refl, refr, false);
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) refl_out[3 * i + c] = refl.P_reflection[i][c];
}

void ref_spline(const double* in, int n, int factor, double* out)
{
    from_cvec(splineInterpolation(to_cvec(in, n), factor), out);
}

void ref_fft(double* a, int n)
{
    std::vector<std::complex<double>> v = to_cvec(a, n);
    FFTProcessor::fft(v);
    from_cvec(v, a);
}

void ref_ifft(double* a, int n)
{
    std::vector<std::complex<double>> v = to_cvec(a, n);
    FFTProcessor::ifft(v);
    from_cvec(v, a);
}

void ref_fft_correlation(const double* s, int ns, const double* r, int nr, double* out)
{
    from_cvec(FFTProcessor::fftCorrelation(to_cvec(s, ns), to_cvec(r, nr)), out);
}

void ref_fft_convolution(const double* s, int ns, const double* r, int nr, double* out)
{
    from_cvec(FFTProcessor::fftConvolution(to_cvec(s, ns), to_cvec(r, nr)), out);
}

int ref_chirp(double f0, double f1, double T, double fs, double* out)
{
    RadarTx tx{};
    tx.FreqMenor = f0; tx.FreqMaior = f1; tx.DuracaoPulso = T; tx.fs = fs;
    std::vector<std::complex<double>> c = createCHIRP(tx);
    if (out) from_cvec(c, out);
    return (int)c.size();
}

// processRawData for one point target P with real rcs; aux/raw/root [S][n]
void ref_process_raw(const double* x, int nx, const double* y, int ny, const double* z, int zrows, int zcols,
                     const double* tx, const double* rx, int n, const double* P, double rcs,
                     double fc, double c, double fs_rx, double fs_tx, double f0, double f1, double T,
                     double n1, double n2, int S, double* aux, double* raw, double* root)
{
    ProcessedDEM d = make_dem(x, nx, y, ny, z, zrows, zcols);
    TrajectoryMatrix t;
    t.Tx.resize(n, std::vector<double>(3));
    t.Rx.resize(n, std::vector<double>(3));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) { t.Tx[i][k] = tx[3 * i + k]; t.Rx[i][k] = rx[3 * i + k]; }
    t.P.assign(P, P + 3);
    Environment env; env.n1 = n1; env.n2 = n2;
    ReflectionPoints refl; RefractionPoints refr;
    RangeData r = calculaSlantRangeFermat(d, t, env, refl, refr, false);
    RadarTx rtx{}; rtx.FreqPortadora = fc; rtx.fs = fs_tx; rtx.FreqMenor = f0; rtx.FreqMaior = f1;
    rtx.DuracaoPulso = T; rtx.lamb = c / fc;
    RadarRx rrx{}; rrx.fs = fs_rx;
    Target tg{}; tg.pos[0] = P[0]; tg.pos[1] = P[1]; tg.pos[2] = P[2]; tg.rcs = rcs;
    System sys{}; sys.VelocidadeLuz = c; sys.IndiceMaximo = S;
    RawDataResult res = processRawData(t, r, rtx, rrx, tg, sys, env);
    for (int row = 0; row < S; row++)
        for (int i = 0; i < n; i++) {
            size_t k = (size_t)row * n + i;
            aux[2 * k] = res.auxData[row][i].real(); aux[2 * k + 1] = res.auxData[row][i].imag();
            raw[2 * k] = res.rawData[row][i].real(); raw[2 * k + 1] = res.rawData[row][i].imag();
            root[2 * k] = res.rootData[row][i].real(); root[2 * k + 1] = res.rootData[row][i].imag();
        }
}

// funcao_espiral: returns the point count, fills up to cap positions
int64_t ref_spiral(double N, double r_minor, double r_major, double h_high, double h_low, double vt, double prf,
                   double* px, double* py, double* pz, int64_t cap)
{
    TrajectoryPoints t = funcao_espiral(N, r_minor, r_major, h_high, h_low, vt, prf);
    for (size_t i = 0; i < t.Px.size() && (int64_t)i < cap; i++) { px[i] = t.Px[i]; py[i] = t.Py[i]; pz[i] = t.Pz[i]; }
    return (int64_t)t.Px.size();
}

// generateTransmitterTrajectory(initializeRadarTx()) (rx = 0) or
// generateReceiverTrajectory(initializeRadarRx()) (rx = 1), out[n][3]
int64_t ref_default_trajectory(int rx, double* out, int64_t cap)
{
    TrajectoryPoints t = rx ? generateReceiverTrajectory(initializeRadarRx())
                            : generateTransmitterTrajectory(initializeRadarTx());
    for (size_t i = 0; i < t.Px.size() && (int64_t)i < cap; i++) {
        out[3 * i] = t.Px[i]; out[3 * i + 1] = t.Py[i]; out[3 * i + 2] = t.Pz[i];
    }
    return (int64_t)t.Px.size();
}

// crearGrid around pos: axes written back to back into ax (cap values), sizes to n[3]
void ref_grid(const double* pos, double dxy, double lxy, double dz, double lz, double* ax, int cap, int* n)
{
    Target t = initializeTarget();
    updatePointToProc(t, pos[0], pos[1], pos[2]);
    Grid3D g = crearGrid(t, dxy, lxy, dz, lz);
    n[0] = (int)g.gridToProc.xAxis.size(); n[1] = (int)g.gridToProc.yAxis.size(); n[2] = (int)g.gridToProc.zAxis.size();
    int k = 0;
    for (const auto* v : {&g.gridToProc.xAxis, &g.gridToProc.yAxis, &g.gridToProc.zAxis})
        for (double a : *v) if (k < cap) ax[k++] = a;
}

// loadDEM (load_dem_simple.cpp): sizes to n[4] = (nx, ny, zrows, zcols of row 0);
// arrays filled up to their capacities
void ref_load_dem(const char* path, double* x, int capx, double* y, int capy, double* z, int capz, int* n)
{
    ProcessedDEM d = loadDEM(path);
    n[0] = (int)d.X_vec.size(); n[1] = (int)d.Y_vec.size(); n[2] = (int)d.Z_DEM.size();
    n[3] = d.Z_DEM.empty() ? 0 : (int)d.Z_DEM[0].size();
    for (int i = 0; i < n[0] && i < capx; i++) x[i] = d.X_vec[i];
    for (int i = 0; i < n[1] && i < capy; i++) y[i] = d.Y_vec[i];
    int k = 0;
    for (const auto& row : d.Z_DEM)
        for (double v : row) if (k < capz) z[k++] = v;
}

}  // extern "C"

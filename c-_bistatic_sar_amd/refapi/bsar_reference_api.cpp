// bsar_reference_api.cpp -- the reference's hot-path declarations defined over
// libbsar, for a build that keeps every call site, header and type of
// DavidAtenciaMondragon/c-_bistatic_sar as it is.
//
// Compile this file in place of src/raw_data_processing.cpp,
// src/fft_utils.cpp and src/interpolation.cpp, with the reference's include/
// directory and this repository's include/ on the include path, and link
// libbsar.so.  It defines every function those three files define, with the
// signatures of include/raw_data_processing.h:45-151, include/fft_utils.h:
// 13-84 and include/interpolation.h:8; the transforms, searches and spline run
// on the GPU through bsar_facade.hpp (one context per host thread, so the
// functions stay callable from the driver's OpenMP threads).
//
// Kept from the reference (the behaviour a caller can observe): environment.txt
// keys and defaults (raw_data_processing.cpp:14-46), the pass-through helpers
// (:48-110), processRawData's delay bins, phases and fftshifted reference
// built from the caller's ranges (:292-452), the FFT's power-of-two rule
// (fft_utils.cpp:17-20).  Not kept: the informational console lines of
// calculaSlantRangeFermat and processRawData.
#include "raw_data_processing.h"
#include "fft_utils.h"
#include "interpolation.h"
#include "project_paths.h"

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "bsar_facade.hpp"

namespace fac = bsar::facade;

// ---- raw_data_processing.h ------------------------------------------------

Environment initializeEnvironment()
{
    Environment env;
    env.n1 = 1.0;
    env.n2 = 4.0;
    std::ifstream file(ProjectPaths::getProcPath("environment.txt"));
    if (!file.is_open()) {
        std::cout << "Environment file not found, using defaults: n1=" << env.n1 << ", n2=" << env.n2 << std::endl;
        return env;
    }
    std::string line;
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        std::istringstream iss(line);
        std::string key, value;
        if (!std::getline(iss, key, '=') || !std::getline(iss, value)) continue;
        if (key == "n1") env.n1 = std::stod(value);
        else if (key == "n2") env.n2 = std::stod(value);
    }
    std::cout << "Environment loaded from file: n1=" << env.n1 << ", n2=" << env.n2 << std::endl;
    return env;
}

TrajectoryMatrix prepareTrajectoryData(const TrajectoryPoints& trajTx, const TrajectoryPoints& trajRx,
                                       const Target& target)
{
    TrajectoryMatrix t;
    const size_t n = trajTx.Px.size();
    t.Tx.assign(n, std::vector<double>(3));
    t.Rx.assign(n, std::vector<double>(3));
    for (size_t i = 0; i < n; ++i) {
        t.Tx[i] = {trajTx.Px[i], trajTx.Py[i], trajTx.Pz[i]};
        t.Rx[i] = {trajRx.Px[i], trajRx.Py[i], trajRx.Pz[i]};
    }
    t.P = {target.pos[0], target.pos[1], target.pos[2]};
    return t;
}

TrajectoryMatrix prepareTrajectoryData(const std::vector<std::vector<double>>& Tx,
                                       const std::vector<std::vector<double>>& Rx,
                                       const std::vector<double>& target_pos)
{
    TrajectoryMatrix t;
    t.Tx = Tx;
    t.Rx = Rx;
    t.P = target_pos;
    return t;
}

ProcessedDEM prepareDEMData(const ProcessedDEM& demData)
{
    std::cout << "DEM data prepared for processing\n  X vector size: " << demData.X_vec.size()
              << "\n  Y vector size: " << demData.Y_vec.size() << "\n  Z matrix size: " << demData.Z_DEM.size()
              << "x" << (demData.Z_DEM.empty() ? 0 : demData.Z_DEM[0].size()) << std::endl;
    return demData;
}

RangeData calculaSlantRangeFermat(const ProcessedDEM& demData, const TrajectoryMatrix& trajectories,
                                  const Environment& env, ReflectionPoints& strReflexao,
                                  RefractionPoints& strRefraccoes, bool bPlotVerbose)
{
    return fac::calculaSlantRangeFermat(demData, trajectories, env, strReflexao, strRefraccoes, bPlotVerbose);
}

std::vector<std::complex<double>> createCHIRP(const RadarTx& radarTx) { return fac::createCHIRP(radarTx); }

RawDataResult processRawData(const TrajectoryMatrix& trajectories, const RangeData& ranges, const RadarTx& radarTx,
                             const RadarRx& radarRx, const Target& target, const System& system,
                             const Environment& env)
{
    RawDataResult res;
    const int S = system.IndiceMaximo;
    const size_t n = trajectories.Tx.size();
    res.chirp = createCHIRP(radarTx);
    res.auxData.assign((size_t)(S > 0 ? S : 0), std::vector<std::complex<double>>(n, 0.0));
    res.rawData = res.auxData;
    res.rootData = res.auxData;
    if (S <= 0 || n == 0) return res;
    // point-target responses from the caller's ranges: row 1 + round(t fs),
    // phase -(2 pi / lambda) L (raw_data_processing.cpp:322-361)
    std::vector<std::complex<double>> cols((size_t)S * n, 0.0);   // [pulse][range]
    for (size_t i = 0; i < n; ++i) {
        const double L = env.n1 * ranges.r1Tx[i] + env.n2 * ranges.r2Tx[i] + env.n1 * ranges.r1Rx[i] +
                         env.n2 * ranges.r2Rx[i];
        const int bin = 1 + static_cast<int>(std::round(L / system.VelocidadeLuz * radarRx.fs));
        const double phi = -(2.0 * M_PI / radarTx.lamb) * L;
        if (bin >= 0 && bin < S) {
            const std::complex<double> v = target.rcs * std::complex<double>(std::cos(phi), std::sin(phi));
            res.auxData[(size_t)bin][i] = v;
            cols[i * (size_t)S + (size_t)bin] = v;
        }
    }
    // the chirp's second half at the start of the window and its first half at
    // the end (:371-394); the reference indexes past its S-sample pulse when
    // the chirp is longer than the window, which is refused here
    const size_t Lc = res.chirp.size(), half = Lc / 2;
    if (Lc > (size_t)S) throw std::runtime_error("processRawData: chirp longer than System::IndiceMaximo");
    std::vector<std::complex<double>> ref((size_t)S, 0.0);
    for (size_t k = 0; k < Lc - half; ++k) ref[k] = res.chirp[half + k];
    for (size_t k = 0; k < half; ++k) ref[(size_t)S - half + k] = res.chirp[k];
    // raw = conv(aux, ref), root = corr(raw, ref) per pulse (:406-446), batched on the GPU
    bsar_ctx* c = fac::context();
    std::vector<std::complex<double>> raw((size_t)S * n), root((size_t)S * n);
    fac::check(bsar_range_compress(c, reinterpret_cast<const double*>(cols.data()), S, (int)n,
                                   reinterpret_cast<const double*>(ref.data()), 0, BSAR_FP64,
                                   reinterpret_cast<double*>(raw.data())), c, "processRawData");
    fac::check(bsar_range_compress(c, reinterpret_cast<const double*>(raw.data()), S, (int)n,
                                   reinterpret_cast<const double*>(ref.data()), 1, BSAR_FP64,
                                   reinterpret_cast<double*>(root.data())), c, "processRawData");
    for (size_t i = 0; i < n; ++i)
        for (int r = 0; r < S; ++r) {
            res.rawData[(size_t)r][i] = raw[i * (size_t)S + (size_t)r];
            res.rootData[(size_t)r][i] = root[i * (size_t)S + (size_t)r];
        }
    return res;
}

std::vector<double> calculateRefractionPoint(double x1, double y1, double z1, double x2, double y2, double z2,
                                             const ProcessedDEM& demData, double n1, double n2, bool bPlotVerbose)
{
    return fac::calculateRefractionPoint(x1, y1, z1, x2, y2, z2, demData, n1, n2, bPlotVerbose);
}

std::vector<double> calculateReflectionPoint(double txX, double txY, double txZ, double rxX, double rxY,
                                             double rxZ, const ProcessedDEM& demData)
{
    return fac::calculateReflectionPoint(txX, txY, txZ, rxX, rxY, rxZ, demData);
}

double interpolateDEMElevation(double x, double y, const ProcessedDEM& demData)
{
    return fac::interpolateDEMElevation(x, y, demData);
}

std::vector<std::complex<double>> correlateWithReference(const std::vector<std::complex<double>>& signal,
                                                         const std::vector<std::complex<double>>& reference)
{
    return fac::FFTProcessor::fftCorrelation(signal, reference);
}

// ---- fft_utils.h ----------------------------------------------------------

void FFTProcessor::fft(std::vector<std::complex<double>>& input) { fac::FFTProcessor::fft(input); }
void FFTProcessor::ifft(std::vector<std::complex<double>>& input) { fac::FFTProcessor::ifft(input); }

std::vector<std::complex<double>> FFTProcessor::fftCorrelation(const std::vector<std::complex<double>>& signal,
                                                               const std::vector<std::complex<double>>& reference)
{
    return fac::FFTProcessor::fftCorrelation(signal, reference);
}

std::vector<std::complex<double>> FFTProcessor::fftConvolution(const std::vector<std::complex<double>>& signal,
                                                               const std::vector<std::complex<double>>& reference)
{
    return fac::FFTProcessor::fftConvolution(signal, reference);
}

std::vector<std::complex<double>> FFTProcessor::multiply(const std::vector<std::complex<double>>& a,
                                                         const std::vector<std::complex<double>>& b)
{
    return fac::FFTProcessor::multiply(a, b);
}

std::vector<std::complex<double>> FFTProcessor::padToPowerOfTwo(const std::vector<std::complex<double>>& input)
{
    return fac::FFTProcessor::padToPowerOfTwo(input);
}

void FFTProcessor::bitReverse(std::vector<std::complex<double>>& input) { fac::FFTProcessor::bitReverse(input); }
bool FFTProcessor::isPowerOfTwo(size_t n) { return fac::FFTProcessor::isPowerOfTwo(n); }
size_t FFTProcessor::nextPowerOfTwo(size_t n) { return fac::FFTProcessor::nextPowerOfTwo(n); }

// ---- interpolation.h ------------------------------------------------------

std::vector<std::complex<double>> splineInterpolation(const std::vector<std::complex<double>>& data, int factor)
{
    return fac::splineInterpolation(data, factor);
}

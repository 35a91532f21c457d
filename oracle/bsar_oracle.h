/*
 * bsar_oracle.h -- CPU FP64 restatement of the bistatic-SAR imaging hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never part of the product path.
 *
 * Every function restates one reference routine (DavidAtenciaMondragon/
 * c-_bistatic_sar) with the same operation order, so that an FP64 build
 * without FMA contraction reproduces the reference bit for bit wherever only
 * + - * / sqrt are involved.  Reference file:line citations are on each
 * declaration.  Complex numbers are interleaved (re, im) doubles.
 *
 * Parity pins (what checks each function; see tests/):
 *   or_dem_interp        golden vectors from the compiled reference + plane
 *                        DEM closed form + node values + out-of-range cases
 *   or_refraction_point  golden vectors from the compiled reference + flat
 *                        DEM Snell closed form + n1 == n2 straight line
 *   or_spline            golden vectors + node/constant/ramp invariants
 *   or_fft / or_ifft     O(n^2) DFT brute force, round trip, Parseval
 *   or_fft_correlation   golden vectors + brute-force circular correlation
 *   or_fft_convolution   golden vectors + brute-force circular convolution
 *   or_chirp             golden vectors + closed form
 *   or_simulate          golden vectors (reference processRawData)
 *   or_backproject       golden output_3D_matrix.bin of the reference's own
 *                        proc_test_script on small inputs + linearity +
 *                        single-pulse closed form
 */
#ifndef BSAR_ORACLE_H
#define BSAR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DEM as the reference holds it (radar_params.h:75-79): X_vec[nx], Y_vec[ny],
 * Z_DEM[zrows][zcols] indexed Z[j][i] = [y][x].  Entries beyond zrows/zcols
 * read as 0.0, as the bounds checks at raw_data_processing.cpp:649-652 do. */
typedef struct {
    const double* x; int nx;
    const double* y; int ny;
    const double* z; int zrows; int zcols;
} or_dem;

/* interpolateDEMElevation, raw_data_processing.cpp:615-680 */
double or_dem_interp(const or_dem* dem, double x, double y);

/* calculateRefractionPoint, raw_data_processing.cpp:455-572 */
void or_refraction_point(double x1, double y1, double z1,
                         double x2, double y2, double z2,
                         const or_dem* dem, double n1, double n2,
                         double out[3]);

/* or_refraction_point over m independent legs legs[k] = (x1,y1,z1,x2,y2,z2),
 * OpenMP over legs (threads <= 0: the runtime default). */
void or_refraction_batch(const double* legs, size_t m, const or_dem* dem, double n1, double n2,
                         double* out, int threads);

/* calculateReflectionPoint, raw_data_processing.cpp:575-612 (dead for the
 * image; run only so CPU timing matches the reference's work). */
void or_reflection_point(double txX, double txY, double txZ,
                         double rxX, double rxY, double rxZ,
                         const or_dem* dem, double out[3]);

/* Refraction points of calculaSlantRangeFermat, raw_data_processing.cpp:153-
 * 193, for one target P and n pulses: ida[i] = refraction(Tx_i -> P, n1, n2),
 * volta[i] = refraction(P -> Rx_i, n2, n1).  tx, rx are [n][3]. */
void or_slant_range_fermat(const or_dem* dem, const double* tx, const double* rx,
                           int n, const double P[3], double n1, double n2,
                           double* ida, double* volta, double* ranges4 /* [4][n] or NULL */);

/* splineInterpolation, interpolation.cpp:4-51.  in: n complex, out: n*factor
 * complex.  Requires n >= 2. */
void or_spline(const double* in, size_t n, int factor, double* out);

/* FFTProcessor::fft / ifft, fft_utils.cpp:11-60 (radix-2 DIT, bit reversal,
 * twiddle recurrence).  Non-power-of-two n is a silent no-op (fft_utils.cpp:
 * 17-20); ifft still conjugates and scales around it (fft_utils.cpp:46-60). */
void or_fft(double* a, size_t n);
void or_ifft(double* a, size_t n);

/* FFTProcessor::fftCorrelation / fftConvolution, fft_utils.cpp:63-137.
 * out has max(nsig, nref) complex entries. */
void or_fft_correlation(const double* sig, size_t nsig, const double* ref, size_t nref, double* out);
void or_fft_convolution(const double* sig, size_t nsig, const double* ref, size_t nref, double* out);

/* createCHIRP, raw_data_processing.cpp:257-290.  Returns L = int(T*fs)+1;
 * writes L complex samples to out when out != NULL. */
int or_chirp(double f0, double f1, double T, double fs, double* out);

/* processRawData, raw_data_processing.cpp:292-452, generalised to several
 * point scatterers (responses superpose in auxData).  Writes root/raw/aux as
 * [S][N] complex in the reference's in-memory layout (row = range sample).
 * rcs is complex per scatterer (the reference uses a real rcs = 1.0).
 * ranges per (scatterer, pulse) come from or_slant_range_fermat. */
typedef struct {
    double fc, c, fs;        /* carrier, light speed, ADC rate (RadarRx.fs) */
    double f0, f1, T;        /* chirp FreqMenor, FreqMaior, DuracaoPulso */
    double fs_chirp;         /* chirp sampling rate (RadarTx.fs) */
    double n1, n2;
    int S;                   /* IndiceMaximo */
} or_sim_params;

void or_simulate(const or_sim_params* p, const or_dem* dem,
                 const double* tx, const double* rx, int n,
                 const double* scat /* [ns][3] */, const double* rcs /* [ns] complex */, int ns,
                 double* aux, double* raw, double* root /* each [S][n] complex, may be NULL */);

/* Range-compressed data upsampling of proc_test_script.cpp:303-354: rootData
 * read as flat r + p*S (range fastest), every pulse upsampled with
 * or_spline.  out: [N][S*factor] complex, range fastest. */
void or_upsample(const double* root, size_t S, size_t N, int factor, double* out);

/* Back-projection body of proc_test_script.cpp:451-798.
 *   rcomp   : [ndata][S4] complex, range fastest (the upsampled rCompData)
 *   fs4     : RadarRx.fs * factorInterp (proc_test_script.cpp:223)
 *   lambda  : c / fc (proc_test_script.cpp:213)
 *   out     : [nz][ny][nx] complex, x fastest (the export order of
 *             data_export.cpp:241-256) -- written only for processed voxels
 *   subset  : optional list of flat voxel ids (coord_idx = z*nx*ny + x*ny + y,
 *             proc_test_script.cpp:760-763); NULL = whole grid
 *   reflection : run the dead calculateReflectionPoint search as the
 *             reference does (timing fidelity only)
 *   threads : OpenMP threads (<= 0: runtime default) */
typedef struct {
    double c, fs4, lambda, n1, n2;
} or_bp_params;

void or_backproject(const or_bp_params* p, const or_dem* dem,
                    const double* tx, const double* rx, int n,
                    const double* rcomp, size_t S4, size_t ndata,
                    const double* gx, int nx, const double* gy, int ny, const double* gz, int nz,
                    double* out, const int64_t* subset, int64_t nsubset,
                    int reflection, int threads);

/* One voxel's back-projected value (same body, no OpenMP). */
void or_backproject_voxel(const or_bp_params* p, const or_dem* dem,
                          const double* tx, const double* rx, int n,
                          const double* rcomp, size_t S4, size_t ndata,
                          const double P[3], int reflection, double out[2]);

/* Flop count per voxel-pulse used for roofline accounting:
 * 65,406 + 20*N (two 709-candidate Fermat solves, the R1 means, delay/phase,
 * lerp and accumulate). */
double or_flops_per_voxel_pulse(int n);

#ifdef __cplusplus
}
#endif
#endif

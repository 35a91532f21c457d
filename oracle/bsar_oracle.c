/*
 * bsar_oracle.c -- CPU FP64 restatement of the bistatic-SAR imaging hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product.  Build: gcc -O3 -fopenmp -ffp-contract=off (oracle/Makefile).
 *
 * Operation order follows the reference source line by line (citations are
 * file:line into DavidAtenciaMondragon/c-_bistatic_sar).  pow(x, 2) in the
 * reference is written x*x here: GCC folds pow(x, 2.0) to x*x, and both are
 * the correctly rounded square.  std::complex arithmetic is spelled out:
 * (a+ib)(c+id) = (ac-bd, ad+bc), complex*real scales both parts.
 */
#include "bsar_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a,
 * max(a,b) = (a < b) ? b : a. */
static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }

static inline double dem_z(const or_dem* d, size_t j, size_t i)
{
    /* raw_data_processing.cpp:649-652: missing rows/columns read as 0.0 */
    if ((int)j < d->zrows && (int)i < d->zcols) return d->z[j * (size_t)d->zcols + i];
    return 0.0;
}

/* interpolateDEMElevation, raw_data_processing.cpp:615-680 */
double or_dem_interp(const or_dem* d, double x, double y)
{
    if (d->nx <= 0 || d->ny <= 0 || d->zrows <= 0) return 0.0;      /* :616-618 */
    size_t nx = (size_t)d->nx, ny = (size_t)d->ny;
    size_t i0 = 0, i1 = 0, j0 = 0, j1 = 0;                            /* :621-622 */
    for (size_t i = 0; i + 1 < nx; i++) {                             /* :625-631 */
        if (x >= d->x[i] && x <= d->x[i + 1]) { i0 = i; i1 = i + 1; break; }
    }
    for (size_t j = 0; j + 1 < ny; j++) {                             /* :634-640 */
        if (y >= d->y[j] && y <= d->y[j + 1]) { j0 = j; j1 = j + 1; break; }
    }
    /* :643-646 clamps are no-ops for indices found above */
    double z00 = dem_z(d, j0, i0);                                    /* :649-652 */
    double z01 = dem_z(d, j0, i1);
    double z10 = dem_z(d, j1, i0);
    double z11 = dem_z(d, j1, i1);
    if (i0 == i1 && j0 == j1) return z00;                             /* :655-657 */
    double dx = d->x[i1] - d->x[i0];                                  /* :659-662 */
    double dy = d->y[j1] - d->y[j0];
    if (dx == 0) dx = 1.0;                                            /* :664-665 */
    if (dy == 0) dy = 1.0;
    double wx = (x - d->x[i0]) / dx;                                  /* :667-668 */
    double wy = (y - d->y[j0]) / dy;
    wx = std_max(0.0, std_min(1.0, wx));                              /* :670-671 */
    wy = std_max(0.0, std_min(1.0, wy));
    double z = z00 * (1 - wx) * (1 - wy) +                            /* :674-677 */
               z01 * wx * (1 - wy) +
               z10 * (1 - wx) * wy +
               z11 * wx * wy;
    return z;
}

/* calculateRefractionPoint, raw_data_processing.cpp:455-572 */
void or_refraction_point(double x1, double y1, double z1,
                         double x2, double y2, double z2,
                         const or_dem* dem, double n1, double n2, double out[3])
{
    double midX = (x1 + x2) / 2.0;                                    /* :478-479 */
    double midY = (y1 + y2) / 2.0;
    double ddx = x2 - x1, ddy = y2 - y1;
    double directDistance = sqrt(ddx * ddx + ddy * ddy);              /* :483 */
    double searchRadius = std_max(directDistance, 500.0);             /* :484 */
    double bestX = midX, bestY = midY;                                /* :496 */
    double minOpticalPath = 1e9;                                      /* :497 */
    const int numLevels = 5;                                          /* :500 */
    double currentRadius = searchRadius;
    for (int level = 0; level < numLevels; level++) {                 /* :507 */
        int gridSize = (level == 0) ? 15 : 11;                        /* :508 */
        double minX = bestX - currentRadius;                          /* :509-512 */
        double maxX = bestX + currentRadius;
        double minY = bestY - currentRadius;
        double maxY = bestY + currentRadius;
        for (int i = 0; i < gridSize; i++) {                          /* :523-543 */
            for (int j = 0; j < gridSize; j++) {
                double px = minX + i * (maxX - minX) / (gridSize - 1);
                double py = minY + j * (maxY - minY) / (gridSize - 1);
                double pz = or_dem_interp(dem, px, py);
                double a = px - x1, b = py - y1, cz = pz - z1;
                double d1 = sqrt(a * a + b * b + cz * cz);
                double e = x2 - px, f = y2 - py, g = z2 - pz;
                double d2 = sqrt(e * e + f * f + g * g);
                double opticalPath = n1 * d1 + n2 * d2;
                if (opticalPath < minOpticalPath) {
                    minOpticalPath = opticalPath;
                    bestX = px;
                    bestY = py;
                }
            }
        }
        currentRadius /= (level == 0) ? 5.0 : 3.0;                    /* :551 */
    }
    out[0] = bestX;                                                   /* :555-557 */
    out[1] = bestY;
    out[2] = or_dem_interp(dem, bestX, bestY);
}

/* calculateReflectionPoint, raw_data_processing.cpp:575-612 */
void or_reflection_point(double txX, double txY, double txZ,
                         double rxX, double rxY, double rxZ,
                         const or_dem* d, double out[3])
{
    (void)txZ; (void)rxZ;
    double midX = (txX + rxX) / 2.0;
    double midY = (txY + rxY) / 2.0;
    double minDist = 1e9, bestZ = 0.0;
    if (d->nx > 0 && d->ny > 0 && d->zrows > 0) {
        for (int i = 0; i < d->nx; i++) {
            for (int j = 0; j < d->ny; j++) {
                if (j < d->zrows && i < d->zcols) {
                    double dx = d->x[i] - midX;
                    double dy = d->y[j] - midY;
                    double dist = sqrt(dx * dx + dy * dy);
                    if (dist < minDist) { minDist = dist; bestZ = d->z[(size_t)j * d->zcols + i]; }
                }
            }
        }
    }
    out[0] = midX; out[1] = midY; out[2] = bestZ;
}

static inline double dist3(double ax, double ay, double az, double bx, double by, double bz)
{
    double a = ax - bx, b = ay - by, c = az - bz;
    return sqrt(a * a + b * b + c * c);
}

/* calculaSlantRangeFermat, raw_data_processing.cpp:153-222 */
void or_slant_range_fermat(const or_dem* dem, const double* tx, const double* rx,
                           int n, const double P[3], double n1, double n2,
                           double* ida, double* volta, double* ranges4)
{
    for (int i = 0; i < n; i++) {
        const double* T = tx + 3 * (size_t)i;
        const double* R = rx + 3 * (size_t)i;
        double qa[3], qb[3];
        or_refraction_point(T[0], T[1], T[2], P[0], P[1], P[2], dem, n1, n2, qa);  /* :166-172 */
        or_refraction_point(P[0], P[1], P[2], R[0], R[1], R[2], dem, n2, n1, qb);  /* :175-181 */
        if (ida)   memcpy(ida + 3 * (size_t)i, qa, sizeof qa);
        if (volta) memcpy(volta + 3 * (size_t)i, qb, sizeof qb);
        if (ranges4) {                                                 /* :197-222 */
            ranges4[i]         = dist3(T[0], T[1], T[2], qa[0], qa[1], qa[2]);
            ranges4[n + i]     = dist3(P[0], P[1], P[2], qa[0], qa[1], qa[2]);
            ranges4[2 * n + i] = dist3(R[0], R[1], R[2], qb[0], qb[1], qb[2]);
            ranges4[3 * n + i] = dist3(P[0], P[1], P[2], qb[0], qb[1], qb[2]);
        }
    }
}

/* splineInterpolation, interpolation.cpp:4-51 */
void or_spline(const double* in, size_t n, int factor, double* out)
{
    size_t newSize = n * (size_t)factor;
    for (size_t i = 0; i < newSize; i++) {
        double pos = (double)i / factor;                                /* :18 */
        size_t idx = (size_t)pos;                                       /* :19 */
        double frac = pos - idx;                                        /* :20 */
        if (idx >= n - 1) { idx = n - 2; frac = 1.0; }                  /* :22-25 */
        double t = frac, t2 = t * t, t3 = t2 * t;                       /* :28-30 */
        double h1 = 2 * t3 - 3 * t2 + 1;                                /* :33-36 */
        double h2 = -2 * t3 + 3 * t2;
        double h3 = t3 - 2 * t2 + t;
        double h4 = t3 - t2;
        for (int c = 0; c < 2; c++) {                                   /* :39-46 */
            double a = in[2 * idx + c], b = in[2 * (idx + 1) + c];
            double dPrev = (idx > 0) ? (in[2 * idx + c] - in[2 * (idx - 1) + c]) : 0;
            double dNext = (idx < n - 2) ? (in[2 * (idx + 2) + c] - in[2 * (idx + 1) + c]) : 0;
            out[2 * i + c] = h1 * a + h2 * b + h3 * dPrev + h4 * dNext;
        }
    }
}

static int is_pow2(size_t n) { return n > 0 && (n & (n - 1)) == 0; }   /* fft_utils.cpp:183-185 */
static size_t next_pow2(size_t n)                                       /* fft_utils.cpp:187-195 */
{
    if (n <= 1) return 1;
    size_t p = 1;
    while (p < n) p *= 2;
    return p;
}

static void bit_reverse(double* a, size_t N)                            /* fft_utils.cpp:164-181 */
{
    size_t j = 0;
    for (size_t i = 1; i < N; i++) {
        size_t bit = N >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j ^= bit;
        if (i < j) {
            double r = a[2 * i], m = a[2 * i + 1];
            a[2 * i] = a[2 * j]; a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = r; a[2 * j + 1] = m;
        }
    }
}

/* FFTProcessor::fft, fft_utils.cpp:11-44 */
void or_fft(double* a, size_t N)
{
    if (N <= 1) return;
    if (!is_pow2(N)) return;                                            /* :17-20 silent no-op */
    bit_reverse(a, N);
    for (size_t len = 2; len <= N; len *= 2) {
        double angle = -2.0 * M_PI / len;                               /* :27-28 */
        double wr0 = cos(angle), wi0 = sin(angle);
        for (size_t i = 0; i < N; i += len) {
            double wr = 1.0, wi = 0.0;
            for (size_t j = 0; j < len / 2; j++) {
                double* u = a + 2 * (i + j);
                double* v = a + 2 * (i + j + len / 2);
                double vr = v[0] * wr - v[1] * wi;                      /* :35 input*w */
                double vi = v[0] * wi + v[1] * wr;
                double ur = u[0], ui = u[1];
                u[0] = ur + vr; u[1] = ui + vi;                         /* :37-38 */
                v[0] = ur - vr; v[1] = ui - vi;
                double nr = wr * wr0 - wi * wi0;                        /* :40 w *= wlen */
                double ni = wr * wi0 + wi * wr0;
                wr = nr; wi = ni;
            }
        }
    }
}

/* FFTProcessor::ifft, fft_utils.cpp:46-60 */
void or_ifft(double* a, size_t N)
{
    for (size_t i = 0; i < N; i++) a[2 * i + 1] = -a[2 * i + 1];
    or_fft(a, N);
    double dN = (double)N;
    for (size_t i = 0; i < N; i++) {
        a[2 * i] = a[2 * i] / dN;
        a[2 * i + 1] = -a[2 * i + 1] / dN;
    }
}

/* fftCorrelation (conj = 1) / fftConvolution (conj = 0), fft_utils.cpp:63-137 */
static void fft_corr_conv(const double* sig, size_t nsig, const double* ref, size_t nref,
                          double* out, int conj)
{
    size_t N = nsig > nref ? nsig : nref;
    size_t P = next_pow2(N);
    double* fs = (double*)calloc(2 * P, sizeof(double));
    double* fr = (double*)calloc(2 * P, sizeof(double));
    memcpy(fs, sig, 2 * nsig * sizeof(double));
    memcpy(fr, ref, 2 * nref * sizeof(double));
    or_fft(fs, P);
    or_fft(fr, P);
    for (size_t i = 0; i < P; i++) {
        double a = fs[2 * i], b = fs[2 * i + 1];
        double c = fr[2 * i], d = conj ? -fr[2 * i + 1] : fr[2 * i + 1];
        fs[2 * i] = a * c - b * d;
        fs[2 * i + 1] = a * d + b * c;
    }
    or_ifft(fs, P);
    memcpy(out, fs, 2 * N * sizeof(double));
    free(fs);
    free(fr);
}

void or_fft_correlation(const double* sig, size_t nsig, const double* ref, size_t nref, double* out)
{
    fft_corr_conv(sig, nsig, ref, nref, out, 1);
}

void or_fft_convolution(const double* sig, size_t nsig, const double* ref, size_t nref, double* out)
{
    fft_corr_conv(sig, nsig, ref, nref, out, 0);
}

/* createCHIRP, raw_data_processing.cpp:257-290 */
int or_chirp(double f0, double f1, double T, double fs, double* out)
{
    double k = (f1 - f0) / T;                                           /* :267 */
    int N = (int)(T * fs) + 1;                                          /* :271 */
    if (out) {
        for (int n = 0; n < N; n++) {
            double t = n / fs;                                          /* :276 */
            double phase = 2.0 * M_PI * (f0 * t + (k / 2.0) * t * t);  /* :279 */
            out[2 * n] = cos(phase);
            out[2 * n + 1] = sin(phase);
        }
    }
    return N;
}

/* Reference signal of processRawData, raw_data_processing.cpp:371-394:
 * second half of the chirp at [0, L-L/2), first half at [S-L/2, S). */
static void reference_signal(const double* chirp, int L, int S, double* refsig)
{
    memset(refsig, 0, 2 * (size_t)S * sizeof(double));
    int half = L / 2;
    for (int i = 0; i < L - half; ++i) {
        refsig[2 * i] = chirp[2 * (half + i)];
        refsig[2 * i + 1] = chirp[2 * (half + i) + 1];
    }
    int start = S - half;
    for (int i = 0; i < half; ++i) {
        refsig[2 * (start + i)] = chirp[2 * i];
        refsig[2 * (start + i) + 1] = chirp[2 * i + 1];
    }
}

/* processRawData, raw_data_processing.cpp:292-452 (several scatterers) */
void or_simulate(const or_sim_params* p, const or_dem* dem,
                 const double* tx, const double* rx, int n,
                 const double* scat, const double* rcs, int ns,
                 double* aux_out, double* raw_out, double* root_out)
{
    int S = p->S;
    size_t SN = (size_t)S * (size_t)n;
    double* aux = (double*)calloc(2 * SN, sizeof(double));
    double lamb = p->c / p->fc;                                         /* gs_test_script.cpp:35 */
    double* r4 = (double*)malloc(4 * (size_t)n * sizeof(double));
    for (int s = 0; s < ns; s++) {
        or_slant_range_fermat(dem, tx, rx, n, scat + 3 * (size_t)s, p->n1, p->n2, NULL, NULL, r4);
        for (int i = 0; i < n; i++) {                                   /* :322-361 */
            double totalRange = p->n1 * r4[i] + p->n2 * r4[n + i] +
                                p->n1 * r4[2 * n + i] + p->n2 * r4[3 * n + i];
            double t = totalRange / p->c;
            int rngBin = 1 + (int)round(t * p->fs);
            double phi = -(2.0 * M_PI / lamb) * totalRange;
            if (rngBin >= 0 && rngBin < S) {
                double c = cos(phi), sn = sin(phi);
                double ar = rcs[2 * s], ai = rcs[2 * s + 1];
                size_t k = (size_t)rngBin * n + i;
                aux[2 * k] += ar * c - ai * sn;
                aux[2 * k + 1] += ar * sn + ai * c;
            }
        }
    }
    free(r4);
    int L = or_chirp(p->f0, p->f1, p->T, p->fs_chirp, NULL);
    double* chirp = (double*)malloc(2 * (size_t)L * sizeof(double));
    or_chirp(p->f0, p->f1, p->T, p->fs_chirp, chirp);
    double* refsig = (double*)malloc(2 * (size_t)S * sizeof(double));
    reference_signal(chirp, L < S ? L : S, S, refsig);
    double* col = (double*)malloc(2 * (size_t)S * sizeof(double));
    double* res = (double*)malloc(2 * (size_t)S * sizeof(double));
    double* raw = (double*)malloc(2 * SN * sizeof(double));
    for (int i = 0; i < n; i++) {                                       /* :406-421 */
        for (int r = 0; r < S; r++) {
            col[2 * r] = aux[2 * ((size_t)r * n + i)];
            col[2 * r + 1] = aux[2 * ((size_t)r * n + i) + 1];
        }
        or_fft_convolution(col, S, refsig, S, res);
        for (int r = 0; r < S; r++) {
            raw[2 * ((size_t)r * n + i)] = res[2 * r];
            raw[2 * ((size_t)r * n + i) + 1] = res[2 * r + 1];
        }
    }
    if (root_out) {
        for (int i = 0; i < n; i++) {                                   /* :431-446 */
            for (int r = 0; r < S; r++) {
                col[2 * r] = raw[2 * ((size_t)r * n + i)];
                col[2 * r + 1] = raw[2 * ((size_t)r * n + i) + 1];
            }
            or_fft_correlation(col, S, refsig, S, res);
            for (int r = 0; r < S; r++) {
                root_out[2 * ((size_t)r * n + i)] = res[2 * r];
                root_out[2 * ((size_t)r * n + i) + 1] = res[2 * r + 1];
            }
        }
    }
    if (aux_out) memcpy(aux_out, aux, 2 * SN * sizeof(double));
    if (raw_out) memcpy(raw_out, raw, 2 * SN * sizeof(double));
    free(aux); free(chirp); free(refsig); free(col); free(res); free(raw);
}

void or_refraction_batch(const double* legs, size_t m, const or_dem* dem, double n1, double n2,
                         double* out, int threads)
{
    if (threads > 0) omp_set_num_threads(threads);
    #pragma omp parallel for schedule(dynamic, 64)
    for (long long k = 0; k < (long long)m; k++) {
        const double* s = legs + 6 * (size_t)k;
        or_refraction_point(s[0], s[1], s[2], s[3], s[4], s[5], dem, n1, n2, out + 3 * (size_t)k);
    }
}

/* proc_test_script.cpp:326-350: pulse p of rootData is flat[r + p*S] */
void or_upsample(const double* root, size_t S, size_t N, int factor, double* out)
{
    size_t S4 = S * (size_t)factor;
    #pragma omp parallel for schedule(static)
    for (long long p = 0; p < (long long)N; p++)
        or_spline(root + 2 * (size_t)p * S, S, factor, out + 2 * (size_t)p * S4);
}

/* Per-voxel body of proc_test_script.cpp:464-668 */
void or_backproject_voxel(const or_bp_params* p, const or_dem* dem,
                          const double* tx, const double* rx, int n,
                          const double* rcomp, size_t S4, size_t ndata,
                          const double P[3], int reflection, double out[2])
{
    double* q = (double*)malloc(6 * (size_t)n * sizeof(double));
    double* ida = q;
    double* volta = q + 3 * (size_t)n;
    double* r4 = reflection ? (double*)malloc(4 * (size_t)n * sizeof(double)) : NULL;
    or_slant_range_fermat(dem, tx, rx, n, P, p->n1, p->n2, ida, volta, r4);  /* :484-485 */
    if (reflection) {                                                   /* rdp:188-193 */
        double rp[3];
        for (int i = 0; i < n; i++)
            or_reflection_point(tx[3 * i], tx[3 * i + 1], tx[3 * i + 2],
                                rx[3 * i], rx[3 * i + 1], rx[3 * i + 2], dem, rp);
        free(r4);
    }
    double accr = 0.0, acci = 0.0;                                      /* :488 */
    double n1 = p->n1, n2 = p->n2;
    for (int i = 0; i < n; i++) {
        const double* qa = ida + 3 * (size_t)i;
        const double* qb = volta + 3 * (size_t)i;
        double R2t = dist3(P[0], P[1], P[2], qa[0], qa[1], qa[2]);    /* :496-505 */
        double R2r = dist3(P[0], P[1], P[2], qb[0], qb[1], qb[2]);    /* :507-516 */
        double st = 0.0, sr = 0.0;
        for (int j = 0; j < n; j++)                                     /* :518-532 */
            st += dist3(tx[3 * j], tx[3 * j + 1], tx[3 * j + 2], qa[0], qa[1], qa[2]);
        double R1t = st / (double)n;
        for (int j = 0; j < n; j++)                                     /* :534-548 */
            sr += dist3(rx[3 * j], rx[3 * j + 1], rx[3 * j + 2], qb[0], qb[1], qb[2]);
        double R1r = sr / (double)n;
        double t_val = (1.0 / p->c) * (n1 * R1t + n2 * R2t + n2 * R2r + n1 * R1r);  /* :567 */
        double rngBin = 1.0 + (t_val * p->fs4);                                      /* :571 */
        double phi = (2.0 * M_PI / p->lambda) * (n1 * R1t + n2 * R2t + n2 * R2r + n1 * R1r); /* :575 */
        double r = rngBin - 1.0;                                        /* :593 */
        /* :595-599 -- floor(r) >= 0 && ceil(r) < S4, with the x86 result of
         * an out-of-range int conversion (INT_MIN) folded in */
        if (!(r >= 0.0 && r <= (double)(S4 - 1))) continue;
        int idxLast = (int)floor(r);
        int idxNext = (int)ceil(r);
        double qf = r - idxLast;                                        /* :602 */
        size_t idxAz = (size_t)i < ndata - 1 ? (size_t)i : ndata - 1;   /* :607 */
        const double* dL = rcomp + 2 * ((size_t)idxLast + idxAz * S4);  /* :614-615 */
        const double* dN = rcomp + 2 * ((size_t)idxNext + idxAz * S4);
        double vr, vi;
        if (idxLast == idxNext) { vr = dL[0]; vi = dL[1]; }            /* :626-628 */
        else {                                                          /* :632-634 */
            vr = (1.0 - qf) * dL[0] + qf * dN[0];
            vi = (1.0 - qf) * dL[1] + qf * dN[1];
        }
        double c = cos(phi), s = sin(phi);                              /* :638 */
        accr += vr * c - vi * s;                                        /* :641 */
        acci += vr * s + vi * c;
    }
    free(q);
    out[0] = accr;
    out[1] = acci;
}

/* proc_test_script.cpp:437-798: voxel list in z -> x -> y order, OpenMP
 * dynamic schedule over voxels, output[x][y][z] written in the x-fastest
 * export order of data_export.cpp:241-256. */
void or_backproject(const or_bp_params* p, const or_dem* dem,
                    const double* tx, const double* rx, int n,
                    const double* rcomp, size_t S4, size_t ndata,
                    const double* gx, int nx, const double* gy, int ny, const double* gz, int nz,
                    double* out, const int64_t* subset, int64_t nsubset,
                    int reflection, int threads)
{
    int64_t total = subset ? nsubset : (int64_t)nx * ny * nz;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < total; k++) {
        int64_t coord_idx = subset ? subset[k] : k;
        int64_t yi = coord_idx % ny;                                    /* :761-763 */
        int64_t xi = (coord_idx / ny) % nx;
        int64_t zi = coord_idx / ((int64_t)nx * ny);
        double P[3] = { gx[xi], gy[yi], gz[zi] };                       /* :144-149 */
        double v[2];
        or_backproject_voxel(p, dem, tx, rx, n, rcomp, S4, ndata, P, reflection, v);
        size_t o = (size_t)((zi * ny + yi) * nx + xi);
        out[2 * o] = v[0];
        out[2 * o + 1] = v[1];
    }
}

double or_flops_per_voxel_pulse(int n)
{
    return 65406.0 + 20.0 * (double)n;
}

/* The FP64 back-projection's expanded squared antenna distances
 * (r1_mean_fp64 in c-_bistatic_sar_amd/csrc/bsar_bp.h, terms built by
 * bsar_dataset_create in csrc/api.cpp), run on the CPU with the same
 * operations: b = A - o and c = |b|^2 on the host, then per point q
 * u = q - o, Q = fma(uz, uz, fma(uy, uy, ux * ux)) and
 * a = fma(bx, -2ux, fma(by, -2uy, fma(bz, -2uz, c + Q))), used only when
 * bsq + Q <= 32 gap^2 (gap = q's vertical distance from the antennas' height
 * range).  Reports the worst relative error of a against the exact squared
 * distance of the original A and q (long double, 64-bit significand) in
 * units of 2^-53, over the (antenna, point) pairs that pass the check.
 *
 *   usage: r1_expansion_model FAMILY N [nocheck]
 *     -> "worst_eps <w> pairs <used> of <total>"
 *
 * Families: 0 the workloads' geometry (spirals of radius 5-15 m at 12 m,
 * points a few metres below ground up to 40 m away), 1 wide apertures (up to
 * 2 km across, heights 10 m - 1 km), 2 the same at UTM-sized coordinates
 * (offsets of 5e5 / 4e6 m).  nocheck = 1 expands every pair regardless of
 * the check (the check's own sanity test: family 2 must then fail).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t next(void)
{
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}
static double u01(void) { return (double)(next() >> 11) * 0x1p-53; }
static double uni(double a, double b) { return a + (b - a) * u01(); }

enum { NA = 64 };

int main(int argc, char** argv)
{
    const int fam = argc > 1 ? atoi(argv[1]) : 0;
    const long n = argc > 2 ? atol(argv[2]) : 100000;
    const int nocheck = argc > 3 ? atoi(argv[3]) : 0;
    double worst = 0.0;
    long used = 0, total = 0;
    for (long rep = 0; rep < n / NA + 1; ++rep) {
        /* one antenna set and one point per repetition */
        double A[NA][3];
        double off[3] = {0.0, 0.0, 0.0};
        double R = 15.0, h = 12.0, hs = 0.0;
        if (fam >= 1) { R = uni(10.0, 1000.0); h = uni(10.0, 1000.0); hs = uni(0.0, 0.3) * h; }
        if (fam == 2) { off[0] = 5e5 + uni(-1e4, 1e4); off[1] = 4e6 + uni(-1e4, 1e4); off[2] = uni(0.0, 3e3); }
        for (int j = 0; j < NA; ++j) {
            const double r = fam == 0 ? uni(5.0, 15.0) : uni(0.0, R), t = uni(0.0, 6.283185307179586);
            A[j][0] = off[0] + r * cos(t);
            A[j][1] = off[1] + r * sin(t);
            A[j][2] = off[2] + h + uni(-hs, hs);
        }
        /* host terms (bsar_dataset_create): o = first antenna pair's midpoint */
        double o[3], B[NA][4], bsq = 0.0, zlo = INFINITY, zhi = -INFINITY;
        for (int k = 0; k < 3; ++k) o[k] = (A[0][k] + A[1][k]) / 2.0;
        for (int j = 0; j < NA; ++j) {
            double cc = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double b = A[j][k] - o[k];
                B[j][k] = b;
                cc += b * b;
            }
            B[j][3] = cc;
            if (cc > bsq) bsq = cc;
            if (A[j][2] < zlo) zlo = A[j][2];
            if (A[j][2] > zhi) zhi = A[j][2];
        }
        bsq *= 1.0 + 1e-12;
        /* a point below the antennas: a refraction point near the surface */
        const double span = fam == 0 ? 40.0 : 2.0 * R;
        const double q[3] = {off[0] + uni(-span, span), off[1] + uni(-span, span), off[2] + uni(-8.0, 2.0)};
        const double ux = q[0] - o[0], uy = q[1] - o[1], uz = q[2] - o[2];
        const double Q = fma(uz, uz, fma(uy, uy, ux * ux));
        const double gap = fmax(fmax(zlo - q[2], q[2] - zhi), 0.0);
        total += NA;
        if (!nocheck && !(bsq + Q <= 32.0 * (gap * gap))) continue;
        const double mx = -2.0 * ux, my = -2.0 * uy, mz = -2.0 * uz;
        for (int j = 0; j < NA; ++j) {
            const double a = fma(B[j][0], mx, fma(B[j][1], my, fma(B[j][2], mz, B[j][3] + Q)));
            long double ex = 0.0L;
            for (int k = 0; k < 3; ++k) {
                const long double f = (long double)A[j][k] - (long double)q[k];
                ex += f * f;
            }
            const double rel = (double)(fabsl((long double)a - ex) / ex) / 0x1p-53;
            if (rel > worst || rel != rel) worst = rel != rel ? INFINITY : rel;
            ++used;
        }
    }
    printf("worst_eps %.3f pairs %ld of %ld\n", worst, used, total);
    return 0;
}

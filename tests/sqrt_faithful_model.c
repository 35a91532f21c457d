/* The FP64 back-projection's square root of the R1 sums (sqrt_faithful in
 * c-_bistatic_sar_amd/csrc/bsar_bp.h), run on the CPU with its v_rsq_f64
 * seed modelled as the exact reciprocal square root times (1 + e) for any
 * relative error |e| <= 2^-21 (AMD documents the instruction at 2^29 ulp,
 * i.e. 2^-23).  Reports the largest error against the exact root (long
 * double, 64-bit significand) in ulps of the correctly rounded root.
 *
 *   usage: sqrt_faithful_model N [drop]  -> "worst_ulp <w> over <n> <violations>"
 *
 * drop = 1 leaves out the final Newton step, which the result must then
 * visibly miss (the check's own sanity test).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t st = 0x243F6A8885A308D3ull;
static uint64_t next(void)
{
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}
static double u01(void) { return (double)(next() >> 11) * 0x1p-53; }

static int drop = 0;

static double sqrt_model(double a, double e)
{
    const double y = (1.0 / sqrt(a)) * (1.0 + e);
    const double s0 = a * y, h = 0.5 * y;
    const double r = fma(-h, s0, 0.5);
    const double s1 = fma(s0, r, s0);
    if (drop) return s1;
    const double d = fma(-s1, s1, a);
    return fma(d, h, s1);
}

int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    drop = argc > 2 ? atoi(argv[2]) : 0;
    double worst = 0.0;
    long bad = 0;
    for (long k = 0; k < n; ++k) {
        double a;
        switch (k % 4) {
        case 0: a = ldexp(1.0 + u01(), (int)(next() % 2000) - 1000); break;   /* [2^-1000, 2^1000) */
        case 1: a = 100.0 + u01() * 1e6; break;                              /* squared antenna distances */
        case 2: a = nextafter(ldexp(1.0, (int)(next() % 200) - 100), INFINITY); break;
        default: { const double q = floor(1.0 + u01() * 1e7); a = q * q; }  /* exact squares */
        }
        double e = (2.0 * u01() - 1.0) * 0x1p-21;
        if (k % 16 == 5) e = 0x1p-21;
        if (k % 16 == 6) e = -0x1p-21;
        const double got = sqrt_model(a, e), cr = sqrt(a);
        const double ulp = nextafter(cr, INFINITY) - cr;
        const double err = (double)(fabsl((long double)got - sqrtl((long double)a)) / ulp);
        if (!(err < 1.0)) ++bad;
        if (err > worst) worst = err;
    }
    printf("worst_ulp %.6f over %ld %ld\n", worst, n, bad);
    return 0;
}

// kernels_fft.hip -- batched range compression (pulse compression) on gfx950.
//
// The reference compresses every pulse column with FFTProcessor::fftCorrelation
// (fft_utils.cpp:63-99; driven at raw_data_processing.cpp:431-446): FFT of the
// column, FFT of the reference (recomputed per call), multiply by the
// conjugate, inverse FFT, keep the first S samples.  fftConvolution
// (fft_utils.cpp:103-137) is the same without the conjugate.  Here the
// reference spectrum H is computed once per call and each column is one
// workgroup: the column is staged in LDS, transformed in place (radix-2 DIT
// forward from bit-reversed load, DIF inverse to bit-reversed store, three
// stages per LDS round trip, twiddles from a precomputed table), so HBM sees
// one read and one write per sample.  16384-sample columns run in a
// persistent filter with the pass structure fixed at compile time
// (k_fft_filter_pc: complex64 with the next column prefetched during the
// inverse stages, complex128 -- twice the LDS budget -- half a column at a
// time); other columns of twice the budget are split by one global radix-2
// stage into two half-length columns filtered in LDS and merged by one global
// radix-2 stage (three passes over the data); longer ones run as log2(S)
// global Stockham passes per direction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "bsar_fft.h"

namespace bsar {

template <typename T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };

__device__ __forceinline__ void twiddle(float a, float& s, float& c) { sincospif(a, &s, &c); }
__device__ __forceinline__ void twiddle(double a, double& s, double& c) { sincospi(a, &s, &c); }

__device__ __forceinline__ uint32_t bitrev(uint32_t k, int logS) { return __brev(k) >> (32 - logS); }

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) { return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }

// exp(sgn i pi r / d) for d <= 8 (multiples of a sixteenth turn), exact for
// r = 0 and the quarter turns.
template <typename T>
__device__ __forceinline__ typename C2<T>::type rot16(int r, int d, T sgn)
{
    using V = typename C2<T>::type;
    constexpr T c1 = (T)0.92387953251128675613;   // cos(pi/8)
    constexpr T s1 = (T)0.38268343236508977173;   // sin(pi/8)
    constexpr T c2 = (T)0.70710678118654752440;   // cos(pi/4)
    const int k = r * (8 / d);                    // angle k pi / 8, k < 8
    switch (k) {
    case 0: return V{(T)1, (T)0};
    case 1: return V{c1, sgn * s1};
    case 2: return V{c2, sgn * c2};
    case 3: return V{s1, sgn * c1};
    case 4: return V{(T)0, sgn};
    case 5: return V{-s1, sgn * c1};
    case 6: return V{-c2, sgn * c2};
    default: return V{-c1, sgn * s1};
    }
}

// a * b with fused products (two roundings per component instead of three).
__device__ __forceinline__ float2 cmul_f(float2 a, float2 b)
{
    return float2{__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ double2 cmul_f(double2 a, double2 b)
{
    return double2{__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x)};
}

// exp(sgn i pi k / 16), k < 16.
template <typename T>
__device__ __forceinline__ typename C2<T>::type rot_16th(int k, T sgn)
{
    using V = typename C2<T>::type;
    constexpr T a1 = (T)0.98078528040323044913;   // cos(pi/16)
    constexpr T b1 = (T)0.19509032201612826785;   // sin(pi/16)
    constexpr T a2 = (T)0.92387953251128675613;   // cos(pi/8)
    constexpr T b2 = (T)0.38268343236508977173;   // sin(pi/8)
    constexpr T a3 = (T)0.83146961230254523708;   // cos(3 pi/16)
    constexpr T b3 = (T)0.55557023301960222474;   // sin(3 pi/16)
    switch (k) {
    case 1: return V{a1, sgn * b1};
    case 2: return V{a2, sgn * b2};
    case 3: return V{a3, sgn * b3};
    case 5: return V{b3, sgn * a3};
    case 6: return V{b2, sgn * a2};
    case 7: return V{b1, sgn * a1};
    case 9: return V{-b1, sgn * a1};
    case 10: return V{-b2, sgn * a2};
    case 11: return V{-b3, sgn * a3};
    case 13: return V{-a3, sgn * b3};
    case 14: return V{-a2, sgn * b2};
    default: return V{-a1, sgn * b1};   // 15
    }
}

// a * exp(sgn i pi r / d) for compile-time r < d <= 16: the quarter turns as
// swaps and signs (exact), the odd eighths as a sum and a difference scaled
// by cos(pi/4), the rest as a fused product.
template <typename T>
__device__ __forceinline__ typename C2<T>::type crot_f(typename C2<T>::type a, int r, int d, T sgn)
{
    using V = typename C2<T>::type;
    constexpr T c4 = (T)0.70710678118654752440;
    const int k = r * (16 / d);                   // angle k pi / 16
    if (k == 0) return a;
    if (k == 8) return V{-sgn * a.y, sgn * a.x};
    if (k == 4) return V{c4 * (a.x - sgn * a.y), c4 * (a.y + sgn * a.x)};
    if (k == 12) return V{-c4 * (a.x + sgn * a.y), c4 * (sgn * a.x - a.y)};
    return cmul_f(a, rot_16th<T>(k, sgn));
}

// Twiddle sources: the table itself, or (TwLds) its two-level copy in LDS for
// S = 2^14, tw[k] = c[k / 64] * f[k mod 64] with c[a] = tw[64 a] (a < 128) at
// t[0, 128) and f[b] = tw[b] at t[128, 192) -- one more rounding.
template <typename V>
__device__ __forceinline__ V tw_get(const V* __restrict__ tw, int64_t k) { return tw[k]; }

template <typename V>
struct TwLds { const V* t; };
template <typename V>
__device__ __forceinline__ V tw_get(TwLds<V> tl, int64_t k)
{
    const uint32_t u = (uint32_t)k;
    return cmul_f(tl.t[u >> 6], tl.t[128 + (u & 63)]);
}

// K consecutive radix-2 stages on the 2^K elements v[m] = x[e0 + m h] (h =
// 2^lh, j = e0 mod h).  The stage at distance 2^t h pairs v[m], v[m + 2^t]
// with twiddle exp(sgn i pi (j + (m mod 2^t) h) / (2^t h)) = W_t *
// exp(sgn i pi (m mod 2^t) / 2^t), W_t = exp(sgn i pi j / (2^t h)) =
// tw[j S / (2^(t+1) h)] (conjugated for the inverse), tw[k] = exp(-2 pi i k / S).
// kDif: decimation in frequency (distances descending, twiddle after the
// difference: natural order in, bit-reversed out); else decimation in time
// (distances ascending, twiddle on the second input: bit-reversed in,
// natural out).
// kUnit: j = 0 (W = 1), the twiddles are the constant rotations alone.
// kFast: fused products (cmul_f) and exact quarter turns (crot_f) -- fewer
// instructions, different last bits; TWS: the twiddle source (tw_get).
template <typename T, int K, bool kDif, bool kInv, bool kUnit = false, bool kFast = false,
          typename TWS = const typename C2<T>::type*>
__device__ __forceinline__ void butterflies(typename C2<T>::type* v, int j, int lh, int logS, TWS tw)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << K;
    static_assert(K <= 4 || (kFast && K <= 5), "constant rotations: radix 16, radix 32 with kFast");
    const T sgn = kInv ? (T)1 : (T)-1;
    V wt[K];
    if (kUnit) {
#pragma unroll
        for (int t = 0; t < K; ++t) wt[t] = V{(T)1, (T)0};
    } else {
        // W_t = exp(sgn i pi j / (2^t h)) = tw[j S / (2^(t+1) h)], read from the table
#pragma unroll
        for (int t = 0; t < K; ++t) {
            V w = tw_get(tw, (int64_t)j << (logS - lh - 1 - t));
            if (kInv) w.y = -w.y;
            wt[t] = w;
        }
    }
#pragma unroll
    for (int tt = 0; tt < K; ++tt) {
        const int t = kDif ? (K - 1 - tt) : tt;
        const int d = 1 << t;
#pragma unroll
        for (int m = 0; m < G; ++m) {
            if (m & d) continue;
            const int r = m & (d - 1);
            if (kFast) {
                if (kDif) {
                    const V u = v[m], b = v[m + d];
                    v[m] = cadd(u, b);
                    const V df = csub(u, b);
                    v[m + d] = kUnit ? crot_f<T>(df, r, d, sgn) : cmul_f(df, crot_f<T>(wt[t], r, d, sgn));
                } else {
                    const V u = v[m];
                    const V b = kUnit ? crot_f<T>(v[m + d], r, d, sgn) : cmul_f(v[m + d], crot_f<T>(wt[t], r, d, sgn));
                    v[m] = cadd(u, b);
                    v[m + d] = csub(u, b);
                }
                continue;
            }
            const V tw_m = kUnit ? rot16<T>(r, d, sgn) : ((r == 0) ? wt[t] : cmul(wt[t], rot16<T>(r, d, sgn)));
            const bool trivial = kUnit && r == 0;
            if (kDif) {
                const V u = v[m], b = v[m + d];
                v[m] = cadd(u, b);
                v[m + d] = trivial ? csub(u, b) : cmul(csub(u, b), tw_m);
            } else {
                const V u = v[m], b = trivial ? v[m + d] : cmul(v[m + d], tw_m);
                v[m] = cadd(u, b);
                v[m + d] = csub(u, b);
            }
        }
    }
}

// LDS layout: element e at e + e / 32 (one pad slot per 32 keeps the strided
// accesses of every pass within 2-4 way bank conflicts).
__device__ __forceinline__ int lds_pad(int e) { return e + (e >> 5); }

// One work item q of a K-stage LDS pass whose smallest distance is 2^lh.
template <typename T, int K, bool kDif, bool kInv>
__device__ __forceinline__ void lds_group(typename C2<T>::type* x, int lh, int q, int logS,
                                          const typename C2<T>::type* __restrict__ tw)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << K;
    const int h = 1 << lh;
    const int j = q & (h - 1);
    const int e0 = ((q >> lh) << (lh + K)) + j;
    V v[G];
#pragma unroll
    for (int m = 0; m < G; ++m) v[m] = x[lds_pad(e0 + m * h)];
    butterflies<T, K, kDif, kInv>(v, j, lh, logS, tw);
#pragma unroll
    for (int m = 0; m < G; ++m) x[lds_pad(e0 + m * h)] = v[m];
}

// The innermost Kb stages (distances 2^(Kb-1) .. 1) of the forward DIF, the
// product with the spectrum and the same stages of the inverse DIT in one LDS
// pass: the two transforms meet on the same 2^Kb contiguous elements.
template <typename T, int Kb>
__device__ __forceinline__ void lds_turn(typename C2<T>::type* x, int q, int logS,
                                         const typename C2<T>::type* __restrict__ Hbr,
                                         const typename C2<T>::type* __restrict__ tw)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << Kb;
    const int e0 = q << Kb;
    V v[G];
#pragma unroll
    for (int m = 0; m < G; ++m) v[m] = x[lds_pad(e0 + m)];
    butterflies<T, Kb, true, false, true>(v, 0, 0, logS, tw);
#pragma unroll
    for (int m = 0; m < G; ++m) v[m] = cmul(v[m], Hbr[e0 + m]);
    butterflies<T, Kb, false, true, true>(v, 0, 0, logS, tw);
#pragma unroll
    for (int m = 0; m < G; ++m) x[lds_pad(e0 + m)] = v[m];
}

// The stages below distance 2^top in LDS (top = log2 of the register pass's
// smallest distance, >= 1): forward DIF in radix-2^R passes from the top, the
// innermost Kb (top - Kb a multiple of R, 1 <= Kb <= R) stages fused with the
// product and the inverse's innermost stages, then the inverse DIT in
// radix-2^R passes outwards.  x holds 2^logN elements: whole blocks of 2^top,
// which the LDS stages never mix, so any run of blocks of the column; logS is
// the column's length (it sets the twiddles) and Hbr the spectrum from x's
// first element on.
template <typename T, int Kb, int R = 4>
__device__ __forceinline__ void lds_filter(typename C2<T>::type* x, int top, int logN, int logS,
                                           const typename C2<T>::type* __restrict__ Hbr,
                                           const typename C2<T>::type* __restrict__ tw)
{
    const int N = 1 << logN;
    for (int s = top; s > Kb; s -= R) {
#pragma unroll 1
        for (int q = threadIdx.x; q < (N >> R); q += blockDim.x) lds_group<T, R, true, false>(x, s - R, q, logS, tw);
        __syncthreads();
    }
#pragma unroll 1
    for (int q = threadIdx.x; q < (N >> Kb); q += blockDim.x) lds_turn<T, Kb>(x, q, logS, Hbr, tw);
    __syncthreads();
    for (int s = Kb; s < top; s += R) {
#pragma unroll 1
        for (int q = threadIdx.x; q < (N >> R); q += blockDim.x) lds_group<T, R, false, true>(x, s, q, logS, tw);
        __syncthreads();
    }
}

// Threads per block at most: 16384 complex64 or 8192 complex128 samples (the
// LDS budget) over 16 samples per thread.
template <typename T>
constexpr int fft_max_threads() { return (int)(kFftLdsMax / sizeof(typename C2<T>::type) / 16); }

// Fused filter: out[c] = IFFT(FFT(in[c]) .* H) for one column c per block of
// S / 2^KP threads.  Thread t owns samples t + i S/2^KP (i < 2^KP): it loads
// them coalesced, runs the KP largest-distance DIF stages in registers and
// stores them to LDS; the remaining stages, the product with the spectrum in
// bit-reversed order (Hbr = H permuted) and the inverse's matching stages run
// in LDS (lds_filter: radix-2^R passes, Kb innermost stages fused); the thread
// then reads its samples back, runs the KP largest-distance DIT stages of the
// inverse in registers and stores the samples / S coalesced.
// tw[k] = exp(-2 pi i k / S), k < S / 2.
template <typename T, int KP, int Kb>
__global__ __launch_bounds__(fft_max_threads<T>()) void k_fft_filter_lds(const typename C2<T>::type* __restrict__ in,
                                                                        typename C2<T>::type* __restrict__ out,
                                                                        const typename C2<T>::type* __restrict__ Hbr,
                                                                        const typename C2<T>::type* __restrict__ tw,
                                                                        int logS, int halves)
{
    using V = typename C2<T>::type;
    constexpr int P = 1 << KP;
    extern __shared__ unsigned char smem[];
    V* x = reinterpret_cast<V*>(smem);
    const int S = 1 << logS;
    const int nt = S >> KP;                 // == blockDim.x
    const int t = threadIdx.x;
    const int top = logS - KP;              // register stages have distances 2^top .. S/2
    const int64_t col = blockIdx.x;
    // halves: columns alternate between the two spectra Hbr[0, S) and Hbr[S, 2S)
    const V* Hc = Hbr + (halves ? (int64_t)(col & 1) * S : 0);
    V v[P];
#pragma unroll
    for (int i = 0; i < P; ++i) v[i] = in[col * S + t + i * nt];
    butterflies<T, KP, true, false>(v, t, top, logS, tw);
    if (top == 0) {
        // S = 2^KP: the registers hold the whole spectrum (bit-reversed)
#pragma unroll
        for (int i = 0; i < P; ++i) v[i] = cmul(v[i], Hc[i]);
    } else {
#pragma unroll
        for (int i = 0; i < P; ++i) x[lds_pad(t + i * nt)] = v[i];
        __syncthreads();
        lds_filter<T, Kb>(x, top, logS, logS, Hc, tw);
#pragma unroll
        for (int i = 0; i < P; ++i) v[i] = x[lds_pad(t + i * nt)];
    }
    butterflies<T, KP, false, true>(v, t, top, logS, tw);
    const T inv = (T)1 / (T)S;
#pragma unroll
    for (int i = 0; i < P; ++i) out[col * S + t + i * nt] = V{v[i].x * inv, v[i].y * inv};
}

// ---- persistent filter for 16384-sample columns --------------------------
//
// The transform of k_fft_filter_lds with the column length, the thread count
// and every pass fixed at compile time (LDS offsets and twiddle shifts are
// immediates), fused products and exact quarter turns (butterflies' kFast),
// and one workgroup per CU looping over columns.  Twiddles come from the
// two-level table in LDS and LDS exchanges use lds_barrier, so the only
// global accesses are the samples and the spectrum: with PF the next
// column's samples load into registers while the inverse stages run.

// Barrier for LDS exchanges: waits for the wave's LDS operations but not its
// global loads and stores (__syncthreads waits for those as well).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A work item's elements e0 + m h (h = 2^lh, e0 = a multiple of 2^(lh + K)
// plus j < h) sit at lds_pad(e0) + m h + (m h) / 32: below h = 32 the item's
// elements either stay inside one 32-element block or, from e0 mod 32 = j <
// h, never carry into the next block's pad; from h = 32 on each step m h is
// whole blocks.
__device__ __forceinline__ int lds_step(int m, int lh) { return (m << lh) + ((m << lh) >> 5); }

template <typename T, int K, bool kDif, bool kInv, int LH, int LOGS, int LOGN, int NT, typename TWS>
__device__ __forceinline__ void lds_pass_ct(typename C2<T>::type* x, TWS tw)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << K, h = 1 << LH, items = (1 << LOGN) >> K;
    static_assert(items % NT == 0, "whole items per thread");
#pragma unroll 1
    for (int q0 = 0; q0 < items; q0 += NT) {
        const int q = q0 + (int)threadIdx.x;
        const int j = q & (h - 1);
        V* xb = x + lds_pad(((q >> LH) << (LH + K)) + j);
        V v[G];
#pragma unroll
        for (int m = 0; m < G; ++m) v[m] = xb[lds_step(m, LH)];
        butterflies<T, K, kDif, kInv, false, true, TWS>(v, j, LH, LOGS, tw);
#pragma unroll
        for (int m = 0; m < G; ++m) xb[lds_step(m, LH)] = v[m];
    }
}

// Forward DIF passes from distance 2^S_ down to 2^Kb (radix 2^R).
template <typename T, int Kb, int R, int S_, int LOGS, int LOGN, int NT, typename TWS>
__device__ __forceinline__ void lds_fwd_ct(typename C2<T>::type* x, TWS tw)
{
    if constexpr (S_ > Kb) {
        lds_pass_ct<T, R, true, false, S_ - R, LOGS, LOGN, NT, TWS>(x, tw);
        lds_barrier();
        lds_fwd_ct<T, Kb, R, S_ - R, LOGS, LOGN, NT, TWS>(x, tw);
    }
}

// Inverse DIT passes from distance 2^S_ up to 2^TOP.
template <typename T, int R, int S_, int TOP, int LOGS, int LOGN, int NT, typename TWS>
__device__ __forceinline__ void lds_inv_ct(typename C2<T>::type* x, TWS tw)
{
    if constexpr (S_ < TOP) {
        lds_pass_ct<T, R, false, true, S_, LOGS, LOGN, NT, TWS>(x, tw);
        lds_barrier();
        lds_inv_ct<T, R, S_ + R, TOP, LOGS, LOGN, NT, TWS>(x, tw);
    }
}

// The innermost Kb stages of both transforms and the product with the
// spectrum (lds_turn with fused products).
template <typename T, int Kb, int LOGS, int LOGN, int NT>
__device__ __forceinline__ void lds_turn_ct(typename C2<T>::type* x, const typename C2<T>::type* __restrict__ Hbr)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << Kb, items = (1 << LOGN) >> Kb;
    static_assert(items % NT == 0, "whole items per thread");
    const V* const none = nullptr;   // the constant rotations need no table
#pragma unroll 1
    for (int q0 = 0; q0 < items; q0 += NT) {
        const int e0 = (q0 + (int)threadIdx.x) << Kb;
        V* xb = x + lds_pad(e0);
        V v[G];
#pragma unroll
        for (int m = 0; m < G; ++m) v[m] = xb[lds_step(m, 0)];
        butterflies<T, Kb, true, false, true, true>(v, 0, 0, LOGS, none);
#pragma unroll
        for (int m = 0; m < G; ++m) v[m] = cmul_f(v[m], Hbr[e0 + m]);
        butterflies<T, Kb, false, true, true, true>(v, 0, 0, LOGS, none);
#pragma unroll
        for (int m = 0; m < G; ++m) xb[lds_step(m, 0)] = v[m];
    }
}

// Columns of 2^LS samples, 2^(LS - KP) threads, KP register stages, radix-2^R
// LDS passes, Kb turn stages; NH LDS phases (as k_fft_filter_lds); PF: load
// the next column during the inverse stages (one whole-column phase only).
template <typename T, int LS, int KP, int Kb, int R, int NH, bool PF>
__global__ __launch_bounds__(1 << (LS - KP), 1) void k_fft_filter_pc(const typename C2<T>::type* __restrict__ in,
                                                                    typename C2<T>::type* __restrict__ out,
                                                                    const typename C2<T>::type* __restrict__ Hbr,
                                                                    const typename C2<T>::type* __restrict__ tw,
                                                                    int64_t ncols)
{
    using V = typename C2<T>::type;
    constexpr int S = 1 << LS, NT = S >> KP, P = 1 << KP, PH = P / NH, TOP = LS - KP;
    constexpr int LNH = NH == 2 ? 1 : 0, LOGN = LS - LNH;
    static_assert(LS == 14, "the two-level LDS twiddle table is sized for 2^14");
    static_assert((TOP - Kb) % R == 0 && Kb >= 1 && Kb <= R && R <= 5, "pass structure");
    static_assert(NH == 1 || (NH == 2 && !PF), "a prefetch needs the whole column in one phase");
    extern __shared__ unsigned char smem[];
    V* x = reinterpret_cast<V*>(smem);
    V* lt = x + (S / NH + S / NH / 32 + 1);
    const int t = threadIdx.x;
    for (int k = t; k < 192; k += NT) lt[k] = k < 128 ? tw[k << 6] : tw[k - 128];
    const TwLds<V> ltw{lt};
    V* xt = x + lds_pad(t);   // the thread's elements t + i NT: xt[lds_step(i, TOP)]
    int64_t col = blockIdx.x;
    V v[P];
    if (PF && col < ncols) {
#pragma unroll
        for (int i = 0; i < P; ++i) v[i] = in[col * S + t + i * NT];
    }
    for (; col < ncols; col += gridDim.x) {
        if (!PF) {
#pragma unroll
            for (int i = 0; i < P; ++i) v[i] = in[col * S + t + i * NT];
        }
        lds_barrier();   // the table, and the previous column's last LDS reads
        butterflies<T, KP, true, false, false, true, TwLds<V>>(v, t, TOP, LS, ltw);
        V vn[PF ? P : 1];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            if (h) lds_barrier();
#pragma unroll
            for (int i = 0; i < PH; ++i) xt[lds_step(i, TOP)] = v[h * PH + i];
            lds_barrier();
            lds_fwd_ct<T, Kb, R, TOP, LS, LOGN, NT, TwLds<V>>(x, ltw);
            lds_turn_ct<T, Kb, LS, LOGN, NT>(x, Hbr + h * (S >> LNH));
            lds_barrier();
            if constexpr (PF) {
                const int64_t nc = col + gridDim.x;
                if (nc < ncols) {
#pragma unroll
                    for (int i = 0; i < P; ++i) vn[i] = in[nc * S + t + i * NT];
                }
            }
            lds_inv_ct<T, R, Kb, TOP, LS, LOGN, NT, TwLds<V>>(x, ltw);
#pragma unroll
            for (int i = 0; i < PH; ++i) v[h * PH + i] = xt[lds_step(i, TOP)];
        }
        butterflies<T, KP, false, true, false, true, TwLds<V>>(v, t, TOP, LS, ltw);
        const T inv = (T)1 / (T)S;
        if constexpr (PF) {
            // streaming (non-temporal) stores: the result is not read again
            // here, and keeping it out of L2 measured 2 % faster for complex64
            // columns (no change for complex128)
            typedef T NV __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const NV w = {v[i].x * inv, v[i].y * inv};
                __builtin_nontemporal_store(w, reinterpret_cast<NV*>(&out[col * S + t + i * NT]));
            }
        } else {
#pragma unroll
            for (int i = 0; i < P; ++i) out[col * S + t + i * NT] = V{v[i].x * inv, v[i].y * inv};
        }
        if constexpr (PF) {
#pragma unroll
            for (int i = 0; i < P; ++i) v[i] = vn[i];
        }
    }
}

// Columns of length S = 2n above the LDS budget run as two independent
// length-n filters: the first radix-2 DIF stage splits X = DFT_S(x) into its
// even bins (DFT_n of x[k] + x[k+n]) and odd bins (DFT_n of (x[k] - x[k+n])
// W^k, W = exp(-2 pi i / S)); each half is filtered by its half of the
// spectrum in the LDS kernel (which scales by 1/n); the last radix-2 DIT
// stage of the inverse recombines them, y[k] = (a[k] + b[k] W^-k) / 2 and
// y[k+n] = (a[k] - b[k] W^-k) / 2.  tw is the table of length-S twiddles.
template <typename T>
__global__ void k_split_dif(typename C2<T>::type* __restrict__ x, const typename C2<T>::type* __restrict__ tw,
                            int n, int64_t ncols)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncols * n) return;
    const int64_t col = t / n;
    const int k = (int)(t - col * n);
    auto* c = x + col * 2 * n;
    const auto u = c[k], v = c[k + n];
    c[k] = cadd(u, v);
    c[k + n] = cmul(csub(u, v), tw[k]);
}

template <typename T>
__global__ void k_split_dit(typename C2<T>::type* __restrict__ x, const typename C2<T>::type* __restrict__ tw,
                            int n, int64_t ncols)
{
    using V = typename C2<T>::type;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncols * n) return;
    const int64_t col = t / n;
    const int k = (int)(t - col * n);
    V* c = x + col * 2 * n;
    const V w = tw[k];
    const V a = c[k], b = cmul(c[k + n], V{w.x, -w.y});
    const T h = (T)0.5;
    c[k] = V{(a.x + b.x) * h, (a.y + b.y) * h};
    c[k + n] = V{(a.x - b.x) * h, (a.y - b.y) * h};
}

// Hbr2 = [H[2 bitrev(p)] | H[2 bitrev(p) + 1]], p < n = 2^logn: the two
// half-spectra, each in the bit-reversed order of a length-n transform.
template <typename T>
__global__ void k_split_spectrum(const typename C2<T>::type* __restrict__ H, typename C2<T>::type* __restrict__ Hbr2,
                                 int logn)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = 1 << logn;
    if (k >= 2 * n) return;
    const int h = k >> logn, p = k & (n - 1);
    Hbr2[k] = H[2 * (int)bitrev((uint32_t)p, logn) + h];
}

// Hbr[k] = H[bitrev(k)].
template <typename T>
__global__ void k_bitrev_copy(const typename C2<T>::type* __restrict__ H, typename C2<T>::type* __restrict__ Hbr,
                              int logS)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < (1 << logS)) Hbr[k] = H[bitrev((uint32_t)k, logS)];
}

// tw[k] = exp(-2 pi i k / S), k < S / 2, from FP64 sincospi.
template <typename T>
__global__ void k_twiddles(typename C2<T>::type* tw, int S)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S / 2) return;
    double sn, cs;
    sincospi(-2.0 * (double)k / (double)S, &sn, &cs);
    tw[k] = typename C2<T>::type{(T)cs, (T)sn};
}

// One radix-2 Stockham pass over all columns (global memory): stage with
// sub-transform length L = 2^s, sign -1 forward / +1 inverse.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_pass(const typename C2<T>::type* __restrict__ a,
                                                  typename C2<T>::type* __restrict__ b, int logS, int s, int sign,
                                                  int64_t ncols)
{
    using V = typename C2<T>::type;
    const int S = 1 << logS;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncols * (S / 2)) return;
    const int64_t col = t / (S / 2);
    const int k = (int)(t - col * (S / 2));
    // Stockham autosort: m = 2^(s-1) butterflies per group, S/(2m) groups.
    const int m = 1 << (s - 1);
    const int j = k & (m - 1);          // position inside the group
    const int g = k >> (s - 1);         // group
    const V* A = a + col * S;
    V* B = b + col * S;
    const V u = A[g * m + j];
    const V w = A[g * m + j + S / 2];
    T sn, cs;
    twiddle((T)(sign * (T)j / (T)m), sn, cs);
    const V v = {w.x * cs - w.y * sn, w.x * sn + w.y * cs};
    B[g * 2 * m + j] = V{u.x + v.x, u.y + v.y};
    B[g * 2 * m + j + m] = V{u.x - v.x, u.y - v.y};
}

template <typename T>
__global__ void k_spec_mul(typename C2<T>::type* __restrict__ a, const typename C2<T>::type* __restrict__ H,
                           int S, int64_t ncols, T scale)
{
    using V = typename C2<T>::type;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncols * S) return;
    const int k = (int)(t % S);
    const V x = a[t], h = H[k];
    a[t] = V{(x.x * h.x - x.y * h.y) * scale, (x.x * h.y + x.y * h.x) * scale};
}

template <typename T>
__global__ void k_scale(typename C2<T>::type* __restrict__ a, int64_t n, T scale)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) { a[t].x *= scale; a[t].y *= scale; }
}

template <typename T>
__global__ void k_conj(typename C2<T>::type* __restrict__ a, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) a[t].y = -a[t].y;
}

__global__ void k_d2f(const double2* __restrict__ a, float2* __restrict__ b, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) b[t] = make_float2((float)a[t].x, (float)a[t].y);
}

__global__ void k_f2d(const float2* __restrict__ a, double2* __restrict__ b, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) b[t] = make_double2((double)a[t].x, (double)a[t].y);
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static int ilog2(int S)
{
    int l = 0;
    while ((1 << l) < S) ++l;
    return l;
}

template <typename T>
static hipError_t fft_global(typename C2<T>::type* a, typename C2<T>::type* tmp, int S, int64_t ncols, int sign,
                             hipStream_t st, typename C2<T>::type** result)
{
    using V = typename C2<T>::type;
    const int logS = ilog2(S);
    V* src = a;
    V* dst = tmp;
    for (int s = 1; s <= logS; ++s) {
        hipLaunchKernelGGL(k_fft_pass<T>, dim3(nblk(ncols * (S / 2), 256)), dim3(256), 0, st, src, dst, logS, s,
                           sign, ncols);
        V* t = src;
        src = dst;
        dst = t;
    }
    *result = src;
    return hipGetLastError();
}

// LDS of one column with one pad slot per 32 elements
template <typename T>
static size_t lds_bytes(int S) { return (size_t)(S + S / 32 + 1) * sizeof(typename C2<T>::type); }

// The dynamic-LDS limit of a kernel, raised once per (kernel, device) to the
// largest size asked so far, and the device's CU count, read once per device:
// both are fixed for the life of the process, and the runtime queries cost
// microseconds of host time on every launch otherwise.
static hipError_t ensure_dyn_lds(const void* fn, size_t lb)
{
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, size_t> limits;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = limits[{fn, dev}];
    if (lb <= have) return hipSuccess;
    if ((e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb)) != hipSuccess) return e;
    have = lb;
    return hipSuccess;
}

static hipError_t device_cus(int* cus)
{
    static std::mutex mu;
    static std::map<int, int> counts;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto it = counts.find(dev);
    if (it == counts.end()) {
        int n = 0;
        if ((e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        it = counts.emplace(dev, std::max(n, 1)).first;
    }
    *cus = it->second;
    return hipSuccess;
}

template <typename T, int KP, int Kb>
static hipError_t launch_lds_filter(typename C2<T>::type* data, const typename C2<T>::type* Hbr,
                                    const typename C2<T>::type* tw, int logS, int64_t ncols, hipStream_t st,
                                    int halves = 0)
{
    const int S = 1 << logS;
    if ((S >> KP) > fft_max_threads<T>()) return hipErrorInvalidValue;
    const size_t lb = lds_bytes<T>(S);
    hipError_t e = ensure_dyn_lds((const void*)k_fft_filter_lds<T, KP, Kb>, lb);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fft_filter_lds<T, KP, Kb>), dim3((unsigned)ncols), dim3(S >> KP), lb, st, data, data, Hbr,
                       tw, logS, halves);
    return hipGetLastError();
}

// One workgroup per CU of the persistent filter (the LDS holds one column).
template <typename T, int LS, int KP, int Kb, int R, int NH, bool PF>
static hipError_t launch_filter_pc(typename C2<T>::type* data, const typename C2<T>::type* Hbr,
                                   const typename C2<T>::type* tw, int64_t ncols, hipStream_t st)
{
    using V = typename C2<T>::type;
    const size_t lb = lds_bytes<T>((1 << LS) / NH) + 192 * sizeof(V);
    hipError_t e = ensure_dyn_lds((const void*)k_fft_filter_pc<T, LS, KP, Kb, R, NH, PF>, lb);
    if (e != hipSuccess) return e;
    int cus = 1;
    if ((e = device_cus(&cus)) != hipSuccess) return e;
    const unsigned grid = (unsigned)std::min<int64_t>(ncols, cus);
    hipLaunchKernelGGL((k_fft_filter_pc<T, LS, KP, Kb, R, NH, PF>), dim3(grid), dim3(1 << (LS - KP)), lb, st, data,
                       data, Hbr, tw, ncols);
    return hipGetLastError();
}

// Twiddle tables tw[k] = exp(-2 pi i k / S), k < S / 2, one per (device, S,
// precision), built on first use and kept for the life of the process.
template <typename T>
static hipError_t twiddle_table(int S, hipStream_t st, const typename C2<T>::type** out)
{
    using V = typename C2<T>::type;
    static std::mutex mu;
    static std::map<std::pair<int, int>, V*> tables;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto it = tables.find({dev, S});
    if (it != tables.end()) { *out = it->second; return hipSuccess; }
    V* tw = nullptr;
    const int n = S / 2 > 0 ? S / 2 : 1;
    if ((e = hipMalloc((void**)&tw, (size_t)n * sizeof(V))) != hipSuccess) return e;
    hipLaunchKernelGGL(k_twiddles<T>, dim3(nblk(n, 256)), dim3(256), 0, st, tw, S);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess) {
        (void)hipFree(tw);
        return e;
    }
    tables[{dev, S}] = tw;
    *out = tw;
    return hipSuccess;
}

// tw_given / twn_given: twiddle tables of S and S / 2 the caller built
// (fft_twiddles); null takes the cached per-process tables.  stage: kFilterAll
// runs everything; kFilterPrepare only the spectrum reordering into tmp (and
// the first-use twiddle tables); kFilterApply the rest, with tmp as prepared.
enum { kFilterAll = 0, kFilterPrepare = 1, kFilterApply = 2 };
template <typename T>
hipError_t fft_filter_t(typename C2<T>::type* data, typename C2<T>::type* tmp, const typename C2<T>::type* H, int S,
                        int64_t ncols, hipStream_t st, const typename C2<T>::type* tw_given = nullptr,
                        const typename C2<T>::type* twn_given = nullptr, int stage = kFilterAll)
{
    using V = typename C2<T>::type;
    if (S < 2 || (S & (S - 1)) || ncols <= 0) return hipErrorInvalidValue;
    const int logS = ilog2(S);
    // S <= 16: the register pass holds the column; else 16 samples per thread
    // and the innermost (logS - 4) mod 4 (or 4) stages fused
    auto lds = [&](int lg, const V* Hbr, const V* tw, int64_t nc, int halves) -> hipError_t {
        switch (lg) {
        case 1: return launch_lds_filter<T, 1, 1>(data, Hbr, tw, lg, nc, st, halves);
        case 2: return launch_lds_filter<T, 2, 1>(data, Hbr, tw, lg, nc, st, halves);
        case 3: return launch_lds_filter<T, 3, 1>(data, Hbr, tw, lg, nc, st, halves);
        case 4: return launch_lds_filter<T, 4, 1>(data, Hbr, tw, lg, nc, st, halves);
        default:
            switch ((lg - 4) % 4) {
            case 1: return launch_lds_filter<T, 4, 1>(data, Hbr, tw, lg, nc, st, halves);
            case 2: return launch_lds_filter<T, 4, 2>(data, Hbr, tw, lg, nc, st, halves);
            case 3: return launch_lds_filter<T, 4, 3>(data, Hbr, tw, lg, nc, st, halves);
            default: return launch_lds_filter<T, 4, 4>(data, Hbr, tw, lg, nc, st, halves);
            }
        }
    };
    // 16384-sample columns run in the persistent filter: complex64 with the
    // whole column in LDS and the next column prefetched, complex128 (256 KiB,
    // twice the LDS budget) half a column at a time with 512 threads of 32
    // samples (one pass over HBM instead of the split path's three).
    const bool pc = logS == 14;
    if (pc || (size_t)S * sizeof(V) <= kFftLdsMax) {
        // S / 2^KP <= 1024 threads: S <= 16384 (complex64) / 8192 (complex128)
        if (!tmp) return hipErrorInvalidValue;
        const V* tw = tw_given;
        hipError_t e = tw ? hipSuccess : twiddle_table<T>(S, st, &tw);
        if (e != hipSuccess) return e;
        if (stage != kFilterApply) {
            hipLaunchKernelGGL(k_bitrev_copy<T>, dim3(nblk(S, 256)), dim3(256), 0, st, H, tmp, logS);
            if ((e = hipGetLastError()) != hipSuccess || stage == kFilterPrepare) return e;
        }
        if (pc) {
            if constexpr (std::is_same<T, double>::value)
                return launch_filter_pc<T, 14, 5, 3, 3, 2, false>(data, tmp, tw, ncols, st);
            else
                return launch_filter_pc<T, 14, 4, 2, 4, 1, true>(data, tmp, tw, ncols, st);
        }
        return lds(logS, tmp, tw, ncols, 0);
    }
    if ((size_t)(S / 2) * sizeof(V) <= kFftLdsMax) {
        // two half-length LDS filters between a DIF split and a DIT merge
        if (!tmp) return hipErrorInvalidValue;
        const int n = S / 2;
        const V *tw = tw_given, *twn = twn_given;
        hipError_t e = tw ? hipSuccess : twiddle_table<T>(S, st, &tw);
        if (e == hipSuccess && !twn) e = twiddle_table<T>(n, st, &twn);
        if (e != hipSuccess) return e;
        if (stage != kFilterApply) {
            hipLaunchKernelGGL(k_split_spectrum<T>, dim3(nblk(S, 256)), dim3(256), 0, st, H, tmp, logS - 1);
            if ((e = hipGetLastError()) != hipSuccess || stage == kFilterPrepare) return e;
        }
        hipLaunchKernelGGL(k_split_dif<T>, dim3(nblk(ncols * n, 256)), dim3(256), 0, st, data, tw, n, ncols);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = lds(logS - 1, tmp, twn, 2 * ncols, 1)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_split_dit<T>, dim3(nblk(ncols * n, 256)), dim3(256), 0, st, data, tw, n, ncols);
        return hipGetLastError();
    }
    if (stage == kFilterPrepare) return hipSuccess;   // the global path reads H as it is
    V* res = nullptr;
    hipError_t e = fft_global<T>(data, tmp, S, ncols, -1, st, &res);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_spec_mul<T>, dim3(nblk(ncols * S, 256)), dim3(256), 0, st, res, H, S, ncols, (T)1);
    V* other = (res == data) ? tmp : data;
    V* res2 = nullptr;
    // inverse through conj(FFT(conj(x)))/S: conj, forward passes, conj + scale
    hipLaunchKernelGGL(k_conj<T>, dim3(nblk(ncols * S, 256)), dim3(256), 0, st, res, ncols * S);
    e = fft_global<T>(res, other, S, ncols, -1, st, &res2);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_conj<T>, dim3(nblk(ncols * S, 256)), dim3(256), 0, st, res2, ncols * S);
    hipLaunchKernelGGL(k_scale<T>, dim3(nblk(ncols * S, 256)), dim3(256), 0, st, res2, ncols * S, (T)1 / (T)S);
    if (res2 != data) {
        e = hipMemcpyAsync(data, res2, (size_t)ncols * S * sizeof(V), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t fft_filter(double2* data, double2* tmp, const double2* H, int S, int64_t ncols, hipStream_t st)
{
    return fft_filter_t<double>(data, tmp, H, S, ncols, st);
}

hipError_t fft_filter(float2* data, float2* tmp, const float2* H, int S, int64_t ncols, hipStream_t st)
{
    return fft_filter_t<float>(data, tmp, H, S, ncols, st);
}

hipError_t fft_filter_prepare(const double2* H, double2* tmp, int S, int64_t ncols, hipStream_t st,
                              const double2* tw, const double2* twn)
{
    return fft_filter_t<double>(nullptr, tmp, H, S, ncols, st, tw, twn, kFilterPrepare);
}

hipError_t fft_filter_prepare(const float2* H, float2* tmp, int S, int64_t ncols, hipStream_t st,
                              const float2* tw, const float2* twn)
{
    return fft_filter_t<float>(nullptr, tmp, H, S, ncols, st, tw, twn, kFilterPrepare);
}

hipError_t fft_filter_apply(double2* data, double2* tmp, const double2* H, int S, int64_t ncols, hipStream_t st,
                            const double2* tw, const double2* twn)
{
    return fft_filter_t<double>(data, tmp, H, S, ncols, st, tw, twn, kFilterApply);
}

hipError_t fft_filter_apply(float2* data, float2* tmp, const float2* H, int S, int64_t ncols, hipStream_t st,
                            const float2* tw, const float2* twn)
{
    return fft_filter_t<float>(data, tmp, H, S, ncols, st, tw, twn, kFilterApply);
}

template <typename T>
static hipError_t twiddles_into(typename C2<T>::type* tw, int S, hipStream_t st)
{
    const int n = S / 2 > 0 ? S / 2 : 1;
    hipLaunchKernelGGL(k_twiddles<T>, dim3(nblk(n, 256)), dim3(256), 0, st, tw, S);
    return hipGetLastError();
}

hipError_t fft_twiddles(double2* tw, int S, hipStream_t st) { return twiddles_into<double>(tw, S, st); }
hipError_t fft_twiddles(float2* tw, int S, hipStream_t st) { return twiddles_into<float>(tw, S, st); }

hipError_t fft_forward(double2* data, double2* tmp, int S, int64_t ncols, hipStream_t st, double2** result)
{
    return fft_global<double>(data, tmp, S, ncols, -1, st, result);
}

hipError_t to_f32(const double2* a, float2* b, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n, 256)), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

hipError_t to_f64(const float2* a, double2* b, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_f2d, dim3(nblk(n, 256)), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

hipError_t scale_inplace(double2* a, int64_t n, double s, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale<double>, dim3(nblk(n, 256)), dim3(256), 0, st, a, n, s);
    return hipGetLastError();
}

hipError_t conj_inplace(double2* a, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_conj<double>, dim3(nblk(n, 256)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

}  // namespace bsar

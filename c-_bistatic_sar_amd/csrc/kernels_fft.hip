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
// one read and one write per sample.  16384-sample complex128 columns (twice
// the LDS budget) go through LDS half a column at a time; other columns of
// twice the budget are split by one global radix-2 stage into two
// half-length columns filtered in LDS and merged by one global radix-2 stage
// (three passes over the data); longer ones run as log2(S) global Stockham
// passes per direction.
#include <hip/hip_runtime.h>
#include <stdint.h>

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
template <typename T, int K, bool kDif, bool kInv, bool kUnit = false>
__device__ __forceinline__ void butterflies(typename C2<T>::type* v, int j, int lh, int logS,
                                            const typename C2<T>::type* __restrict__ tw)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << K;
    const T sgn = kInv ? (T)1 : (T)-1;
    V wt[K];
    if (kUnit) {
#pragma unroll
        for (int t = 0; t < K; ++t) wt[t] = V{(T)1, (T)0};
    } else {
        // W_t = exp(sgn i pi j / (2^t h)) = tw[j S / (2^(t+1) h)], read from the table
#pragma unroll
        for (int t = 0; t < K; ++t) {
            V w = tw[(int64_t)j << (logS - lh - 1 - t)];
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
//
// The register stages leave 2^KP independent blocks of S / 2^KP samples, so
// with NH = 2 the column goes through LDS one half at a time (a thread's
// registers i < P/2, then the rest): half the LDS per workgroup.  The lean
// variants (R = 3, MINW waves per SIMD, NT threads at most) use that for two
// workgroups per CU, so one column's loads and stores overlap another's LDS
// passes, or for columns twice the whole-column LDS budget.
template <typename T, int KP, int Kb, int NH = 1, int R = 4, int MINW = 1, int NT = fft_max_threads<T>()>
__global__ __launch_bounds__(NT, MINW) void k_fft_filter_lds(const typename C2<T>::type* __restrict__ in,
                                                            typename C2<T>::type* __restrict__ out,
                                                            const typename C2<T>::type* __restrict__ Hbr,
                                                            const typename C2<T>::type* __restrict__ tw, int logS,
                                                            int halves)
{
    using V = typename C2<T>::type;
    constexpr int P = 1 << KP;
    constexpr int PH = P / NH;              // registers per LDS phase
    constexpr int LNH = NH == 2 ? 1 : 0;
    static_assert(NH == 1 || NH == 2, "one or two LDS phases");
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
        for (int h = 0; h < NH; ++h) {
            if (h) __syncthreads();         // every read of the previous phase is done
#pragma unroll
            for (int i = 0; i < PH; ++i) x[lds_pad(t + i * nt)] = v[h * PH + i];
            __syncthreads();
            lds_filter<T, Kb, R>(x, top, logS - LNH, logS, Hc + (int64_t)h * (S >> LNH), tw);
#pragma unroll
            for (int i = 0; i < PH; ++i) v[h * PH + i] = x[lds_pad(t + i * nt)];
        }
    }
    butterflies<T, KP, false, true>(v, t, top, logS, tw);
    const T inv = (T)1 / (T)S;
#pragma unroll
    for (int i = 0; i < P; ++i) out[col * S + t + i * nt] = V{v[i].x * inv, v[i].y * inv};
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

template <typename T, int KP, int Kb>
static hipError_t launch_lds_filter(typename C2<T>::type* data, const typename C2<T>::type* Hbr,
                                    const typename C2<T>::type* tw, int logS, int64_t ncols, hipStream_t st,
                                    int halves = 0)
{
    const int S = 1 << logS;
    if ((S >> KP) > fft_max_threads<T>()) return hipErrorInvalidValue;
    const size_t lb = lds_bytes<T>(S);
    hipError_t e = hipFuncSetAttribute((const void*)k_fft_filter_lds<T, KP, Kb>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fft_filter_lds<T, KP, Kb>), dim3((unsigned)ncols), dim3(S >> KP), lb, st, data, data, Hbr,
                       tw, logS, halves);
    return hipGetLastError();
}

// A lean variant of the filter: NH LDS phases, radix-2^R LDS passes, sized
// for MINW waves per SIMD, up to NT threads.
template <typename T, int KP, int Kb, int NH, int R, int MINW, int NT>
static hipError_t launch_lds_filter_lean(typename C2<T>::type* data, const typename C2<T>::type* Hbr,
                                         const typename C2<T>::type* tw, int logS, int64_t ncols, hipStream_t st,
                                         int halves)
{
    const int S = 1 << logS;
    if ((S >> KP) > NT || logS - KP < 1 || (logS - KP - Kb) % R != 0) return hipErrorInvalidValue;
    const size_t lb = lds_bytes<T>(S / NH);
    hipError_t e = hipFuncSetAttribute((const void*)k_fft_filter_lds<T, KP, Kb, NH, R, MINW, NT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fft_filter_lds<T, KP, Kb, NH, R, MINW, NT>), dim3((unsigned)ncols), dim3(S >> KP), lb, st,
                       data, data, Hbr, tw, logS, halves);
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
// (fft_twiddles); null takes the cached per-process tables.
template <typename T>
hipError_t fft_filter_t(typename C2<T>::type* data, typename C2<T>::type* tmp, const typename C2<T>::type* H, int S,
                        int64_t ncols, hipStream_t st, const typename C2<T>::type* tw_given = nullptr,
                        const typename C2<T>::type* twn_given = nullptr)
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
    // 16384 complex128 samples (256 KiB) go through LDS in two halves with
    // radix-8 passes in one 1024-thread workgroup per CU: one pass over HBM
    // instead of the split path's three (the kernel spills 224 B per lane at
    // its 128 VGPRs, radix-16 passes 472 B).
    const bool lean = std::is_same<T, double>::value && logS == 14;
    if (lean || (size_t)S * sizeof(V) <= kFftLdsMax) {
        // S / 2^KP <= 1024 threads: S <= 16384 (complex64) / 8192 (complex128)
        if (!tmp) return hipErrorInvalidValue;
        const V* tw = tw_given;
        hipError_t e = tw ? hipSuccess : twiddle_table<T>(S, st, &tw);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_bitrev_copy<T>, dim3(nblk(S, 256)), dim3(256), 0, st, H, tmp, logS);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if constexpr (std::is_same<T, double>::value) {
            if (lean) return launch_lds_filter_lean<T, 4, 1, 2, 3, 4, 1024>(data, tmp, tw, logS, ncols, st, 0);
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
        hipLaunchKernelGGL(k_split_spectrum<T>, dim3(nblk(S, 256)), dim3(256), 0, st, H, tmp, logS - 1);
        hipLaunchKernelGGL(k_split_dif<T>, dim3(nblk(ncols * n, 256)), dim3(256), 0, st, data, tw, n, ncols);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = lds(logS - 1, tmp, twn, 2 * ncols, 1)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_split_dit<T>, dim3(nblk(ncols * n, 256)), dim3(256), 0, st, data, tw, n, ncols);
        return hipGetLastError();
    }
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

hipError_t fft_filter_tw(double2* data, double2* tmp, const double2* H, int S, int64_t ncols, hipStream_t st,
                         const double2* tw, const double2* twn)
{
    if (!tw || !twn) return hipErrorInvalidValue;
    return fft_filter_t<double>(data, tmp, H, S, ncols, st, tw, twn);
}

hipError_t fft_filter_tw(float2* data, float2* tmp, const float2* H, int S, int64_t ncols, hipStream_t st,
                         const float2* tw, const float2* twn)
{
    if (!tw || !twn) return hipErrorInvalidValue;
    return fft_filter_t<float>(data, tmp, H, S, ncols, st, tw, twn);
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

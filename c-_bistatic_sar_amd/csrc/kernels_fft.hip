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
// one read and one write per sample.  Columns larger than the LDS budget
// run as log2(S) global Stockham passes per direction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
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

// Group of 2^K elements e_m = base + m h of K consecutive radix-2 DIT stages
// s .. s+K-1 (h = 2^(s-1), j = q mod h): stage s+t pairs (e_m, e_m+2^t) with
// twiddle exp(-i pi (j + (m mod 2^t) h) / (2^t h)) = W^(2^(K-1-t)) *
// exp(-i pi (m mod 2^t) / 2^t), W = exp(-i pi j / (2^(K-1) h)) = tw[j S / (2^K h)]
// where tw[k] = exp(-2 pi i k / S).  Inverse (DIF) groups run the stages in
// the opposite order with conjugate twiddles.
template <typename T, int K, bool kInv>
__device__ __forceinline__ void fft_group(typename C2<T>::type* x, int s, int q, int logS,
                                          const typename C2<T>::type* __restrict__ tw)
{
    using V = typename C2<T>::type;
    constexpr int G = 1 << K;
    // forward: h = 2^(s-1) is the smallest distance; inverse: s is the top stage, h = 2^(s-K)
    const int lh = kInv ? (s - K) : (s - 1);
    const int h = 1 << lh;
    const int j = q & (h - 1);
    const int base = ((q >> lh) << (lh + K)) + j;
    V w = tw[(int64_t)j << (logS - K - lh)];                         // exp(-i pi j / (2^(K-1) h))
    if (kInv) w.y = -w.y;
    // powers: wt[t] = W^(2^(K-1-t)), t = K-1 .. 0
    V wt[K];
    wt[K - 1] = w;
#pragma unroll
    for (int t = K - 2; t >= 0; --t) wt[t] = cmul(wt[t + 1], wt[t + 1]);
    V v[G];
#pragma unroll
    for (int m = 0; m < G; ++m) v[m] = x[base + m * h];
    const T sgn = kInv ? (T)1 : (T)-1;
#pragma unroll
    for (int tt = 0; tt < K; ++tt) {
        const int t = kInv ? (K - 1 - tt) : tt;                      // DIT: distance 2^t ascending; DIF: descending
        const int d = 1 << t;
#pragma unroll
        for (int m = 0; m < G; ++m) {
            if (m & d) continue;
            const int r = m & (d - 1);
            // exp(sgn i pi r / d), d <= 4: multiples of an eighth turn
            constexpr T c8 = (T)0.70710678118654752440;
            V rot;
            if (r == 0) rot = V{(T)1, (T)0};
            else if (2 * r == d) rot = V{(T)0, sgn};
            else if (r == 1) rot = V{c8, sgn * c8};
            else rot = V{-c8, sgn * c8};                               // r = 3, d = 4
            const V tw_m = (r == 0) ? wt[t] : cmul(wt[t], rot);
            if (!kInv) {
                const V u = v[m], b = cmul(v[m + d], tw_m);
                v[m] = cadd(u, b);
                v[m + d] = csub(u, b);
            } else {
                const V u = v[m], b = v[m + d];
                v[m] = cadd(u, b);
                v[m + d] = cmul(csub(u, b), tw_m);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < G; ++m) x[base + m * h] = v[m];
}

// Fused correlation: out[c] = IFFT(FFT(in[c]) * H) for one column c per block.
// H is the (already conjugated when correlating) reference spectrum, length
// S, natural order; tw[k] = exp(-2 pi i k / S), k < S / 2.  The column is
// loaded bit-reversed, transformed by radix-2 DIT stages taken up to three at
// a time in registers (one LDS round trip per group), multiplied by H,
// inverse transformed by DIF stages three at a time, and stored
// bit-reversed / S.
template <typename T>
__global__ __launch_bounds__(1024) void k_fft_filter_lds(const typename C2<T>::type* __restrict__ in,
                                                         typename C2<T>::type* __restrict__ out,
                                                         const typename C2<T>::type* __restrict__ H,
                                                         const typename C2<T>::type* __restrict__ tw, int logS)
{
    using V = typename C2<T>::type;
    extern __shared__ unsigned char smem[];
    V* x = reinterpret_cast<V*>(smem);
    const int S = 1 << logS;
    const int64_t col = blockIdx.x;
    const V* src = in + col * S;
    for (int k = threadIdx.x; k < S; k += blockDim.x) x[bitrev(k, logS)] = src[k];
    __syncthreads();
    // forward DIT: groups of 3 stages, the remainder (1 or 2) first
    int s = 1;
    const int rem = logS % 3;
    if (rem == 1) {
        for (int q = threadIdx.x; q < S / 2; q += blockDim.x) fft_group<T, 1, false>(x, s, q, logS, tw);
        __syncthreads();
        s += 1;
    } else if (rem == 2) {
        for (int q = threadIdx.x; q < S / 4; q += blockDim.x) fft_group<T, 2, false>(x, s, q, logS, tw);
        __syncthreads();
        s += 2;
    }
    for (; s + 2 <= logS; s += 3) {
        for (int q = threadIdx.x; q < S / 8; q += blockDim.x) fft_group<T, 3, false>(x, s, q, logS, tw);
        __syncthreads();
    }
    for (int k = threadIdx.x; k < S; k += blockDim.x) x[k] = cmul(x[k], H[k]);
    __syncthreads();
    // inverse DIF: groups of 3 stages from the top (s = top stage), the
    // remainder (stages rem .. 1) last
    for (s = logS; s >= rem + 3; s -= 3) {
        for (int q = threadIdx.x; q < S / 8; q += blockDim.x) fft_group<T, 3, true>(x, s, q, logS, tw);
        __syncthreads();
    }
    if (rem == 1) {
        for (int q = threadIdx.x; q < S / 2; q += blockDim.x) fft_group<T, 1, true>(x, s, q, logS, tw);
        __syncthreads();
    } else if (rem == 2) {
        for (int q = threadIdx.x; q < S / 4; q += blockDim.x) fft_group<T, 2, true>(x, s, q, logS, tw);
        __syncthreads();
    }
    const T inv = (T)1 / (T)S;
    V* dst = out + col * S;
    for (int k = threadIdx.x; k < S; k += blockDim.x) {
        const V a = x[bitrev(k, logS)];
        dst[k] = V{a.x * inv, a.y * inv};
    }
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

template <typename T>
static size_t lds_bytes(int S) { return (size_t)S * sizeof(typename C2<T>::type); }

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

template <typename T>
hipError_t fft_filter_t(typename C2<T>::type* data, typename C2<T>::type* tmp, const typename C2<T>::type* H, int S,
                        int64_t ncols, hipStream_t st)
{
    using V = typename C2<T>::type;
    if (S < 2 || (S & (S - 1)) || ncols <= 0) return hipErrorInvalidValue;
    const int logS = ilog2(S);
    const size_t lb = lds_bytes<T>(S);
    if (lb <= kFftLdsMax) {
        hipError_t e = hipFuncSetAttribute((const void*)k_fft_filter_lds<T>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
        if (e != hipSuccess) return e;
        const V* tw = nullptr;
        if ((e = twiddle_table<T>(S, st, &tw)) != hipSuccess) return e;
        // one thread per 8-point group of the radix-8 passes
        const int threads = S / 8 >= 1024 ? 1024 : (S / 8 < 64 ? 64 : S / 8);
        hipLaunchKernelGGL(k_fft_filter_lds<T>, dim3((unsigned)ncols), dim3(threads), lb, st, data, data, H, tw,
                           logS);
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

hipError_t conj_inplace(double2* a, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_conj<double>, dim3(nblk(n, 256)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

}  // namespace bsar

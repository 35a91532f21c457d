// kernels_fft.hip -- batched range compression (pulse compression) on gfx950.
//
// The reference compresses every pulse column with FFTProcessor::fftCorrelation
// (fft_utils.cpp:63-99; driven at raw_data_processing.cpp:431-446): FFT of the
// column, FFT of the reference (recomputed per call), multiply by the
// conjugate, inverse FFT, keep the first S samples.  fftConvolution
// (fft_utils.cpp:103-137) is the same without the conjugate.  Here the
// reference spectrum H is computed once per call and each column is one
// workgroup: the column is staged in LDS, transformed in place (radix-2 DIT
// forward from bit-reversed load, DIF inverse to bit-reversed store), so HBM
// sees one read and one write per sample.  Columns larger than the LDS budget
// run as log2(S) global Stockham passes per direction.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Fused correlation: out[c] = IFFT(FFT(in[c]) * H) for one column c per block.
// H is the (already conjugated when correlating) reference spectrum, length
// S, natural order.  The column is loaded bit-reversed, transformed by
// radix-2 DIT stages taken two at a time (one LDS round trip and one
// sin/cos per 4 elements: with w' = exp(-i pi j / 2h) the stage-s twiddle is
// w'^2 and the second stage-(s+1) twiddle -i w'), multiplied by H, inverse
// transformed by DIF stages two at a time, and stored bit-reversed / S.
template <typename T>
__global__ __launch_bounds__(1024) void k_fft_filter_lds(const typename C2<T>::type* __restrict__ in,
                                                         typename C2<T>::type* __restrict__ out,
                                                         const typename C2<T>::type* __restrict__ H, int logS)
{
    using V = typename C2<T>::type;
    extern __shared__ unsigned char smem[];
    V* x = reinterpret_cast<V*>(smem);
    const int S = 1 << logS;
    const int64_t col = blockIdx.x;
    const V* src = in + col * S;
    for (int k = threadIdx.x; k < S; k += blockDim.x) x[bitrev(k, logS)] = src[k];
    __syncthreads();
    // forward DIT; an odd stage count starts with the twiddle-free stage 1
    int s = 1;
    if (logS & 1) {
        for (int b = threadIdx.x; b < S / 2; b += blockDim.x) {
            const V u = x[2 * b], v = x[2 * b + 1];
            x[2 * b] = cadd(u, v);
            x[2 * b + 1] = csub(u, v);
        }
        __syncthreads();
        s = 2;
    }
    for (; s < logS; s += 2) {                       // stages s, s + 1; h = 2^(s-1)
        const int h = 1 << (s - 1);
        for (int q = threadIdx.x; q < S / 4; q += blockDim.x) {
            const int j = q & (h - 1);
            const int e0 = ((q >> (s - 1)) << (s + 1)) + j;
            T sn, cs;
            twiddle((T)(-(T)j / (T)(2 * h)), sn, cs);
            const V w1 = V{cs, sn};                  // exp(-i pi j / 2h)
            const V w0 = cmul(w1, w1);               // exp(-i pi j / h)
            const V w2 = V{sn, -cs};                 // -i w1
            const V x0 = x[e0], x1 = x[e0 + h], x2 = x[e0 + 2 * h], x3 = x[e0 + 3 * h];
            const V t1 = cmul(x1, w0), t3 = cmul(x3, w0);
            const V a0 = cadd(x0, t1), a1 = csub(x0, t1), a2 = cadd(x2, t3), a3 = csub(x2, t3);
            const V u2 = cmul(a2, w1), u3 = cmul(a3, w2);
            x[e0] = cadd(a0, u2);
            x[e0 + 2 * h] = csub(a0, u2);
            x[e0 + h] = cadd(a1, u3);
            x[e0 + 3 * h] = csub(a1, u3);
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < S; k += blockDim.x) x[k] = cmul(x[k], H[k]);
    __syncthreads();
    // inverse DIF, stages taken two at a time from the top; an odd count
    // ends with the twiddle-free stage 1
    for (s = logS; s >= 2; s -= 2) {                 // stages s, s - 1; h = 2^(s-2)
        const int h = 1 << (s - 2);
        for (int q = threadIdx.x; q < S / 4; q += blockDim.x) {
            const int j = q & (h - 1);
            const int e0 = ((q >> (s - 2)) << s) + j;
            T sn, cs;
            twiddle((T)((T)j / (T)(2 * h)), sn, cs);
            const V c1 = V{cs, sn};                  // exp(+i pi j / 2h)
            const V c2 = V{-sn, cs};                 // +i c1
            const V c0 = cmul(c1, c1);               // exp(+i pi j / h)
            const V x0 = x[e0], x1 = x[e0 + h], x2 = x[e0 + 2 * h], x3 = x[e0 + 3 * h];
            const V b0 = cadd(x0, x2), b2 = cmul(csub(x0, x2), c1);
            const V b1 = cadd(x1, x3), b3 = cmul(csub(x1, x3), c2);
            x[e0] = cadd(b0, b1);
            x[e0 + h] = cmul(csub(b0, b1), c0);
            x[e0 + 2 * h] = cadd(b2, b3);
            x[e0 + 3 * h] = cmul(csub(b2, b3), c0);
        }
        __syncthreads();
    }
    if (s == 1) {
        for (int b = threadIdx.x; b < S / 2; b += blockDim.x) {
            const V u = x[2 * b], v = x[2 * b + 1];
            x[2 * b] = cadd(u, v);
            x[2 * b + 1] = csub(u, v);
        }
        __syncthreads();
    }
    const T inv = (T)1 / (T)S;
    V* dst = out + col * S;
    for (int k = threadIdx.x; k < S; k += blockDim.x) {
        const V a = x[bitrev(k, logS)];
        dst[k] = V{a.x * inv, a.y * inv};
    }
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
        const int threads = S / 2 >= 1024 ? 1024 : (S / 2 < 64 ? 64 : S / 2);
        hipLaunchKernelGGL(k_fft_filter_lds<T>, dim3((unsigned)ncols), dim3(threads), lb, st, data, data, H, logS);
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

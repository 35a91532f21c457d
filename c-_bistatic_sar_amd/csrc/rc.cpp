// rc.cpp -- bsar_range_compress: batched FFT correlation / convolution of
// pulse columns against one reference (fft_utils.cpp:63-137 as driven by
// raw_data_processing.cpp:406-446).
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "api_internal.h"
#include "bsar_fft.h"

namespace {

struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
};

size_t next_pow2(size_t n)   // fft_utils.cpp:187-195
{
    size_t p = 1;
    while (p < n) p *= 2;
    return p;
}

}  // namespace

using namespace bsar;

extern "C" int bsar_range_compress(bsar_ctx* c, const double* sig, int S, int N, const double* ref, int conj,
                                   int precision, double* out)
{
    if (!c) return BSAR_ERR_INVALID;
    if (S < 1 || N < 0 || (N && (!sig || !out)) || !ref)
        return set_error(c, BSAR_ERR_INVALID, "range_compress needs S >= 1, N >= 0 and buffers");
    if (precision != BSAR_FP64 && precision != BSAR_MIXED) return set_error(c, BSAR_ERR_INVALID, "bad precision");
    BSAR_CHECK_HIP(c, hipSetDevice(c->device));
    if (N == 0) return BSAR_OK;
    // Both inputs are zero-padded to the next power of two and the first S
    // samples of the circular result are kept (fft_utils.cpp:68-78, 95-98).
    const size_t P = next_pow2((size_t)S);
    if (P > (size_t)1 << 26) return set_error(c, BSAR_ERR_INVALID, "column too long");
    if (P == 1) {
        // length-1 transforms are the identity (fft_utils.cpp:14): the
        // product of the two samples, the reference conjugated for correlation
        const double hr = ref[0], hi = conj ? -ref[1] : ref[1];
        for (int k = 0; k < N; ++k) {
            const double a = sig[2 * k], b = sig[2 * k + 1];
            out[2 * k] = a * hr - b * hi;
            out[2 * k + 1] = a * hi + b * hr;
        }
        return BSAR_OK;
    }
    std::vector<double> padded;
    const double* src = sig;
    if (P != (size_t)S) {
        padded.assign(2 * P * N, 0.0);
        for (int k = 0; k < N; ++k) std::memcpy(&padded[2 * P * k], sig + 2 * (size_t)S * k, 16 * (size_t)S);
        src = padded.data();
    }
    std::vector<double> refp(2 * P, 0.0);
    std::memcpy(refp.data(), ref, 16 * (size_t)S);

    const size_t n = P * (size_t)N;
    Buf dref, dtmp1, data, tmp;
    BSAR_CHECK_HIP(c, hipMalloc(&dref.p, P * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMalloc(&dtmp1.p, P * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMemcpyAsync(dref.p, refp.data(), P * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    double2* H = nullptr;
    BSAR_CHECK_HIP(c, fft_forward((double2*)dref.p, (double2*)dtmp1.p, (int)P, 1, c->stream, &H));
    if (conj) BSAR_CHECK_HIP(c, conj_inplace(H, (int64_t)P, c->stream));
    BSAR_CHECK_HIP(c, hipMalloc(&data.p, n * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMemcpyAsync(data.p, src, n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    if (precision == BSAR_FP64) {
        BSAR_CHECK_HIP(c, hipMalloc(&tmp.p, fft_filter_tmp_elems(P, N, sizeof(double2)) * sizeof(double2)));
        BSAR_CHECK_HIP(c, fft_filter((double2*)data.p, (double2*)tmp.p, H, (int)P, N, c->stream));
    } else {
        Buf d32, h32;
        BSAR_CHECK_HIP(c, hipMalloc(&d32.p, n * sizeof(float2)));
        BSAR_CHECK_HIP(c, hipMalloc(&h32.p, P * sizeof(float2)));
        BSAR_CHECK_HIP(c, hipMalloc(&tmp.p, fft_filter_tmp_elems(P, N, sizeof(float2)) * sizeof(float2)));
        BSAR_CHECK_HIP(c, to_f32((const double2*)data.p, (float2*)d32.p, (int64_t)n, c->stream));
        BSAR_CHECK_HIP(c, to_f32(H, (float2*)h32.p, (int64_t)P, c->stream));
        BSAR_CHECK_HIP(c, fft_filter((float2*)d32.p, (float2*)tmp.p, (const float2*)h32.p, (int)P, N, c->stream));
        BSAR_CHECK_HIP(c, to_f64((const float2*)d32.p, (double2*)data.p, (int64_t)n, c->stream));
        BSAR_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (P == (size_t)S) {
        BSAR_CHECK_HIP(c, hipMemcpyAsync(out, data.p, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        BSAR_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<double> full(2 * n);
        BSAR_CHECK_HIP(c, hipMemcpyAsync(full.data(), data.p, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        BSAR_CHECK_HIP(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < N; ++k) std::memcpy(out + 2 * (size_t)S * k, &full[2 * P * k], 16 * (size_t)S);
    }
    return BSAR_OK;
}

extern "C" int bsar_range_compress_device(bsar_ctx* c, void* d_sig, int S, int N, const double* ref, int conj,
                                          int precision, bsar_report* rep)
{
    if (!c) return BSAR_ERR_INVALID;
    if (S < 2 || (S & (S - 1)) || N < 0 || (N && !d_sig) || !ref)
        return set_error(c, BSAR_ERR_INVALID, "range_compress_device needs a power-of-two S >= 2 and buffers");
    if (precision != BSAR_FP64 && precision != BSAR_MIXED) return set_error(c, BSAR_ERR_INVALID, "bad precision");
    if (S > (1 << 26)) return set_error(c, BSAR_ERR_INVALID, "column too long");
    BSAR_CHECK_HIP(c, hipSetDevice(c->device));
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (N == 0) return BSAR_OK;
    const size_t P = (size_t)S, n = P * (size_t)N;
    // filter spectrum H = FFT(ref) (conjugated for correlation), FP64
    Buf dref, dtmp1, h32, tmp;
    BSAR_CHECK_HIP(c, hipMalloc(&dref.p, P * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMalloc(&dtmp1.p, P * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMemcpyAsync(dref.p, ref, P * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    double2* H = nullptr;
    BSAR_CHECK_HIP(c, fft_forward((double2*)dref.p, (double2*)dtmp1.p, (int)P, 1, c->stream, &H));
    if (conj) BSAR_CHECK_HIP(c, conj_inplace(H, (int64_t)P, c->stream));
    const size_t esz = precision == BSAR_FP64 ? sizeof(double2) : sizeof(float2);
    BSAR_CHECK_HIP(c, hipMalloc(&tmp.p, fft_filter_tmp_elems(P, N, esz) * esz));
    if (precision == BSAR_MIXED) {
        BSAR_CHECK_HIP(c, hipMalloc(&h32.p, P * sizeof(float2)));
        BSAR_CHECK_HIP(c, to_f32(H, (float2*)h32.p, (int64_t)P, c->stream));
    }
    // The reordered spectrum is made before the timed region, which then
    // holds the column transforms and products only.
    if (precision == BSAR_FP64)
        BSAR_CHECK_HIP(c, fft_filter_prepare(H, (double2*)tmp.p, (int)P, N, c->stream));
    else
        BSAR_CHECK_HIP(c, fft_filter_prepare((const float2*)h32.p, (float2*)tmp.p, (int)P, N, c->stream));
    BSAR_CHECK_HIP(c, hipEventRecord(c->ev0, c->stream));
    if (precision == BSAR_FP64)
        BSAR_CHECK_HIP(c, fft_filter_apply((double2*)d_sig, (double2*)tmp.p, H, (int)P, N, c->stream));
    else
        BSAR_CHECK_HIP(c, fft_filter_apply((float2*)d_sig, (float2*)tmp.p, (const float2*)h32.p, (int)P, N, c->stream));
    BSAR_CHECK_HIP(c, hipEventRecord(c->ev1, c->stream));
    BSAR_CHECK_HIP(c, hipEventSynchronize(c->ev1));
    if (rep) {
        float ms = 0.f;
        BSAR_CHECK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        rep->seconds = ms * 1e-3;
        rep->pulses = N;
        rep->voxels = 0;
        rep->launches = 1;
    }
    return BSAR_OK;
}

extern "C" int bsar_fft(bsar_ctx* c, double* data, int S, int N, int inverse)
{
    if (!c) return BSAR_ERR_INVALID;
    if (S < 0 || N < 0 || (S && N && !data)) return set_error(c, BSAR_ERR_INVALID, "fft needs S, N >= 0 and a buffer");
    if (S > (1 << 26)) return set_error(c, BSAR_ERR_INVALID, "column too long");
    BSAR_CHECK_HIP(c, hipSetDevice(c->device));
    if (S == 0 || N == 0) return BSAR_OK;
    const size_t n = (size_t)S * (size_t)N;
    const bool pow2 = (S & (S - 1)) == 0;
    // fft_utils.cpp:14-20: lengths 1 and non-powers of two leave the data as
    // it is; ifft (:46-60) still divides by the length around that no-op.
    if (S == 1 || !pow2) {
        if (inverse && S > 1)
            for (size_t k = 0; k < 2 * n; ++k) data[k] /= (double)S;
        return BSAR_OK;
    }
    Buf a, b;
    BSAR_CHECK_HIP(c, hipMalloc(&a.p, n * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMalloc(&b.p, n * sizeof(double2)));
    BSAR_CHECK_HIP(c, hipMemcpyAsync(a.p, data, n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    if (inverse) BSAR_CHECK_HIP(c, conj_inplace((double2*)a.p, (int64_t)n, c->stream));
    double2* res = nullptr;
    BSAR_CHECK_HIP(c, fft_forward((double2*)a.p, (double2*)b.p, S, N, c->stream, &res));
    if (inverse) {
        BSAR_CHECK_HIP(c, conj_inplace(res, (int64_t)n, c->stream));
        BSAR_CHECK_HIP(c, scale_inplace(res, (int64_t)n, 1.0 / (double)S, c->stream));
    }
    BSAR_CHECK_HIP(c, hipMemcpyAsync(data, res, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    BSAR_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

// ---- range compression with no allocation --------------------------------
//
// bsar_range_compress_device_noalloc runs bsar_range_compress_device's kernels
// with every buffer they need beyond the caller's columns (the reference
// spectrum, its FFT ping-pong buffer, the filter's bit-reversed spectrum or
// global-pass buffer, the FP32 spectrum, the twiddle tables) carved out of
// one file-scope device array.  Each GPU holds its own copy of `scratch`; its
// device address is looked up once per device with that device current
// (scratch_init, called when a context is created) and only that address is
// handed to kernels and copies.  Every call on a GPU shares the GPU's array:
// calls that overlap on one GPU (two host threads, two contexts) race.

namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;   // 4 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

struct ScratchLayout {
    size_t ref, tmp1, tmp, h32, tw, twn, total;
};

ScratchLayout scratch_layout(size_t S, size_t N, size_t esz, bool mixed)
{
    ScratchLayout L{};
    size_t off = 0;
    L.ref = off;  off += up256(S * sizeof(double2));
    L.tmp1 = off; off += up256(S * sizeof(double2));
    L.tmp = off;  off += up256(fft_filter_tmp_elems(S, N, esz) * esz);
    L.h32 = off;  off += mixed ? up256(S * sizeof(float2)) : 0;
    L.tw = off;   off += up256((S / 2 > 0 ? S / 2 : 1) * esz);
    L.twn = off;  off += up256((S / 4 > 0 ? S / 4 : 1) * esz);
    L.total = off;
    return L;
}

}  // namespace

namespace bsar {

int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return BSAR_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return g_scratch_err[device] == hipSuccess ? BSAR_OK : BSAR_ERR_DEVICE;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_done[device] = true;
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    return e == hipSuccess ? BSAR_OK : BSAR_ERR_DEVICE;
}

}  // namespace bsar

extern "C" int bsar_range_compress_device_noalloc(bsar_ctx* c, void* d_sig, int S, int N, const double* ref,
                                                  int conj, int precision, bsar_report* rep)
{
    if (!c) return BSAR_ERR_INVALID;
    if (S < 2 || (S & (S - 1)) || N < 0 || (N && !d_sig) || !ref)
        return set_error(c, BSAR_ERR_INVALID,
                         "range_compress_device_noalloc needs a power-of-two S >= 2 and buffers");
    if (precision != BSAR_FP64 && precision != BSAR_MIXED) return set_error(c, BSAR_ERR_INVALID, "bad precision");
    if (S > (1 << 26)) return set_error(c, BSAR_ERR_INVALID, "column too long");
    if (rep) std::memset(rep, 0, sizeof *rep);
    const bool mixed = precision == BSAR_MIXED;
    const size_t P = (size_t)S, esz = mixed ? sizeof(float2) : sizeof(double2);
    const ScratchLayout L = scratch_layout(P, (size_t)N, esz, mixed);
    if (L.total > kScratchBytes)
        return set_error(c, BSAR_ERR_CAPACITY,
                         "range_compress_device_noalloc: S = " + std::to_string(S) + ", N = " + std::to_string(N) +
                             " needs " + std::to_string(L.total) + " bytes of scratch; the no-allocation entry "
                             "point has " + std::to_string(kScratchBytes) + " (use bsar_range_compress_device)");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxDevices && g_scratch_done[c->device])
            base = g_scratch_ptr[c->device];
    }
    if (!base) return set_error(c, BSAR_ERR_DEVICE, "no device scratch on this GPU (address lookup failed)");
    BSAR_CHECK_HIP(c, hipSetDevice(c->device));
    if (N == 0) return BSAR_OK;
    unsigned char* b = static_cast<unsigned char*>(base);
    double2* dref = reinterpret_cast<double2*>(b + L.ref);
    double2* dtmp1 = reinterpret_cast<double2*>(b + L.tmp1);
    BSAR_CHECK_HIP(c, hipMemcpyAsync(dref, ref, P * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    double2* H = nullptr;
    BSAR_CHECK_HIP(c, fft_forward(dref, dtmp1, (int)P, 1, c->stream, &H));
    if (conj) BSAR_CHECK_HIP(c, conj_inplace(H, (int64_t)P, c->stream));
    if (mixed) {
        float2* h32 = reinterpret_cast<float2*>(b + L.h32);
        float2* tw = reinterpret_cast<float2*>(b + L.tw);
        float2* twn = reinterpret_cast<float2*>(b + L.twn);
        BSAR_CHECK_HIP(c, to_f32(H, h32, (int64_t)P, c->stream));
        BSAR_CHECK_HIP(c, fft_twiddles(tw, (int)P, c->stream));
        BSAR_CHECK_HIP(c, fft_twiddles(twn, (int)(P / 2), c->stream));
        float2* tmp = reinterpret_cast<float2*>(b + L.tmp);
        BSAR_CHECK_HIP(c, fft_filter_prepare(h32, tmp, (int)P, N, c->stream, tw, twn));
        BSAR_CHECK_HIP(c, hipEventRecord(c->ev0, c->stream));
        BSAR_CHECK_HIP(c, fft_filter_apply((float2*)d_sig, tmp, h32, (int)P, N, c->stream, tw, twn));
    } else {
        double2* tw = reinterpret_cast<double2*>(b + L.tw);
        double2* twn = reinterpret_cast<double2*>(b + L.twn);
        BSAR_CHECK_HIP(c, fft_twiddles(tw, (int)P, c->stream));
        BSAR_CHECK_HIP(c, fft_twiddles(twn, (int)(P / 2), c->stream));
        double2* tmp = reinterpret_cast<double2*>(b + L.tmp);
        BSAR_CHECK_HIP(c, fft_filter_prepare(H, tmp, (int)P, N, c->stream, tw, twn));
        BSAR_CHECK_HIP(c, hipEventRecord(c->ev0, c->stream));
        BSAR_CHECK_HIP(c, fft_filter_apply((double2*)d_sig, tmp, H, (int)P, N, c->stream, tw, twn));
    }
    BSAR_CHECK_HIP(c, hipEventRecord(c->ev1, c->stream));
    BSAR_CHECK_HIP(c, hipEventSynchronize(c->ev1));
    if (rep) {
        float ms = 0.f;
        BSAR_CHECK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        rep->seconds = ms * 1e-3;
        rep->pulses = N;
        rep->launches = 1;
    }
    return BSAR_OK;
}

// bsar_fft.h -- batched FFT filtering (range compression) launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bsar {

// Columns whose complex samples fit in this many LDS bytes run in one
// workgroup each; larger ones run as global radix-2 passes.
constexpr size_t kFftLdsMax = 128 * 1024;

// data[c] <- IFFT(FFT(data[c]) .* H) for ncols contiguous columns of length
// S (power of two).  Columns within the LDS budget run in LDS (16384
// complex128 samples too, half of the column at a time; 16384-sample columns
// in a persistent workgroup per CU); other columns of
// twice that as two half-length LDS filters between a global radix-2 split
// and merge; longer ones as global radix-2 passes.  tmp: scratch of
// fft_filter_tmp_elems(S, ncols, bytes per complex) complex elements -- S
// (the spectrum in bit-reversed order) for the LDS paths, ncols*S for the
// global path.
inline size_t fft_filter_tmp_elems(size_t S, size_t ncols, size_t elem_bytes)
{
    return (S / 2) * elem_bytes > kFftLdsMax ? S * ncols : S;
}
hipError_t fft_filter(double2* data, double2* tmp, const double2* H, int S, int64_t ncols, hipStream_t st);
hipError_t fft_filter(float2* data, float2* tmp, const float2* H, int S, int64_t ncols, hipStream_t st);
// fft_filter in two steps on one stream: fft_filter_prepare writes the
// reordered spectrum into tmp (nothing on the global path) and builds the
// cached twiddle tables on first use; fft_filter_apply then filters the
// columns with tmp as prepared (the same tmp, S and ncols).  tw / twn: the
// twiddle tables of S and S / 2 (lengths S / 2 and S / 4, at least one
// entry each) as fft_twiddles fills them, so that no path allocates; null
// takes the per-process tables, built on first use.
hipError_t fft_filter_prepare(const double2* H, double2* tmp, int S, int64_t ncols, hipStream_t st,
                              const double2* tw = nullptr, const double2* twn = nullptr);
hipError_t fft_filter_prepare(const float2* H, float2* tmp, int S, int64_t ncols, hipStream_t st,
                              const float2* tw = nullptr, const float2* twn = nullptr);
hipError_t fft_filter_apply(double2* data, double2* tmp, const double2* H, int S, int64_t ncols, hipStream_t st,
                            const double2* tw = nullptr, const double2* twn = nullptr);
hipError_t fft_filter_apply(float2* data, float2* tmp, const float2* H, int S, int64_t ncols, hipStream_t st,
                            const float2* tw = nullptr, const float2* twn = nullptr);
// tw[k] = exp(-2 pi i k / S), k < max(S / 2, 1) -- the cached tables' contents.
hipError_t fft_twiddles(double2* tw, int S, hipStream_t st);
hipError_t fft_twiddles(float2* tw, int S, hipStream_t st);
// Forward FFT of ncols columns (global passes); *result is data or tmp.
hipError_t fft_forward(double2* data, double2* tmp, int S, int64_t ncols, hipStream_t st, double2** result);
hipError_t to_f32(const double2* a, float2* b, int64_t n, hipStream_t st);
hipError_t to_f64(const float2* a, double2* b, int64_t n, hipStream_t st);
hipError_t conj_inplace(double2* a, int64_t n, hipStream_t st);
hipError_t scale_inplace(double2* a, int64_t n, double s, hipStream_t st);

}  // namespace bsar

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
// Same with the twiddle tables supplied: tw of length S / 2 and twn of length
// S / 4 (both entries of at least one), as fft_twiddles fills them; no
// allocation on any path.
hipError_t fft_filter_tw(double2* data, double2* tmp, const double2* H, int S, int64_t ncols, hipStream_t st,
                         const double2* tw, const double2* twn);
hipError_t fft_filter_tw(float2* data, float2* tmp, const float2* H, int S, int64_t ncols, hipStream_t st,
                         const float2* tw, const float2* twn);
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

// sim.cpp -- bsar_process_raw: processRawData (raw_data_processing.cpp:292-452)
// on the GPU.  Declaration and contract in include/bsar.h.
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.h"
#include "bsar_fft.h"
#include "bsar_sim.h"

using namespace bsar;

namespace {

struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" int bsar_process_raw(bsar_ctx* c, const bsar_dataset* ds, const double* scat, const double* rcs, int ns,
                                const bsar_chirp_params* chirp, int S, double* aux, double* raw, double* root)
{
    if (!c || !ds) return BSAR_ERR_INVALID;
    if (!chirp || ns < 0 || (ns && (!scat || !rcs))) return set_error(c, BSAR_ERR_INVALID, "bad scatterer buffers");
    if (S < 2 || (S & (S - 1)) || S > (1 << 24))
        return set_error(c, BSAR_ERR_INVALID, "S (IndiceMaximo) must be a power of two >= 2");
    if (ds->device != c->device) return set_error(c, BSAR_ERR_INVALID, "dataset lives on another device");
    // createCHIRP (raw_data_processing.cpp:267-282) and the fftshifted
    // reference of :371-394, zero-padded to S
    const int L = bsar_chirp(chirp->f0, chirp->f1, chirp->T, chirp->fs, nullptr);
    if (L < 0) return set_error(c, BSAR_ERR_INVALID, "bad chirp parameters");
    if (L > S) return set_error(c, BSAR_ERR_INVALID, "chirp longer than the range window");
    std::vector<double> ch(2 * (size_t)L), ref(2 * (size_t)S, 0.0);
    bsar_chirp(chirp->f0, chirp->f1, chirp->T, chirp->fs, ch.data());
    const int half = L / 2;
    for (int i = 0; i < L - half; ++i) { ref[2 * i] = ch[2 * (half + i)]; ref[2 * i + 1] = ch[2 * (half + i) + 1]; }
    for (int i = 0; i < half; ++i) {
        ref[2 * (S - half + i)] = ch[2 * i];
        ref[2 * (S - half + i) + 1] = ch[2 * i + 1];
    }
    BSAR_CHECK_HIP(c, hipSetDevice(c->device));
    const int n = ds->n;
    const size_t cells = (size_t)S * n;
    if (cells == 0) return BSAR_OK;
    Buf dscat, drcs, dbins, dvals, daux, dref, dtmp, dtmp2, dconj;
    BSAR_CHECK_HIP(c, hipMalloc(&daux.p, cells * sizeof(double2)));
    BSAR_CHECK_HIP(c, fill_zero((double2*)daux.p, (int64_t)cells, c->stream));
    if (ns > 0) {
        const size_t m = (size_t)ns * n;
        BSAR_CHECK_HIP(c, hipMalloc(&dscat.p, (size_t)ns * 3 * sizeof(double)));
        BSAR_CHECK_HIP(c, hipMalloc(&drcs.p, (size_t)ns * sizeof(double2)));
        BSAR_CHECK_HIP(c, hipMalloc(&dbins.p, m * sizeof(int)));
        BSAR_CHECK_HIP(c, hipMalloc(&dvals.p, m * sizeof(double2)));
        BSAR_CHECK_HIP(c, hipMemcpyAsync(dscat.p, scat, (size_t)ns * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        BSAR_CHECK_HIP(c, hipMemcpyAsync(drcs.p, rcs, (size_t)ns * sizeof(double2), hipMemcpyHostToDevice, c->stream));
        SimArgs sa{};
        sa.n1 = ds->p.n1;
        sa.n2 = ds->p.n2;
        sa.c = ds->p.c;
        sa.fs = ds->p.fs;
        sa.lamb = ds->p.c / ds->p.fc;                 // RadarTx.lamb
        sa.S = S;
        BSAR_CHECK_HIP(c, launch_sim_terms(ds->dem, ds->traj, (const double*)dscat.p, (const double2*)drcs.p, ns, sa,
                                           (int*)dbins.p, (double2*)dvals.p, c->stream));
        BSAR_CHECK_HIP(c, launch_sim_place((const int*)dbins.p, (const double2*)dvals.p, ns, n, S, (double2*)daux.p,
                                           c->stream));
    }
    if (aux) BSAR_CHECK_HIP(c, hipMemcpyAsync(aux, daux.p, cells * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    if (raw || root) {
        // H = FFT(reference); raw = IFFT(FFT(aux) H) (fftConvolution, fft_utils.cpp:103-137)
        BSAR_CHECK_HIP(c, hipMalloc(&dref.p, (size_t)S * sizeof(double2)));
        BSAR_CHECK_HIP(c, hipMalloc(&dtmp.p, (size_t)S * sizeof(double2)));
        BSAR_CHECK_HIP(c, hipMemcpyAsync(dref.p, ref.data(), (size_t)S * sizeof(double2), hipMemcpyHostToDevice, c->stream));
        double2* H = nullptr;
        BSAR_CHECK_HIP(c, fft_forward((double2*)dref.p, (double2*)dtmp.p, S, 1, c->stream, &H));
        BSAR_CHECK_HIP(c, hipMalloc(&dtmp2.p, fft_filter_tmp_elems(S, n, sizeof(double2)) * sizeof(double2)));
        BSAR_CHECK_HIP(c, fft_filter((double2*)daux.p, (double2*)dtmp2.p, H, S, n, c->stream));
        if (raw) BSAR_CHECK_HIP(c, hipMemcpyAsync(raw, daux.p, cells * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        if (root) {
            // rootData = IFFT(FFT(raw) conj(H)) (fftCorrelation, fft_utils.cpp:63-99; :431-446)
            BSAR_CHECK_HIP(c, hipMalloc(&dconj.p, (size_t)S * sizeof(double2)));
            BSAR_CHECK_HIP(c, hipMemcpyAsync(dconj.p, H, (size_t)S * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
            BSAR_CHECK_HIP(c, conj_inplace((double2*)dconj.p, S, c->stream));
            BSAR_CHECK_HIP(c, fft_filter((double2*)daux.p, (double2*)dtmp2.p, (const double2*)dconj.p, S, n, c->stream));
            BSAR_CHECK_HIP(c, hipMemcpyAsync(root, daux.p, cells * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        }
    }
    BSAR_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    return BSAR_OK;
}

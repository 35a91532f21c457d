// valu_rates.hip -- VALU instruction rates of the GPU this runs on.
//
// Each kernel runs 8 waves per SIMD (2048 workgroups of 256 threads on 256
// CUs), every wave 8 independent dependency chains of one instruction kind,
// long enough (>= 1.5 ms) for the clock to settle.  Prints the time per
// wave-instruction per SIMD in cycles of the 2.4 GHz peak engine clock and
// the shader clock block 0 saw (s_memtime against the 100 MHz s_memrealtime).
// Built with the library's flags (FMA contraction off, packed-FP32
// instructions disabled):
//
//   hipcc --offload-arch=gfx950 -O3 -ffp-contract=off -Xclang -target-feature \
//       -Xclang -packed-fp32-ops c-_bistatic_sar_amd/tools/valu_rates.hip -o /tmp/valu_rates
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                                \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                 \
            return 1;                                                                           \
        }                                                                                       \
    } while (0)

__device__ unsigned long long g_clk[2];

enum Op { FMA64, FMA32, MULADD64, MULADD32, RSQ64, RSQ32, SQRT32, ADDU32, SHIFTXOR, CVT, SELECT32 };

template <int OP, int ITERS>
__global__ __launch_bounds__(256) void k_rate(const double* __restrict__ in, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c0 = 0, r0 = 0;
    if (t == 0) {
        c0 = __builtin_amdgcn_s_memtime();
        r0 = __builtin_amdgcn_s_memrealtime();
    }
    double a[8];
    float f[8];
    int ia[8];
    unsigned ib[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = in[(t + i) & 1023] + i;
        f[i] = (float)a[i];
        ia[i] = t + i;
        ib[i] = 3u * t + i;
    }
    for (int it = 0; it < ITERS; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (OP == FMA64) a[i] = __builtin_fma(a[i], 0.999999, 1e-9);
            if (OP == FMA32) f[i] = __builtin_fmaf(f[i], 0.999999f, 1e-9f);
            if (OP == MULADD64) a[i] = a[i] * 0.999999 + 1e-9;
            if (OP == MULADD32) f[i] = f[i] * 0.999999f + 1e-9f;
            if (OP == RSQ64) a[i] = __builtin_amdgcn_rsq(a[i]);
            if (OP == RSQ32) f[i] = __builtin_amdgcn_rsqf(f[i]);
            if (OP == SQRT32) f[i] = __builtin_amdgcn_sqrtf(f[i]);
            if (OP == ADDU32) ia[i] = ia[i] + ia[(i + 3) & 7];
            if (OP == SHIFTXOR) ib[i] = (ib[i] >> 1) ^ (ib[i] << 3);
            if (OP == CVT) a[i] = (double)(float)a[i] * 1.0000001;
            if (OP == SELECT32) f[i] = (f[i] > 0.5f) ? f[i] * 0.5f : f[i] + 0.75f;
        }
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += a[i] + f[i] + ia[i] + ib[i];
    out[t] = s;
    if (t == 0) {
        g_clk[0] = __builtin_amdgcn_s_memtime() - c0;
        g_clk[1] = __builtin_amdgcn_s_memrealtime() - r0;
    }
}

template <int OP, int ITERS>
static int run(const char* name, int per_iter, const double* din, double* dout, int blocks, int simds)
{
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL((k_rate<OP, ITERS>), dim3(blocks), dim3(256), 0, 0, din, dout);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL((k_rate<OP, ITERS>), dim3(blocks), dim3(256), 0, 0, din, dout);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long clk[2];
    CHECK(hipMemcpyFromSymbol(clk, HIP_SYMBOL(g_clk), sizeof clk));
    const double waves = blocks * 4.0;
    const double winst = waves * ITERS * per_iter;          // wave-instructions issued
    const double cycles = ms * 1e-3 * 2.4e9 * simds / winst;  // at the 2.4 GHz peak clock
    const double ghz = clk[1] ? (double)clk[0] / ((double)clk[1] * 10.0) : 0.0;
    std::printf("| %-26s | %8.3f | %5.2f | %4.2f |\n", name, ms, cycles, ghz);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return 0;
}

int main()
{
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int simds = 4 * cus;
    const int blocks = 2 * simds;   // 256 threads = 4 waves: 8 waves per SIMD
    double *din = nullptr, *dout = nullptr;
    CHECK(hipMalloc(&din, 1024 * sizeof(double)));
    CHECK(hipMalloc(&dout, (size_t)blocks * 256 * sizeof(double)));
    double h[1024];
    for (int i = 0; i < 1024; ++i) h[i] = 1.0 + i * 1e-3;
    CHECK(hipMemcpy(din, h, sizeof h, hipMemcpyHostToDevice));
    std::printf("%d CUs, %d SIMDs, 8 waves per SIMD\n\n", cus, simds);
    std::printf("| instructions (8 chains per wave) | ms | cycles per wave-instruction per SIMD at 2.4 GHz | shader GHz |\n");
    std::printf("|---|---:|---:|---:|\n");
    int rc = 0;
    rc |= run<FMA64, 16384>("v_fma_f64", 8, din, dout, blocks, simds);
    rc |= run<FMA32, 32768>("v_fma_f32", 8, din, dout, blocks, simds);
    rc |= run<MULADD64, 8192>("v_mul_f64 + v_add_f64", 16, din, dout, blocks, simds);
    rc |= run<MULADD32, 16384>("v_mul_f32 + v_add_f32", 16, din, dout, blocks, simds);
    rc |= run<RSQ64, 4096>("v_rsq_f64", 8, din, dout, blocks, simds);
    rc |= run<RSQ32, 8192>("v_rsq_f32", 8, din, dout, blocks, simds);
    rc |= run<SQRT32, 8192>("v_sqrt_f32", 8, din, dout, blocks, simds);
    rc |= run<ADDU32, 32768>("v_add_u32", 8, din, dout, blocks, simds);
    rc |= run<SHIFTXOR, 8192>("shift, shift, xor (b32)", 24, din, dout, blocks, simds);
    rc |= run<CVT, 8192>("cvt f32<-f64, cvt f64<-f32, mul_f64", 24, din, dout, blocks, simds);
    rc |= run<SELECT32, 8192>("cmp, mul, add, cndmask f32", 32, din, dout, blocks, simds);
    CHECK(hipFree(din));
    CHECK(hipFree(dout));
    return rc;
}

// valu_ceiling -- measured fused multiply-add throughput of the vector ALU on
// this GPU, the ceiling of the pushers on analytic fields (VALU-bound).
//
// Every lane runs 8 independent FMA chains x = fma(x, a, b) for a fixed
// count, in FP64 and in FP32.  The library is built with packed FP32
// instructions kept out of the code object, and so is this program, so the
// FP32 figure is the rate an ordinary float FMA issues at here (the vendor's
// FP32 peak counts packed math).  Prints one JSON line: FLOP/s with an FMA
// counted as 2.
//
// The loop count is fixed and nothing is read from memory, so every wave
// ends; one store per lane keeps the chains live.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                        \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "valu_ceiling: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            exit(1);                                                                 \
        }                                                                            \
    } while (0)

namespace {

constexpr int kChains = 8;

template <typename T>
__global__ void __launch_bounds__(256) k_fma(T a, T b, int iters, T* __restrict__ out)
{
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    T x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = (T)(tid & 1023) * (T)1e-4 + (T)c;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = fma(x[c], a, b);
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[tid] = s;
}

template <typename T>
double flops(int blocks, int iters)
{
    T* out = nullptr;
    CK(hipMalloc((void**)&out, sizeof(T) * (size_t)blocks * 256));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL((k_fma<T>), dim3(blocks), dim3(256), 0, 0, (T)0.9999999, (T)1e-7, iters, out);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < best) best = ms;   // first launch is a warm-up
    }
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    CK(hipFree(out));
    return (double)blocks * 256.0 * iters * kChains * 2.0 / (best * 1e-3);
}

}  // namespace

int main()
{
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    // 16 blocks of 4 waves per CU: every SIMD holds several waves
    const int blocks = prop.multiProcessorCount * 16;
    const int iters = 8192;
    const double f64 = flops<double>(blocks, iters);
    const double f32 = flops<float>(blocks, iters);
    printf("{\"cu\": %d, \"clock_mhz\": %d, \"fp64_fma_flops_per_s\": %.6e, \"fp32_fma_flops_per_s\": %.6e}\n",
           prop.multiProcessorCount, prop.clockRate / 1000, f64, f32);
    return 0;
}

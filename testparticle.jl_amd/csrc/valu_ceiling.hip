// valu_ceiling -- measured fused multiply-add throughput of the vector ALU on
// this GPU, the ceiling of the pushers on analytic fields (VALU-bound), and
// the issue rate of the other instruction kinds the pushers are made of.
//
// Every lane runs 8 independent chains x = op(x, a, b) for a fixed count, in
// FP64 and in FP32.  The library is built with packed FP32 instructions kept
// out of the code object, and so is this program, so the FP32 figure is the
// rate an ordinary float FMA issues at here (the vendor's FP32 peak counts
// packed math).  Prints one JSON line: FLOP/s of the FMA chains with an FMA
// counted as 2, and lane-instructions/s per kind (fma, add, mul, a compare
// and select pair counted as two instructions, a 32-bit xor-and-add pair,
// which the compiler fuses into one v_xad_u32, and v_readlane_b32 counted as
// 64 lane-instructions per wave instruction, as SQ_INSTS_VALU x 64 counts it).
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

enum Op { kFma, kAdd, kMul, kSel, kInt, kReadlane };

template <typename T, int OP>
__device__ __forceinline__ T step(T x, T a, T b)
{
    if constexpr (OP == kFma) return fma(x, a, b);
    else if constexpr (OP == kAdd) return x + b;
    else if constexpr (OP == kMul) return x * a;
    else return x < b ? a : x;                       // kSel: v_cmp + v_cndmask
}

template <typename T, int OP>
__global__ void __launch_bounds__(256) k_op(T a, T b, int iters, T* __restrict__ out)
{
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (OP == kReadlane) {
        // v_readlane_b32 (one lane of a VGPR into an SGPR), the instruction
        // that keeps uniform values in VGPR lanes when SGPRs run out; the
        // lane index changes every iteration so the reads cannot be hoisted
        int v[kChains];
#pragma unroll
        for (int c = 0; c < kChains; ++c) v[c] = (int)tid * (c + 1);
        int acc = 0;
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) acc ^= __builtin_amdgcn_readlane(v[c], (it + c) & 63);
        }
        out[tid] = (T)acc;
    } else if constexpr (OP == kInt) {
        unsigned x[kChains];
        const unsigned ia = (unsigned)(tid * 2654435761u);
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = tid + (unsigned)c;
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = (x[c] ^ ia) + (unsigned)c;
        }
        unsigned s = 0;
#pragma unroll
        for (int c = 0; c < kChains; ++c) s ^= x[c];
        out[tid] = (T)s;
    } else {
        T x[kChains];
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = (T)(tid & 1023) * (T)1e-4 + (T)c;
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = step<T, OP>(x[c], a, b);
        }
        T s = 0;
#pragma unroll
        for (int c = 0; c < kChains; ++c) s += x[c];
        out[tid] = s;
    }
}

// lane-operations per second of one chain kind (best of three timed launches)
template <typename T, int OP>
double lane_ops(int blocks, int iters)
{
    T* out = nullptr;
    CK(hipMalloc((void**)&out, sizeof(T) * (size_t)blocks * 256));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL((k_op<T, OP>), dim3(blocks), dim3(256), 0, 0, (T)0.9999999, (T)1e-7, iters, out);
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
    return (double)blocks * 256.0 * iters * kChains / (best * 1e-3);
}

}  // namespace

int main()
{
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    // 16 blocks of 4 waves per CU: every SIMD holds several waves
    const int blocks = prop.multiProcessorCount * 16;
    const int iters = 8192;
    const double f64 = lane_ops<double, kFma>(blocks, iters);
    const double f32 = lane_ops<float, kFma>(blocks, iters);
    printf("{\"cu\": %d, \"clock_mhz\": %d, \"fp64_fma_flops_per_s\": %.6e, \"fp32_fma_flops_per_s\": %.6e, "
           "\"lane_instructions_per_s\": {\"fma_f64\": %.6e, \"add_f64\": %.6e, \"mul_f64\": %.6e, "
           "\"fma_f32\": %.6e, \"add_f32\": %.6e, \"mul_f32\": %.6e, \"cmp_sel_f32\": %.6e, "
           "\"add_u32\": %.6e, \"readlane_b32\": %.6e}}\n",
           prop.multiProcessorCount, prop.clockRate / 1000, 2.0 * f64, 2.0 * f32, f64,
           lane_ops<double, kAdd>(blocks, iters), lane_ops<double, kMul>(blocks, iters), f32,
           lane_ops<float, kAdd>(blocks, iters), lane_ops<float, kMul>(blocks, iters),
           2.0 * lane_ops<float, kSel>(blocks, iters), lane_ops<float, kInt>(blocks, iters),
           lane_ops<float, kReadlane>(blocks, iters));
    return 0;
}

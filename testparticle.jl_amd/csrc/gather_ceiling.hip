// gather_ceiling -- measured ceiling of the pushers' field gather on this GPU.
//
// The gridded pushers read, per particle-step, one cell record of the
// cell-replicated grid at a position that is effectively random (particles
// cross several cells per step): 96 B in one 128-B line for float cells,
// 192 B over two lines for double cells.  This program issues exactly that
// access pattern with no arithmetic around it: every lane walks a dependent
// chain of random cell indices over a 2 GiB buffer (the size of the 256^3
// float cell grid), so the rate it reaches is the ceiling a pusher kernel can
// approach.  Prints one JSON line: gathers/s and line-equivalent bytes/s.
// An optional argument sets the buffer size in GiB (1-64): random gathers over
// more memory than the address translation caches cover run slower, which is
// what a larger grid (or many grids at once) would meet.
//
// Every index is (hash % ncells), so every access stays inside the buffer; the
// loop count is fixed, so every wave ends.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                        \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "gather_ceiling: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            exit(1);                                                                 \
        }                                                                            \
    } while (0)

namespace {

__device__ __forceinline__ unsigned mix(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// NV 16-B loads per gather from a record of STRIDE16 16-B units
template <int NV, int STRIDE16>
__global__ void __launch_bounds__(256) k_chain(const float4* __restrict__ buf, unsigned ncells,
                                               int iters, float* __restrict__ out)
{
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned idx = mix(tid * 2654435761u + 1u) % ncells;
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
        const float4* p = buf + (size_t)idx * STRIDE16;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const float4 v = p[k];
            s += (v.x + v.y) + (v.z + v.w);
        }
        acc += s;
        // the next index depends on the loaded data, as a pusher's next
        // position depends on the field it just read
        idx = mix(idx ^ (unsigned)(s != 1.0f) ^ ((unsigned)it * 0x9e3779b9u)) % ncells;
    }
    out[tid] = acc;
}

template <class K>
double rate(K kern, const float4* buf, unsigned ncells, float* out, int blocks, int iters)
{
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        CK(hipEventRecord(a));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, buf, ncells, iters, out);
        CK(hipGetLastError());
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, a, b));
        if (rep > 0 && ms < best) best = ms;   // first launch is a warm-up
    }
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return (double)blocks * 256.0 * iters / (best * 1e-3);
}

}  // namespace

int main(int argc, char** argv)
{
    const int gib = argc > 1 ? atoi(argv[1]) : 2;
    if (gib < 1 || gib > 64) {
        fprintf(stderr, "gather_ceiling: buffer size must be 1-64 GiB\n");
        return 2;
    }
    const size_t bytes = (size_t)gib << 30;
    float4* buf = nullptr;
    float* out = nullptr;
    CK(hipMalloc((void**)&buf, bytes));
    CK(hipMemset(buf, 0, bytes));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    // 64 blocks of 4 waves per CU: several rounds of fully resident waves
    const int blocks = prop.multiProcessorCount * 64;
    CK(hipMalloc((void**)&out, sizeof(float) * (size_t)blocks * 256));
    const int iters = 200;
    // float cells: 96 B used of each 128-B record; double cells: 192 B of a 192-B record
    const double g32 = rate(k_chain<6, 8>, buf, (unsigned)(bytes / 128), out, blocks, iters);
    const double g64 = rate(k_chain<12, 12>, buf, (unsigned)(bytes / 192), out, blocks, iters);
    printf("{\"buffer_bytes\": %zu, \"cu\": %d, \"float_cell_gathers_per_s\": %.6e, "
           "\"double_cell_gathers_per_s\": %.6e, \"float_cell_line_bytes_per_s\": %.6e, "
           "\"double_cell_line_bytes_per_s\": %.6e}\n",
           bytes, prop.multiProcessorCount, g32, g64, g32 * 128.0, g64 * 256.0);
    CK(hipFree(buf));
    CK(hipFree(out));
    return 0;
}

// shared_div_check -- SharedDiv (tpamd_device.h), the dipole kernel's
// division of three numerators by one denominator, against the compiler's
// IEEE double division on random pairs across the range where SharedDiv is
// used: d in [1, 2^64], |x| in [2^-512, 2^33], with the range's edges drawn
// on purpose.  Prints the pair count and the number of results that differ
// in any bit; exits 1 if any does.
//
// Every thread draws its own pairs (xorshift64) and writes its count and
// first differing pair to its own slots; the loop count is fixed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tpamd_device.h"

using tpamd::SharedDiv;

namespace {

__device__ __forceinline__ unsigned long long next(unsigned long long& s)
{
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

__device__ __forceinline__ double unit(unsigned long long& s) { return (double)(next(s) >> 11) * 0x1p-53; }

__global__ void __launch_bounds__(256) k_check(int iters, unsigned long long* __restrict__ bad,
                                               double* __restrict__ first)
{
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long s = 0x9E3779B97F4A7C15ull ^ ((unsigned long long)(tid + 1) * 0xBF58476D1CE4E5B9ull);
    for (int k = 0; k < 8; ++k) next(s);
    unsigned long long nb = 0;
    for (int it = 0; it < iters; ++it) {
        // d = 1 + u 2^e, e uniform in [-60, 63]; the edges 1, 2^64 and the
        // double below 2^64 one draw in 64 each
        double d = 1.0 + unit(s) * exp2(-60.0 + 123.0 * unit(s));
        const unsigned sel = (unsigned)(next(s) & 63);
        if (sel == 0) d = 1.0;
        if (sel == 1 || d > 0x1p64) d = 0x1p64;
        if (sel == 2) d = nextafter(0x1p64, 0.0);
        // |x| = (1 + u) 2^e, e uniform in [-512, 32]; the edges 2^-512, 2^33
        double x = (1.0 + unit(s)) * exp2(-512.0 + 544.0 * unit(s));
        if (sel == 3 || x < 0x1p-512) x = 0x1p-512;
        if (sel == 4 || x > 0x1p33) x = 0x1p33;
        if (next(s) & 1) x = -x;
        SharedDiv q;
        q.init(d);
        const double a = q.div(x);
        const double b = x / d;
        if (__double_as_longlong(a) != __double_as_longlong(b)) {
            if (nb == 0) {
                first[2 * tid] = x;
                first[2 * tid + 1] = d;
            }
            ++nb;
        }
    }
    bad[tid] = nb;
}

}  // namespace

int main(int argc, char** argv)
{
    const int iters = argc > 1 ? atoi(argv[1]) : 1024;
    if (iters < 1 || iters > 65536) {
        fprintf(stderr, "shared_div_check: iterations must be in [1, 65536]\n");
        return 2;
    }
    const int blocks = 2048, threads = 256;
    const size_t n = (size_t)blocks * threads;
    unsigned long long* d_bad = nullptr;
    double* d_first = nullptr;
    if (hipMalloc((void**)&d_bad, n * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&d_first, 2 * n * sizeof(double)) != hipSuccess) {
        fprintf(stderr, "shared_div_check: allocation failed\n");
        return 2;
    }
    hipLaunchKernelGGL(k_check, dim3(blocks), dim3(threads), 0, 0, iters, d_bad, d_first);
    std::vector<unsigned long long> bad(n);
    std::vector<double> first(2 * n);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(bad.data(), d_bad, n * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(first.data(), d_first, 2 * n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "shared_div_check: kernel or copy failed\n");
        return 2;
    }
    unsigned long long total = 0;
    long at = -1;
    for (size_t i = 0; i < n; ++i) {
        total += bad[i];
        if (bad[i] && at < 0) at = (long)i;
    }
    printf("{\"pairs\": %.0f, \"differing\": %llu", (double)n * iters, total);
    if (at >= 0) printf(", \"first\": [\"%a\", \"%a\"]", first[2 * at], first[2 * at + 1]);
    printf("}\n");
    hipFree(d_bad);
    hipFree(d_first);
    return total ? 1 : 0;
}

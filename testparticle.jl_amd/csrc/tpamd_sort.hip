// tpamd_sort.hip -- partitions of the adjoint's dL/dB records by cell bin:
// the block-local LDS sort (with hipCUB's block scan) and the radix sort
// (hipCUB over rocPRIM) for grids with more bins, kept in their own
// translation unit so that the pusher source does not carry rocPRIM's
// templates.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "tpamd_adjoint.h"

namespace tpamd {

// Sorts (key, record) pairs on key bits [begin_bit, end_bit).  With
// d_temp == nullptr only the temporary storage size is written.
template <typename R>
hipError_t adj_sort_records(void* d_temp, size_t& temp_bytes, const unsigned* keys_in,
                            unsigned* keys_out, const AdjRec<R>* vals_in, AdjRec<R>* vals_out,
                            long long n, int begin_bit, int end_bit, hipStream_t stream)
{
    return hipcub::DeviceRadixSort::SortPairs(d_temp, temp_bytes, keys_in, keys_out, vals_in,
                                              vals_out, (size_t)n, begin_bit, end_bit, stream);
}

namespace {

// Sorts each block of P = AdjPart::part<R>() consecutive records by bin in
// LDS and writes it back in place; seg_start / seg_count[bin][block] get the
// block-local start and the count of each bin's segment.
template <typename R>
__global__ void __launch_bounds__(AdjPart::kThreads) k_adj_localsort(
    const unsigned* __restrict__ keys_in, const AdjRec<R>* __restrict__ vals_in, long long M,
    unsigned nbins, unsigned* __restrict__ keys_out, AdjRec<R>* __restrict__ vals_out,
    unsigned* __restrict__ seg_start, unsigned* __restrict__ seg_count)
{
    constexpr int T = AdjPart::kThreads;
    constexpr int P = AdjPart::part<R>();
    constexpr int PER = P / T;                 // records per thread
    constexpr int BPT = AdjPart::kMaxBins / T; // bin counters per thread in the scan
    using Scan = hipcub::BlockScan<unsigned, T>;
    __shared__ typename Scan::TempStorage scan_tmp;
    __shared__ unsigned cnt[AdjPart::kMaxBins];
    __shared__ unsigned start[AdjPart::kMaxBins];
    __shared__ unsigned skey[P];
    __shared__ AdjRec<R> sval[P];
    const int tid = threadIdx.x;
    const long long r0 = (long long)blockIdx.x * P;
    const int n = (int)(M - r0 < P ? M - r0 : P);
    for (int i = tid; i < (int)AdjPart::kMaxBins; i += T) cnt[i] = 0u;
    __syncthreads();
    unsigned key[PER], rank[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int idx = tid + j * T;
        key[j] = idx < n ? keys_in[r0 + idx] : 0u;
        rank[j] = idx < n ? atomicAdd(&cnt[key[j] >> AdjBin::kShift], 1u) : 0u;
    }
    __syncthreads();
    unsigned c[BPT], sc[BPT];
#pragma unroll
    for (int q = 0; q < BPT; ++q) c[q] = cnt[tid * BPT + q];
    Scan(scan_tmp).ExclusiveSum(c, sc);
#pragma unroll
    for (int q = 0; q < BPT; ++q) start[tid * BPT + q] = sc[q];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int idx = tid + j * T;
        if (idx < n) {
            const unsigned lp = start[key[j] >> AdjBin::kShift] + rank[j];
            skey[lp] = key[j];
            sval[lp] = vals_in[r0 + idx];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < n; idx += T) {
        keys_out[r0 + idx] = skey[idx];
        vals_out[r0 + idx] = sval[idx];
    }
    for (unsigned b = tid; b < nbins; b += T) {
        seg_start[(size_t)b * gridDim.x + blockIdx.x] = start[b];
        seg_count[(size_t)b * gridDim.x + blockIdx.x] = cnt[b];
    }
}

// As k_adj_bins, over block-sorted records: bin b's records are the
// segments [blk * P + seg_start[b][blk], + seg_count[b][blk]) of every block.
// The bin's row of segment counts is scanned in LDS, kChunk blocks at a time,
// and thread f takes the f-th record of the concatenated segments (found by
// binary search in the scan), so neighbouring threads read neighbouring
// records of a segment.
template <typename R>
__global__ void __launch_bounds__(AdjBin::kThreads) k_adj_bins_seg(
    const unsigned* __restrict__ keys, const AdjRec<R>* __restrict__ vals,
    const unsigned* __restrict__ seg_start, const unsigned* __restrict__ seg_count, int nblk,
    unsigned ncell, double* __restrict__ G)
{
    constexpr int T = AdjBin::kThreads;
    constexpr int kCells = AdjBin::kCells;
    constexpr int P = AdjPart::part<R>();
    constexpr int kChunk = 4096;
    constexpr int IPT = kChunk / T;
    using Scan = hipcub::BlockScan<unsigned, T>;
    __shared__ typename Scan::TempStorage scan_tmp;
    __shared__ double tile[kCells * 24];
    __shared__ unsigned pre[kChunk];
    __shared__ unsigned beg[kChunk];
    const int tid = threadIdx.x;
    const unsigned bin = blockIdx.x;
    const unsigned first = bin << AdjBin::kShift;
    for (int k = tid; k < kCells * 24; k += T) tile[k] = 0.0;
    const size_t row = (size_t)bin * (size_t)nblk;
    for (int c0 = 0; c0 < nblk; c0 += kChunk) {
        const int nb = nblk - c0 < kChunk ? nblk - c0 : kChunk;
        unsigned v[IPT];
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int i = tid * IPT + q;
            v[q] = i < nb ? seg_count[row + c0 + i] : 0u;
        }
        unsigned total;
        __syncthreads();                      // the previous chunk is done with pre / beg
        Scan(scan_tmp).ExclusiveSum(v, v, total);
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int i = tid * IPT + q;
            pre[i] = v[q];
            beg[i] = i < nb ? seg_start[row + c0 + i] : 0u;
        }
        __syncthreads();
        for (unsigned f = tid; f < total; f += T) {
            int lo = 0, hi = nb - 1;          // the last segment starting at or before f
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pre[mid] <= f) lo = mid;
                else hi = mid - 1;
            }
            const long long r = (long long)(c0 + lo) * P + beg[lo] + (f - pre[lo]);
            adj_tile_add<R>(tile, keys[r] - first, vals[r]);
        }
    }
    __syncthreads();
    for (int k = tid; k < kCells * 24; k += T) {
        const unsigned local = (unsigned)(k / 24), comp = (unsigned)(k % 24);
        const unsigned cell = first + local;
        if (cell < ncell) G[(size_t)cell * 24 + comp] += tile[comp * kCells + local];
    }
}

}  // namespace

// Launches k_adj_bins_seg, one workgroup per bin.
template <typename R>
hipError_t adj_bins_seg(const unsigned* keys, const AdjRec<R>* vals, const unsigned* seg_start,
                        const unsigned* seg_count, int nblk, unsigned nbins, unsigned ncell,
                        double* G, hipStream_t stream)
{
    hipLaunchKernelGGL((k_adj_bins_seg<R>), dim3(nbins), dim3(AdjBin::kThreads), 0, stream, keys, vals,
                       seg_start, seg_count, nblk, ncell, G);
    return hipGetLastError();
}

template hipError_t adj_bins_seg<float>(const unsigned*, const AdjRec<float>*, const unsigned*,
                                        const unsigned*, int, unsigned, unsigned, double*, hipStream_t);
template hipError_t adj_bins_seg<double>(const unsigned*, const AdjRec<double>*, const unsigned*,
                                         const unsigned*, int, unsigned, unsigned, double*, hipStream_t);

// Launches k_adj_localsort over the M records (nbins <= AdjPart::kMaxBins).
template <typename R>
hipError_t adj_localsort_records(const unsigned* keys_in, const AdjRec<R>* vals_in, long long M,
                                 unsigned nbins, unsigned* keys_out, AdjRec<R>* vals_out,
                                 unsigned* seg_start, unsigned* seg_count, hipStream_t stream)
{
    constexpr int P = AdjPart::part<R>();
    const long long nblk = (M + P - 1) / P;
    hipLaunchKernelGGL((k_adj_localsort<R>), dim3((unsigned)nblk), dim3(AdjPart::kThreads), 0, stream,
                       keys_in, vals_in, M, nbins, keys_out, vals_out, seg_start, seg_count);
    return hipGetLastError();
}

template hipError_t adj_localsort_records<float>(const unsigned*, const AdjRec<float>*, long long,
                                                 unsigned, unsigned*, AdjRec<float>*, unsigned*,
                                                 unsigned*, hipStream_t);
template hipError_t adj_localsort_records<double>(const unsigned*, const AdjRec<double>*, long long,
                                                  unsigned, unsigned*, AdjRec<double>*, unsigned*,
                                                  unsigned*, hipStream_t);

template hipError_t adj_sort_records<float>(void*, size_t&, const unsigned*, unsigned*,
                                            const AdjRec<float>*, AdjRec<float>*, long long, int,
                                            int, hipStream_t);
template hipError_t adj_sort_records<double>(void*, size_t&, const unsigned*, unsigned*,
                                             const AdjRec<double>*, AdjRec<double>*, long long,
                                             int, int, hipStream_t);

}  // namespace tpamd

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

// Index of the f-th record of a bin's concatenated segments within a chunk
// of blocks starting at block c0: pre[] is the exclusive scan of the
// segment counts and beg[] the segments' block-local starts (nb entries).
template <int P>
__device__ __forceinline__ long long seg_record(const unsigned* pre, const unsigned* beg, int nb, int c0,
                                                unsigned f)
{
    int lo = 0, hi = nb - 1;                  // the last segment starting at or before f
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= f) lo = mid;
        else hi = mid - 1;
    }
    return (long long)(c0 + lo) * P + beg[lo] + (f - pre[lo]);
}

// The sums of k_adj_bins, over block-sorted records: bin b's records are the
// segments [blk * P + seg_start[b][blk], + seg_count[b][blk]) of every block.
// The bin's row of segment counts is scanned in LDS, kChunk blocks at a time,
// and thread f takes the f-th record of the concatenated segments, so
// neighbouring threads read neighbouring records of a segment.
//
// Double records are summed into the 24 corner moments of their cell as LDS
// atomics (ds_add_f64) on a tile of the bin's cells.  For float records that
// capped the kernel at the LDS atomic rate, so there each batch of 4,096
// records is grouped by cell in LDS with a counting sort (one ds_add_u32 per
// record for its rank in the cell, a block scan of the cell counts), and two
// owner threads per cell, one for corners 0-3 and one for 4-7, sum the
// cell's records into double registers: 263 us against 347 us a call at the
// 64^3 adjoint workload.  Staging double records the same way (2,048 a batch,
// 64-bit LDS accesses) lost: 412 us against 354 us.  Either way each product
// is formed in the record's precision and converted to double; only the
// order of the double sums differs.  At the end the workgroup adds its sums
// to the per-cell gradient G[cell][24], which it alone writes.
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
    constexpr bool OWN = sizeof(R) == 4;      // owner sums (float) or LDS atomics (double)
    constexpr int RB = 4096;                  // owner sums: records grouped per batch
    constexpr int RPT = RB / T;
    static_assert(T == 2 * kCells, "two owner threads per cell");
    // one LDS buffer for both: the tile, or 7 staged columns of RB records
    constexpr int kLds = OWN ? 7 * RB * (int)sizeof(float) / (int)sizeof(double) : kCells * 24;
    static_assert(!OWN || 7 * RB * sizeof(float) >= kCells * 24 * sizeof(double), "the tile fits");
    using Scan = hipcub::BlockScan<unsigned, T>;
    __shared__ typename Scan::TempStorage scan_tmp;
    __shared__ unsigned pre[kChunk];
    __shared__ unsigned beg[kChunk];
    __shared__ unsigned cnt[OWN ? kCells : 1];
    __shared__ unsigned off[OWN ? kCells : 1];
    // the atomics' tile[(3 q + c) * kCells + cell], or the owners' staged
    // records sb[v * RB + slot] (v: Bbar 0-2, weights 3-5, exponent 6)
    __shared__ double lds[kLds];
    double* tile = lds;
    float* sb = reinterpret_cast<float*>(lds);
    int* sbe = reinterpret_cast<int*>(lds) + 6 * RB;
    const int tid = threadIdx.x;
    const unsigned first = blockIdx.x << AdjBin::kShift;
    const int own = tid >> 1, half = tid & 1;
    double acc[OWN ? 12 : 1];
    if constexpr (OWN) {
#pragma unroll
        for (int j = 0; j < 12; ++j) acc[j] = 0.0;
        if (tid < kCells) cnt[tid] = 0u;
    } else {
        for (int k = tid; k < kCells * 24; k += T) tile[k] = 0.0;
    }
    const size_t row = (size_t)blockIdx.x * (size_t)nblk;
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
        if constexpr (!OWN) {
            for (unsigned f = tid; f < total; f += T) {
                const long long r = seg_record<P>(pre, beg, nb, c0, f);
                adj_tile_add<R>(tile, keys[r] - first, vals[r]);
            }
        } else {
            for (unsigned base = 0; base < total; base += RB) {
                // stage: each thread's records, and their ranks within their cells
                unsigned loc[RPT], rank[RPT];
                AdjRec<R> rv[RPT];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const unsigned f = base + (unsigned)(tid + q * T);
                    loc[q] = kCells;
                    rank[q] = 0u;
                    if (f < total) {
                        const long long r = seg_record<P>(pre, beg, nb, c0, f);
                        loc[q] = keys[r] - first;
                        rv[q] = vals[r];
                        rank[q] = atomicAdd(&cnt[loc[q]], 1u);
                    }
                }
                __syncthreads();
                unsigned cc = tid < kCells ? cnt[tid] : 0u, oo;
                Scan(scan_tmp).ExclusiveSum(cc, oo);
                if (tid < kCells) off[tid] = oo;
                __syncthreads();
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    if (loc[q] < (unsigned)kCells) {
                        const unsigned d = off[loc[q]] + rank[q];
                        sb[d] = rv[q].bb[0];
                        sb[RB + d] = rv[q].bb[1];
                        sb[2 * RB + d] = rv[q].bb[2];
                        sb[3 * RB + d] = rv[q].w[0];
                        sb[4 * RB + d] = rv[q].w[1];
                        sb[5 * RB + d] = rv[q].w[2];
                        if constexpr (OWN) sbe[d] = rv[q].e;
                    }
                }
                __syncthreads();
                // the owners sum their cell's records: corner q = (qx, qy, qz) in bits 0..2
                const unsigned i0 = off[own], i1 = i0 + cnt[own];
                for (unsigned i = i0; i < i1; ++i) {
                    const float b0 = sb[i], b1 = sb[RB + i], b2 = sb[2 * RB + i];
                    const float w0 = sb[3 * RB + i], w1 = sb[4 * RB + i], w2 = sb[5 * RB + i];
                    const double sc = ldexp(1.0, sbe[i]);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int q = 4 * half + k;
                        const float W = ((q & 1) ? w0 : 1.0f - w0) * (((q >> 1) & 1) ? w1 : 1.0f - w1) *
                                        ((q >> 2) ? w2 : 1.0f - w2);
                        acc[3 * k] += (double)(W * b0) * sc;
                        acc[3 * k + 1] += (double)(W * b1) * sc;
                        acc[3 * k + 2] += (double)(W * b2) * sc;
                    }
                }
                __syncthreads();              // cnt / off / sb are read
                if (tid < kCells) cnt[tid] = 0u;
                __syncthreads();
            }
        }
    }
    if constexpr (OWN) {
        const unsigned cell = first + (unsigned)own;
        if (cell < ncell) {
#pragma unroll
            for (int j = 0; j < 12; ++j) G[(size_t)cell * 24 + 12 * half + j] += acc[j];
        }
    } else {
        __syncthreads();
        for (int k = tid; k < kCells * 24; k += T) {
            const unsigned local = (unsigned)(k / 24), comp = (unsigned)(k % 24);
            const unsigned cell = first + local;
            if (cell < ncell) G[(size_t)cell * 24 + comp] += tile[comp * kCells + local];
        }
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

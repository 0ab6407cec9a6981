// tpamd_adjoint.h -- reverse-mode gradient of a trajectory loss through the
// fixed-step Boris pusher (standard Boris, E = 0, gridded B).
//
// Loss  L = (1/N) sum_i |x_i(T) - x*_i|^2 over final positions.
// Outputs dL/dB at every grid node (3 components) and dL/d(x0, v0).
//
// Forward map of one step (boris.jl:445-478 with E = 0):
//   t = h B(x_n),  s = 2 t / (1 + t.t),  c = u + u x t,
//   u' = u + c x s,  x_{n+1} = x_n + u' dt          (h = q2m dt / 2)
// Vector-Jacobian product, with g the adjoint of u' (plus dt times the
// adjoint of x_{n+1}):
//   cbar = s x g,  ubar = g + cbar + t x cbar,
//   tbar = cbar x u + (2/D) sbar - (4/D^2) t (t.sbar),  sbar = g x c,  D = 1 + t.t
//   Bbar = h tbar  -> scattered to the 8 corner nodes with the trilinear weights,
//   xbar += (dB/dx)^T Bbar  (trilinear gradient).
// The initial back half-step v(-1/2) = U(v0, x0, -dt/2) (boris.jl:514) is
// differentiated the same way with h = -q2m dt / 4.
//
// Memory: FP64 checkpoints of the state every K steps; the reverse sweep
// recomputes each K-step segment into a per-particle buffer (structure of
// arrays, particle index fastest) and walks it backwards.
//
// dL/dB without global atomics: each reverse step writes one record (the
// cell it sampled, Bbar and the three trilinear weights).  Per segment the
// records are partitioned by cell bin (a block-local sort in LDS, or a radix
// sort for grids with more bins than its counters hold), and one workgroup
// per bin sums its records into per-cell corner moments (8 corners x 3
// components), which it alone adds into the global per-cell gradient.  A
// last pass folds the per-cell corner moments into the node gradient.
#pragma once

#include "tpamd_grid.h"

namespace tpamd {

// one reverse-step contribution: Bbar and the cell weights (wx, wy, wz)
template <typename R>
struct AdjRec {
    R bb[3];
    R w[3];
};

// Float records carry a power-of-two exponent: Bbar = bb * 2^e.  The exact
// gradient of a chaotic flow grows like exp(lambda T) (|dL/dB| ~ 1e63 after
// 2,048 steps of the 64^3 turbulent box), far past the float range, so the
// float adjoint state is kept as a mantissa vector with one exponent per
// particle, renormalised every step, and each record inherits its
// particle's exponent.  The bin sums apply it in double.
template <>
struct AdjRec<float> {
    float bb[3];
    float w[3];
    int e;
};

__device__ __forceinline__ double adj_rec_scale(const AdjRec<float>& v) { return ldexp(1.0, v.e); }
__device__ __forceinline__ double adj_rec_scale(const AdjRec<double>&) { return 1.0; }

// cells per bin: the bin's LDS tile holds 24 doubles per cell, 96 KiB.  The
// tile is double for both precisions: on gfx950 an LDS float atomic add
// (ds_add_f32) runs at ~1/7 the rate of the double one (ds_add_f64).
struct AdjBin {
    static constexpr int kShift = 9;
    static constexpr int kCells = 1 << kShift;
    static constexpr int kThreads = 1024;
};

// Partition of the records by bin (key >> kShift), used while the bin
// counters fit in LDS: each block of kPart consecutive records is sorted by
// bin in LDS and written back in place (contiguous, coalesced stores), with
// the count and block-local start of every (bin, block) segment in [bin][block]
// tables; the bin kernel then reads its bin's segment from every block.  One
// read and one write of the records, against two radix passes and a global
// scatter.  Records keep their bin, not their order within a bin (the bin
// sums are order-free LDS atomics already).
struct AdjPart {
    static constexpr int kThreads = 1024;
    static constexpr unsigned kMaxBins = 2048;
    // records per block: the block's records and keys staged in LDS
    template <typename R>
    static constexpr int part() { return sizeof(R) == 4 ? 4096 : 2048; }
};

struct AdjParams {
    const double* targets;    // [N][3]
    double* ckpt;             // [nseg][9][N]: frame position (3), wrap (3), velocity (3)
    void* seg;                // [K][9 or 6][N] recompute buffer (float or double)
    void* state;              // [6][N] adjoint state (xbar, ubar) between segment launches
    int* state_exp;           // [N] its power-of-two exponent (float adjoint state only)
    unsigned* rec_key;        // [(K + 1) N] record keys: the cell index
    void* rec_val;            // [(K + 1) N] AdjRec<R>; rows without a step hold a zero record
    unsigned ncell;           // cells in the grid
    double* grad_xv0;         // [N][6]
    double* loss;             // scalar sum
    double* xfinal;           // [N][3] final positions
    int K;                    // checkpoint interval
    int nseg;
    int nc[3];                // cells per axis (n periodic, n - 1 otherwise)
};

// ---------------------------------------------------------------------------
// generic-precision vector helpers
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void xcross(const R a[3], const R b[3], R o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename R>
__device__ __forceinline__ R xdot(const R a[3], const R b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// VJP of u' = u + (u + u x t) x s(t)
template <typename R>
__device__ __forceinline__ void boris_vjp(const R u[3], const R t[3], const R g[3], R ubar[3],
                                          R tbar[3])
{
    const R D = (R)1 + xdot(t, t);
    const R iD = (R)1 / D;
    R s[3], c[3], uxt[3], cb[3], txcb[3], cbxu[3], sb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = (R)2 * t[k] * iD;
    xcross(u, t, uxt);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = u[k] + uxt[k];
    xcross(s, g, cb);
    xcross(t, cb, txcb);
#pragma unroll
    for (int k = 0; k < 3; ++k) ubar[k] = g[k] + cb[k] + txcb[k];
    xcross(cb, u, cbxu);
    xcross(g, c, sb);
    const R ts = xdot(t, sb);
#pragma unroll
    for (int k = 0; k < 3; ++k) tbar[k] = cbxu[k] + (R)2 * iD * sb[k] - (R)4 * iD * iD * ts * t[k];
}

// Corner nodes (flat node indices) and per-axis weights of the cell holding
// cell coordinate s (same location rules as the forward gather).
template <bool PER, typename R>
__device__ __forceinline__ bool adj_locate(const DevField& f, const R s[3], int i0[3], int i1[3],
                                           R w[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        bool ok;
        if constexpr (sizeof(R) == 4) ok = locate_f32<PER>(f, d, s[d], i0[d], i1[d], w[d]);
        else ok = locate_f64<PER>(f, d, s[d], i0[d], i1[d], w[d]);
        if (!ok) return false;
    }
    return true;
}

template <typename T, typename R>
__device__ __forceinline__ void adj_corners(const DevField& f, const int i0[3], const int i1[3],
                                            R c[8][3], unsigned node[8])
{
    const unsigned nx = (unsigned)f.n[0], nxy = (unsigned)f.n[0] * (unsigned)f.n[1];
    const T* base = (const T*)f.data;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const unsigned id = (unsigned)((q & 1) ? i1[0] : i0[0]) + nx * (unsigned)(((q >> 1) & 1) ? i1[1] : i0[1]) +
                            nxy * (unsigned)((q >> 2) ? i1[2] : i0[2]);
        node[q] = id;
        const T* p = base + 3u * id;
        c[q][0] = (R)p[0]; c[q][1] = (R)p[1]; c[q][2] = (R)p[2];
    }
}

// the 8 corner vectors from the cell-replicated copy (corner q = (qx, qy, qz)
// in bits 0..2, 3 components each, as the cell builder stores them)
template <typename T, typename R>
__device__ __forceinline__ void adj_corners_cell(const DevField& f, const int i0[3], R c[8][3])
{
    const T* rec = (const T*)((const char*)f.cells + cell_offset(f, i0));
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        c[q][0] = (R)rec[3 * q];
        c[q][1] = (R)rec[3 * q + 1];
        c[q][2] = (R)rec[3 * q + 2];
    }
}

// trilinear value, blended in the forward kernels' order (lerp along x, y, z)
template <typename R>
__device__ __forceinline__ void adj_value(const R c[8][3], const R w[3], R B[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if constexpr (sizeof(R) == 8) e[q] = lerp_f64(c[2 * q][k], c[2 * q + 1][k], w[0]);
            else e[q] = lerp_f32(c[2 * q][k], c[2 * q + 1][k], w[0]);
        }
        if constexpr (sizeof(R) == 8)
            B[k] = lerp_f64(lerp_f64(e[0], e[1], w[1]), lerp_f64(e[2], e[3], w[1]), w[2]);
        else
            B[k] = lerp_f32(lerp_f32(e[0], e[1], w[1]), lerp_f32(e[2], e[3], w[1]), w[2]);
    }
}

// (dB/dx)^T Bbar, the trilinear gradient through the cell weights
template <typename R>
__device__ __forceinline__ void adj_xgrad(const DevField& f, const R c[8][3], const R w[3],
                                          const R Bb[3], R xb[3])
{
    R g[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const R ax = (q & 1) ? w[0] : (R)1 - w[0];
        const R ay = ((q >> 1) & 1) ? w[1] : (R)1 - w[1];
        const R az = (q >> 2) ? w[2] : (R)1 - w[2];
        const R cb = c[q][0] * Bb[0] + c[q][1] * Bb[1] + c[q][2] * Bb[2];
        const R sx = (q & 1) ? (R)1 : (R)-1;
        const R sy = ((q >> 1) & 1) ? (R)1 : (R)-1;
        const R sz = (q >> 2) ? (R)1 : (R)-1;
        g[0] += sx * ay * az * cb;
        g[1] += ax * sy * az * cb;
        g[2] += ax * ay * sz * cb;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) xb[d] = g[d] * (R)f.inv_h[d];
}

// one record's corner moments into the bin's LDS tile (component-major)
template <typename R>
__device__ __forceinline__ void adj_tile_add(double* tile, unsigned local, const AdjRec<R>& v)
{
    constexpr int kCells = AdjBin::kCells;
    double* t = tile + local;
    const double sc = adj_rec_scale(v);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const R W = ((q & 1) ? v.w[0] : (R)1 - v.w[0]) *
                    (((q >> 1) & 1) ? v.w[1] : (R)1 - v.w[1]) *
                    ((q >> 2) ? v.w[2] : (R)1 - v.w[2]);
        atomicAdd(t + (3 * q) * kCells, (double)(W * v.bb[0]) * sc);
        atomicAdd(t + (3 * q + 1) * kCells, (double)(W * v.bb[1]) * sc);
        atomicAdd(t + (3 * q + 2) * kCells, (double)(W * v.bb[2]) * sc);
    }
}

// Sum the records of one cell bin (keys sorted on the bin bits; the bin's
// range is found by binary search) into an LDS tile of per-cell corner
// moments, then add the tile into the global per-cell gradient G[cell][24],
// which this workgroup alone writes.  Corner q = (qx, qy, qz) in bits 0..2.
// The tile is component-major, tile[(3 q + c) * kCells + cell], so the lanes
// of one LDS atomic spread over the banks.
template <typename R>
__global__ void __launch_bounds__(AdjBin::kThreads) k_adj_bins(const unsigned* __restrict__ keys,
                                                               const AdjRec<R>* __restrict__ vals,
                                                               long long M, unsigned ncell,
                                                               double* __restrict__ G)
{
    constexpr int kShift = AdjBin::kShift, kCells = AdjBin::kCells;
    __shared__ double tile[kCells * 24];
    const unsigned bin = blockIdx.x;
    const unsigned first = bin << kShift;
    auto lower = [&](unsigned target) {
        long long lo = 0, hi = M;
        while (lo < hi) {
            const long long mid = (lo + hi) >> 1;
            if (keys[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const long long r0 = lower(first), r1 = lower(first + (unsigned)kCells);
    for (int k = threadIdx.x; k < kCells * 24; k += blockDim.x) tile[k] = 0.0;
    __syncthreads();
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) adj_tile_add<R>(tile, keys[r] - first, vals[r]);
    __syncthreads();
    for (int k = threadIdx.x; k < kCells * 24; k += blockDim.x) {
        const unsigned local = (unsigned)(k / 24), comp = (unsigned)(k % 24);
        const unsigned cell = first + local;
        if (cell < ncell) G[(size_t)cell * 24 + comp] += tile[comp * kCells + local];
    }
}

// node gradient [k][j][i][3] from the per-cell corner moments: node (i, j, k)
// is corner q of cell (i - qx, j - qy, k - qz), wrapped on periodic grids
template <typename R>
__global__ void __launch_bounds__(256) k_adj_nodes(const R* __restrict__ G, int nx, int ny, int nz,
                                                   int cx, int cy, int cz, int periodic,
                                                   R* __restrict__ out)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nn = (long long)nx * ny * nz;
    if (id >= nn) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((long long)nx * ny));
    R acc[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        int ci = i - (q & 1), cj = j - ((q >> 1) & 1), ck = k - (q >> 2);
        if (periodic) {
            ci += ci < 0 ? nx : 0;
            cj += cj < 0 ? ny : 0;
            ck += ck < 0 ? nz : 0;
        } else if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) {
            continue;
        }
        const R* g = G + ((size_t)ci + (size_t)cx * ((size_t)cj + (size_t)cy * (size_t)ck)) * 24 + 3 * q;
        acc[0] += g[0];
        acc[1] += g[1];
        acc[2] += g[2];
    }
    out[3 * id] = acc[0];
    out[3 * id + 1] = acc[1];
    out[3 * id + 2] = acc[2];
}

}  // namespace tpamd

// tpamd_adj.hip -- the checkpointed adjoint of the fixed-step Boris solve
// (kernels and host driver; see tpamd_adjoint.h).
#include "tpamd_pusher.h"
#include "tpamd_adjoint.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Adjoint (see tpamd_adjoint.h).  The forward pass reuses the specialised
// Boris policies, so its trajectory is bit-identical to a solve of the same
// problem; FP64 checkpoints are taken every K steps.
// ---------------------------------------------------------------------------
// FP32C: float storage (the host checks it); FP64: storage T.  CELLS: the
// gathers read the grid's cell-replicated copy (one line per gather).
template <bool F32, bool PER, typename T = double, bool CELLS = false>
using AdjState = typename std::conditional<F32, S32G<kEUniform, 0, PER, CELLS>,
                                           S64G<kEUniform, 0, PER, T, CELLS>>::type;

template <bool F32, bool PER, typename T, bool CELLS>
__device__ __forceinline__ void adj_ckpt_store(const AdjParams& a, long long N, int seg, long long i,
                                               const AdjState<F32, PER, T, CELLS>& s)
{
    double* c = a.ckpt + (size_t)seg * 9 * N + i;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (F32) {
            c[(size_t)d * N] = (double)s.pos.hi[d] + (double)s.pos.lo[d];
            c[(size_t)(3 + d) * N] = (double)s.pos.wrap[d];
        } else {
            c[(size_t)d * N] = s.pos.x[d];
            c[(size_t)(3 + d) * N] = 0.0;
        }
        c[(size_t)(6 + d) * N] = (double)s.v[d];
    }
}

template <bool F32, bool PER, typename T, bool CELLS>
__device__ __forceinline__ void adj_ckpt_load(const AdjParams& a, long long N, int seg, long long i,
                                              AdjState<F32, PER, T, CELLS>& s)
{
    const double* c = a.ckpt + (size_t)seg * 9 * N + i;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double r = c[(size_t)d * N];
        if constexpr (F32) {
            s.pos.hi[d] = (float)r;
            s.pos.lo[d] = (float)(r - (double)s.pos.hi[d]);
            s.pos.wrap[d] = (int)c[(size_t)(3 + d) * N];
        } else {
            s.pos.x[d] = r;
        }
        s.v[d] = (typename AdjState<F32, PER, T, CELLS>::R)c[(size_t)(6 + d) * N];
    }
}

// cell coordinates of the current position, as the forward gather computes them
template <bool F32, bool PER, typename T, bool CELLS>
__device__ __forceinline__ void adj_cellcoord(const KParams& p, const AdjState<F32, PER, T, CELLS>& s,
                                              typename AdjState<F32, PER, T, CELLS>::R sc[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (F32) sc[d] = fmaf(s.pos.hi[d], p.B.inv_hf[d], s.pos.lo[d] * p.B.inv_hf[d]);
        else sc[d] = (s.pos.x[d] - p.B.origin[d]) * p.B.inv_h[d];
    }
}

// a record that adds nothing, keyed to a cell that spreads such records over
// the bins
template <typename R>
__device__ __forceinline__ void adj_zero_record(const AdjParams& a, size_t rec)
{
    a.rec_key[rec] = (unsigned)(rec % (size_t)a.ncell);
    AdjRec<R> v;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v.bb[k] = (R)0;
        v.w[k] = (R)0;
    }
    if constexpr (sizeof(R) == 4) v.e = 0;
    ((AdjRec<R>*)a.rec_val)[rec] = v;
}

// Float adjoint state: (xbar, ubar) = (xb, ub) * 2^ex.  Rescales the
// mantissas so that the largest component lies in [0.5, 1) (exact powers of
// two; a zero or non-finite state is left as it is, and a non-finite one
// reaches the outputs, whose scan reports it).
__device__ __forceinline__ void adj_renorm(float xb[3], float ub[3], int& ex)
{
    const float m = fmaxf(fmaxf(fmaxf(fabsf(xb[0]), fabsf(xb[1])), fmaxf(fabsf(xb[2]), fabsf(ub[0]))),
                          fmaxf(fabsf(ub[1]), fabsf(ub[2])));
    if (m > 0.0f && m <= 3.4028235e38f) {
        int k;
        (void)frexpf(m, &k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xb[c] = ldexpf(xb[c], -k);
            ub[c] = ldexpf(ub[c], -k);
        }
        ex += k;
    }
}

// one reverse step at the state held in s: (xbar, ubar) <- VJP of the step
// that started there, with rotation factor h and position step dt.  The
// step's dL/dB contribution goes to record slot `rec` (cell, Bbar, weights).
template <bool F32, bool PER, typename T, bool CELLS>
__device__ __forceinline__ void adj_step_back(const KParams& p, const AdjParams& a,
                                              const AdjState<F32, PER, T, CELLS>& s, const typename AdjState<F32, PER, T, CELLS>::R u[3],
                                              double h, double dt, typename AdjState<F32, PER, T, CELLS>::R xb[3],
                                              typename AdjState<F32, PER, T, CELLS>::R ub[3], int& ex,
                                              bool with_dt, size_t rec)
{
    using R = typename AdjState<F32, PER, T, CELLS>::R;
    R sc[3], w[3], c[8][3], B[3];
    int i0[3], i1[3];
    unsigned node[8];
    adj_cellcoord<F32, PER, T, CELLS>(p, s, sc);
    if (!adj_locate<PER, R>(p.B, sc, i0, i1, w)) {
        adj_zero_record<R>(a, rec);
        return;
    }
    if constexpr (CELLS) adj_corners_cell<T, R>(p.B, i0, c);
    else adj_corners<T, R>(p.B, i0, i1, c, node);
    adj_value<R>(c, w, B);
    const R hh = (R)h;
    R t[3], g[3], tb[3], un[3], Bb[3], xg[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t[k] = hh * B[k];
        g[k] = with_dt ? ub[k] + (R)dt * xb[k] : ub[k];
    }
    boris_vjp<R>(u, t, g, un, tb);
#pragma unroll
    for (int k = 0; k < 3; ++k) Bb[k] = hh * tb[k];
    adj_xgrad<R>(p.B, c, w, Bb, xg);
    if constexpr (!PER) {
        // outside the node box a clamped field does not change along that axis
        if (p.B.extrap == kExtrapClamp) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (!(sc[k] >= (R)0 && sc[k] <= (R)(p.B.n[k] - 1))) xg[k] = (R)0;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xb[k] += xg[k];
        ub[k] = un[k];
    }
    a.rec_key[rec] = (unsigned)i0[0] + (unsigned)a.nc[0] * ((unsigned)i0[1] + (unsigned)a.nc[1] * (unsigned)i0[2]);
    AdjRec<R> v;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v.bb[k] = Bb[k];
        v.w[k] = w[k];
    }
    if constexpr (F32) {
        v.e = ex;                  // Bbar carries the state's scale at this step
        adj_renorm(xb, ub, ex);
    }
    ((AdjRec<R>*)a.rec_val)[rec] = v;
}

template <bool F32, bool PER, typename T, bool CELLS>
__global__ void __launch_bounds__(256) k_adj_forward(const KParams p, const AdjParams a)
{
    using S = AdjState<F32, PER, T, CELLS>;
    using R = typename S::R;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double lsum = 0.0;
    if (i < p.N) {
        S s;
        s.load(p, i);
        const double dt = p.dt;
        s.set_dt(dt);
        R vt[3];
        s.template U<0>(p, s.v, -0.5 * dt, p.t0, vt);            // boris.jl:514
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
        double t = fma(-0.5, dt, p.t0);
        for (long long n = 0; n < p.nt; ++n) {
            if (n % a.K == 0) adj_ckpt_store<F32, PER, T, CELLS>(a, p.N, (int)(n / a.K), i, s);
            t += dt;
            s.template U<0>(p, s.v, dt, t, vt);
#pragma unroll
            for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
            typename S::Pos xn;
            do_step(p, s, s.v, dt, xn);
            s.pos = xn;
        }
        double x[3];
        s.abs_pos(p, s.pos, x);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a.xfinal[3 * i + c] = x[c];
            const double d = x[c] - a.targets[3 * i + c];
            lsum += d * d;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(a.loss, lsum / (double)p.N);
}

// Reverse sweep over segment `seg` (launched for seg = nseg-1 .. 0): reload
// the adjoint state, recompute the segment from its checkpoint, walk it
// backwards writing one record per step into rows 0..K-1 of the record
// buffer (row K holds the t0 half-step on segment 0; unused rows are marked
// invalid), and store the adjoint state for the next launch.
template <bool F32, bool PER, typename T, bool CELLS>
__global__ void __launch_bounds__(256) k_adj_reverse(const KParams p, const AdjParams a, int seg)
{
    using S = AdjState<F32, PER, T, CELLS>;
    using R = typename S::R;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.N) return;
    const long long N = p.N;
    const double dt = p.dt;
    const double h = (p.q2m * 0.5) * dt;
    R* st = (R*)a.state;
    R xb[3], ub[3];
    int ex = 0;                                                // float state: scale 2^ex
    if (seg == a.nseg - 1) {
        double x0[3], m = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x0[c] = 2.0 / (double)N * (a.xfinal[3 * i + c] - a.targets[3 * i + c]);
            m = fmax(m, fabs(x0[c]));
        }
        if constexpr (F32) {
            if (m > 0.0 && m <= 1.0e300) (void)frexp(m, &ex);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xb[c] = (R)(F32 ? ldexp(x0[c], -ex) : x0[c]);
            ub[c] = (R)0;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xb[c] = st[(size_t)c * N + i];
            ub[c] = st[(size_t)(3 + c) * N + i];
        }
        if constexpr (F32) ex = a.state_exp[i];
    }
    constexpr int NC = F32 ? 9 : 6;
    R* buf = (R*)a.seg;
    const long long n0 = (long long)seg * a.K;
    const int len = (int)((p.nt - n0) < a.K ? (p.nt - n0) : a.K);
    S s;
    adj_ckpt_load<F32, PER, T, CELLS>(a, N, seg, i, s);
    s.set_dt(dt);
    R vt[3];
    for (int j = 0; j < len; ++j) {                           // recompute the segment
        R* b = buf + (size_t)j * NC * N + i;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if constexpr (F32) {
                b[(size_t)d * N] = s.pos.hi[d];
                b[(size_t)(3 + d) * N] = s.pos.lo[d];
                b[(size_t)(6 + d) * N] = s.v[d];
            } else {
                b[(size_t)d * N] = s.pos.x[d];
                b[(size_t)(3 + d) * N] = s.v[d];
            }
        }
        s.template U<0>(p, s.v, dt, 0.0, vt);
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
        typename S::Pos xn;
        do_step(p, s, s.v, dt, xn);
        s.pos = xn;
    }
    for (int j = a.K - 1; j >= len; --j) adj_zero_record<R>(a, (size_t)j * N + i);
    for (int j = len - 1; j >= 0; --j) {                      // and walk it backwards
        const R* b = buf + (size_t)j * NC * N + i;
        R u[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if constexpr (F32) {
                s.pos.hi[d] = b[(size_t)d * N];
                s.pos.lo[d] = b[(size_t)(3 + d) * N];
                u[d] = b[(size_t)(6 + d) * N];
            } else {
                s.pos.x[d] = b[(size_t)d * N];
                u[d] = b[(size_t)(3 + d) * N];
            }
        }
        adj_step_back<F32, PER, T, CELLS>(p, a, s, u, h, dt, xb, ub, ex, true, (size_t)j * N + i);
    }
    const size_t last = (size_t)a.K * N + i;
    if (seg == 0) {
        // the initial back half-step u(-1/2) = U(v0, x0, -dt/2) (boris.jl:514)
        S s0;
        s0.load(p, i);
        R v0[3] = {s0.v[0], s0.v[1], s0.v[2]};
        adj_step_back<F32, PER, T, CELLS>(p, a, s0, v0, (p.q2m * 0.5) * (-0.5 * dt), 0.0, xb, ub, ex, false,
                                          last);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a.grad_xv0[6 * i + c] = ldexp((double)xb[c], ex);
            a.grad_xv0[6 * i + 3 + c] = ldexp((double)ub[c], ex);
        }
    } else {
        adj_zero_record<R>(a, last);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            st[(size_t)c * N + i] = xb[c];
            st[(size_t)(3 + c) * N + i] = ub[c];
        }
        if constexpr (F32) a.state_exp[i] = ex;
    }
}

__global__ void k_zero_f32(float* p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_zero_f64(double* p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// counts the entries of a[0..n) that are not finite into *count
__global__ void __launch_bounds__(256) k_count_nonfinite(const double* __restrict__ a, long long n,
                                                         unsigned long long* count)
{
    unsigned long long c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        c += isfinite(a[i]) ? 0ull : 1ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

template <bool F32, bool PER, typename T, bool CELLS>
int adjoint_run(tpamd_ctx* ctx, KParams p, const tpamd_field* B, const double* xv0, int64_t N,
                const double* targets, int K, int64_t nt, double* loss, double* dL_dB,
                double* dL_dxv0, tpamd_stats* stats)
{
    using R = typename AdjState<F32, PER, T, CELLS>::R;
    const int nseg = (int)((nt + K - 1) / K);
    const int ncomp = F32 ? 9 : 6;
    const size_t rs = sizeof(R);
    const int nx = B->d.n[0], ny = B->d.n[1], nz = B->d.n[2];
    const size_t nnodes = (size_t)nx * ny * nz;
    int nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = PER ? B->d.n[k] : B->d.n[k] - 1;
    if (nc[0] < 1 || nc[1] < 1 || nc[2] < 1)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "the adjoint's grid needs at least one cell per axis");
    const double ncell_d = (double)nc[0] * nc[1] * nc[2];
    if (ncell_d >= 1073741824.0) return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the adjoint supports fewer than 2^30 cells");
    const unsigned ncell = (unsigned)ncell_d;
    const long long M = (long long)(K + 1) * N;              // record rows 0..K
    if ((double)(K + 1) * (double)N >= 2147483648.0)
        return fail(ctx, TPAMD_ERR_CAPACITY, "too many adjoint records: lower the checkpoint interval or N");
    // records are partitioned on the bin bits [kShift, cellbits) of the cell
    // index; a grid of one bin needs no partition
    constexpr int kShift = AdjBin::kShift;
    int cellbits = 0;
    while ((1u << cellbits) < ncell) ++cellbits;
    const bool partition = cellbits > kShift;
    const unsigned nbins = (ncell + AdjBin::kCells - 1) / AdjBin::kCells;
    // the block-local partition while its bin counters fit in LDS, else a radix sort
    const bool local = partition && nbins <= AdjPart::kMaxBins;
    const long long nblk = (M + AdjPart::part<R>() - 1) / AdjPart::part<R>();
    const long long ncnt = local ? (long long)nbins * nblk : 0;
    size_t temp_bytes = 0;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (partition && !local)
        HIP_TRY(ctx, adj_sort_records<R>(nullptr, temp_bytes, nullptr, nullptr, nullptr, nullptr, M, kShift,
                                         cellbits, ctx->stream));
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_xv = al(sizeof(double) * 6 * N), b_tg = al(sizeof(double) * 3 * N),
                 b_ck = al(sizeof(double) * 9 * (size_t)nseg * N), b_sg = al(rs * ncomp * (size_t)K * N),
                 b_st = al(rs * 6 * (size_t)N), b_se = al(sizeof(int) * (size_t)N), b_key = al(sizeof(unsigned) * (size_t)M),
                 b_val = al(sizeof(AdjRec<R>) * (size_t)M), b_tmp = al(std::max<size_t>(temp_bytes, 1)),
                 b_G = al(sizeof(double) * 24 * (size_t)ncell), b_out = al(sizeof(double) * 3 * nnodes),
                 b_gx = al(sizeof(double) * 6 * N), b_ls = al(sizeof(double)), b_xf = al(sizeof(double) * 3 * N),
                 b_cnt = al(sizeof(unsigned) * (size_t)std::max<long long>(ncnt, 1)),
                 b_nf = al(sizeof(unsigned long long));
    const size_t total = b_xv + b_tg + b_ck + b_sg + b_st + b_se + b_nf + 2 * b_key + 2 * b_val + b_tmp + b_G + b_out +
                         b_gx + b_ls + b_xf + 2 * b_cnt;
    char* buf = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&buf, total));
    char* q = buf;
    auto take = [&](size_t n) { char* r = q; q += n; return (void*)r; };
    double* d_xv = (double*)take(b_xv);
    double* d_tg = (double*)take(b_tg);
    AdjParams a;
    a.ckpt = (double*)take(b_ck);
    a.seg = take(b_sg);
    a.state = take(b_st);
    a.state_exp = (int*)take(b_se);
    unsigned long long* d_nf = (unsigned long long*)take(b_nf);
    unsigned* keys0 = (unsigned*)take(b_key);
    unsigned* keys1 = (unsigned*)take(b_key);
    AdjRec<R>* vals0 = (AdjRec<R>*)take(b_val);
    AdjRec<R>* vals1 = (AdjRec<R>*)take(b_val);
    void* d_tmp = take(b_tmp);
    double* G = (double*)take(b_G);
    double* d_out = (double*)take(b_out);
    a.grad_xv0 = (double*)take(b_gx);
    a.loss = (double*)take(b_ls);
    a.xfinal = (double*)take(b_xf);
    unsigned* d_cnt = (unsigned*)take(b_cnt);      // segment starts [bin][block]
    unsigned* d_off = (unsigned*)take(b_cnt);      // segment counts [bin][block]
    a.targets = d_tg;
    a.rec_key = keys0;
    a.rec_val = vals0;
    a.K = K;
    a.nseg = nseg;
    a.ncell = ncell;
    for (int k = 0; k < 3; ++k) a.nc[k] = nc[k];
    p.xv0 = d_xv;
    hipStream_t st = ctx->stream;
    const dim3 block(256), grid((unsigned)((N + 255) / 256));
    const long long ng = 24 * (long long)ncell;
    float ms_f = 0.f, ms_r = 0.f;
    hipError_t e = hipMemcpyAsync(d_xv, xv0, sizeof(double) * 6 * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tg, targets, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, G, ng);
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, a.loss, 1LL);
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, (double*)d_nf, 1LL);
        e = hipEventRecord(ctx->ev0, st);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL((k_adj_forward<F32, PER, T, CELLS>), grid, block, 0, st, p, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms_f, ctx->ev0, ctx->ev1);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, st);
    for (int seg = nseg - 1; seg >= 0 && e == hipSuccess; --seg) {
        hipLaunchKernelGGL((k_adj_reverse<F32, PER, T, CELLS>), grid, block, 0, st, p, a, seg);
        e = hipGetLastError();
        if (e == hipSuccess && local) {
            e = adj_localsort_records<R>(keys0, vals0, M, nbins, keys1, vals1, d_cnt, d_off, st);
            if (e == hipSuccess) e = adj_bins_seg<R>(keys1, vals1, d_cnt, d_off, (int)nblk, nbins, ncell, G, st);
        } else if (e == hipSuccess) {
            if (partition)
                e = adj_sort_records<R>(d_tmp, temp_bytes, keys0, keys1, vals0, vals1, M, kShift, cellbits, st);
            if (e == hipSuccess) {
                hipLaunchKernelGGL((k_adj_bins<R>), dim3(nbins), dim3(AdjBin::kThreads), 0, st,
                                   partition ? keys1 : keys0, partition ? vals1 : vals0, M, ncell, G);
                e = hipGetLastError();
            }
        }
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_adj_nodes<double>, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, st, G, nx,
                           ny, nz, nc[0], nc[1], nc[2], PER ? 1 : 0, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, st);
    // the outputs are scanned for entries that are not finite (the loss, the
    // node gradient, the initial-state gradient)
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_count_nonfinite, dim3(1), dim3(64), 0, st, a.loss, 1LL, d_nf);
        hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)std::min<size_t>((3 * nnodes + 255) / 256, 4096)),
                           dim3(256), 0, st, d_out, (long long)(3 * nnodes), d_nf);
        hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)std::min<long long>((6 * N + 255) / 256, 4096)),
                           dim3(256), 0, st, a.grad_xv0, (long long)(6 * N), d_nf);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms_r, ctx->ev0, ctx->ev1);
    unsigned long long nonfinite = 0;
    if (e == hipSuccess) e = hipMemcpy(dL_dB, d_out, sizeof(double) * 3 * nnodes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(dL_dxv0, a.grad_xv0, sizeof(double) * 6 * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(loss, a.loss, sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&nonfinite, d_nf, sizeof nonfinite, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(ctx, TPAMD_ERR_DEVICE, "adjoint failed: %s", hipGetErrorString(e));
    if (stats) {
        stats->kernel_ms = ms_f;
        stats->aux_ms = ms_r;
        stats->particle_steps = nt * N;
        stats->nt = nt;
        stats->nout = 1;
        stats->max_len = 1;
        stats->total_ms = ms_f + ms_r;
    }
    if (nonfinite)
        return fail(ctx, TPAMD_ERR_NONFINITE,
                    "%llu entries of the loss and its gradients are not finite (the forward "
                    "trajectory left the field, or a gradient exceeds the double range)",
                    nonfinite);
    return TPAMD_OK;
}

}  // namespace

namespace tpamd {

int adjoint_launch(tpamd_ctx* ctx, const KParams& p, const tpamd_field* B, const double* xv0,
                   int64_t N, const double* targets, int K, int64_t nt, double* loss,
                   double* dL_dB, double* dL_dxv0, tpamd_stats* stats, bool f32, bool bf32,
                   bool per, bool cells)
{
#define TPAMD_ADJ(F32_, PER_, T_, CELLS_) \
    return adjoint_run<F32_, PER_, T_, CELLS_>(ctx, p, B, xv0, N, targets, K, nt, loss, dL_dB, dL_dxv0, stats)
    if (f32) {
        if (cells) {
            if (per) TPAMD_ADJ(true, true, float, true);
            TPAMD_ADJ(true, false, float, true);
        }
        if (per) TPAMD_ADJ(true, true, float, false);
        TPAMD_ADJ(true, false, float, false);
    }
    if (bf32) {
        if (per) TPAMD_ADJ(false, true, float, false);
        TPAMD_ADJ(false, false, float, false);
    }
    if (cells) {
        if (per) TPAMD_ADJ(false, true, double, true);
        TPAMD_ADJ(false, false, double, true);
    }
    if (per) TPAMD_ADJ(false, true, double, false);
    TPAMD_ADJ(false, false, double, false);
#undef TPAMD_ADJ
}

}  // namespace tpamd

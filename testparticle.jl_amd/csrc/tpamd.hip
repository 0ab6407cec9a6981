// tpamd.hip -- host side and C ABI of the MI355X ensemble Boris pusher: fields,
// plans, solves, batches and the adjoint's entry; the kernels live in the
// family units (tpamd_gen64/32[m], tpamd_spec64/32, tpamd_adj).
#include "tpamd_pusher.h"
#include "tpamd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

namespace {

__global__ void k_batch_reset(unsigned* counter, unsigned long long* steps)
{
    if (blockIdx.x == 0 && threadIdx.x < kQueues) counter[threadIdx.x * kQueueStride] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) *steps = 0ull;
}

__global__ void __launch_bounds__(256) k_field_eval(const KParams p, const double* __restrict__ x,
                                                    long long n, double t, int which, int prec,
                                                    double* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (prec == 0) {
        S64M s;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.pos.x[c] = x[3 * i + c];
        double E[3], B[3];
        s.fields(p, t, E, B);
        const double* o = which ? E : B;
        for (int c = 0; c < 3; ++c) out[3 * i + c] = o[c];
    } else {
        S32CM s;
        const double tmp[6] = {x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0, 0.0, 0.0};
        s.load_from(p, tmp);
        float E[3], B[3];
        s.fields(p, s.pos, t, E, B);
        const float* o = which ? E : B;
        for (int c = 0; c < 3; ++c) out[3 * i + c] = (double)o[c];
    }
}

// fields and work rates at given states: the columns record_extra writes,
// without the solve (p.save_fields / p.save_work / p.rec_w select them)
__global__ void __launch_bounds__(256) k_record_eval(const KParams p, const double* __restrict__ xv,
                                                     const double* __restrict__ t, long long n,
                                                     double* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[3], v[3], r[16];
    for (int c = 0; c < 3; ++c) {
        x[c] = xv[6 * i + c];
        v[c] = xv[6 * i + 3 + c];
    }
    record_extra<kFkMesh>(p, x, v, t[i], r);
    const int w = p.rec_w - 6;
    for (int c = 0; c < w; ++c) out[(size_t)w * i + c] = r[6 + c];
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
namespace {

thread_local std::string g_init_err;

}  // namespace

int tpamd::fail(tpamd_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else g_init_err = buf;
    return code;
}

namespace {


int fill_poly(int n, Poly& P)
{
    // integer coefficients of the rational multistep form: D = (1+u)^n,
    // P1 + i tau P2 = (1 + i tau)^(2n), P3 = (D - P1)/u, P6 = (2n D - P2)/u
    if (n < 1 || n > kRationalNMax) return 0;
    auto binom = [](int a, int b) {
        double r = 1.0;
        for (int j = 1; j <= b; ++j) r = r * (double)(a - b + j) / (double)j;
        return std::floor(r + 0.5);
    };
    std::memset(&P, 0, sizeof P);
    for (int j = 0; j <= n; ++j) P.D[j] = binom(n, j);
    for (int k = 0; k <= 2 * n; ++k) {
        const double b = binom(2 * n, k);
        if (k % 2 == 0) P.P1[k / 2] += ((k / 2) % 2 == 0 ? 1.0 : -1.0) * b;
        else P.P2[(k - 1) / 2] += (((k - 1) / 2) % 2 == 0 ? 1.0 : -1.0) * b;
    }
    for (int j = 1; j <= n; ++j) P.P3[j - 1] = P.D[j] - P.P1[j];
    for (int j = 1; j <= n; ++j) P.P6[j - 1] = 2.0 * n * P.D[j] - (j <= n - 1 ? P.P2[j] : 0.0);
    P.deg = n;
    return 1;
}

int check_desc(tpamd_ctx* ctx, const tpamd_solve_desc* d, int64_t* nt, int64_t* nout)
{
    if (!d) return fail(ctx, TPAMD_ERR_ARGUMENT, "null descriptor");
    if (d->savestepinterval < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "savestepinterval must be >= 1");
    if (d->alg != TPAMD_ALG_BORIS && d->alg != TPAMD_ALG_MULTISTEP)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "unknown algorithm %d", d->alg);
    if (d->alg == TPAMD_ALG_MULTISTEP) {
        if (d->order_N != 2 && d->order_N != 4 && d->order_N != 6)
            return fail(ctx, TPAMD_ERR_ORDER, "Multistep Boris order N must be 2, 4, or 6.");
        if (d->nsub < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "multistep n must be >= 1");
        if (d->form == TPAMD_FORM_RATIONAL && d->nsub > kRationalNMax)
            return fail(ctx, TPAMD_ERR_UNSUPPORTED, "rational multistep form supports n <= %d",
                        kRationalNMax);
        // the reference form's sincos(n atan(tau)) reduces n atan(tau) < n pi/2
        // with the medium-size FDLIBM reduction, which holds below 2^20 pi/2
        if (d->form != TPAMD_FORM_RATIONAL && d->nsub > (1 << 18))
            return fail(ctx, TPAMD_ERR_UNSUPPORTED, "reference multistep form supports n <= 262144");
    }
    if (d->precision != TPAMD_FP64 && d->precision != TPAMD_FP32C)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "unknown precision %d", d->precision);
    if (d->adaptive && !(d->safety > 0.0 && std::isfinite(d->safety)))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "adaptive safety factor must be positive and finite");
    if (d->adaptive && !(std::fabs(d->q2m) > 0.0 && std::isfinite(d->q2m)))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "adaptive solve needs a finite nonzero q/m");
    if (d->max_steps < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "max_steps must be >= 0");
    int64_t n = 0, no = 0;
    if (!d->adaptive) {
        if (d->dt_pp) {
            n = d->nsteps_pp;
            if (n < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "nsteps_pp must be >= 0");
        } else {
            // _prepare_checks, src/boris/boris.jl:349-356
            if (std::fabs(d->dt) < 10.0 * 2.220446049250313e-16)
                return fail(ctx, TPAMD_ERR_DT_TOO_SMALL,
                            "time step dt is too small, violating min_dt = 10 * eps(typeof(dt))");
            const double ttotal = d->t1 - d->t0;
            n = std::llabs((long long)std::nearbyint(ttotal / d->dt));
        }
        if (n > d->maxiters)
            return fail(ctx, TPAMD_ERR_MAXITERS, "number of iterations nt (%lld) exceeds maxiters (%lld)",
                        (long long)n, (long long)d->maxiters);
        // _calculate_nout, src/boris/boris.jl:383-403
        const int64_t K = d->savestepinterval;
        if (d->save_start) no += 1;
        if (d->save_everystep) {
            const bool last_is_step = (n > 0) && (n % K == 0);
            no += n / K;
            if (!d->save_end && last_is_step) no -= 1;
            if (d->save_end && !last_is_step) no += 1;
        } else if (d->save_end) {
            no += 1;
        }
    }
    if (nt) *nt = n;
    if (nout) *nout = no;
    return TPAMD_OK;
}

bool is_gridded(const DevField& f) { return f.kind == kFieldGrid || f.kind == kFieldTimeGrid; }

// the field-kind class of a solve: the generic kernels compiled for it
int solve_class(const KParams& p)
{
    const int b = field_class(p.B.kind);
    return p.e_is_uxb ? b : std::max(b, field_class(p.E.kind));
}

bool same_geometry(const DevField& a, const DevField& b)
{
    for (int d = 0; d < 3; ++d)
        if (a.n[d] != b.n[d] || a.origin[d] != b.origin[d] || a.inv_h[d] != b.inv_h[d]) return false;
    return a.extrap == b.extrap;
}

int setup_fields(tpamd_ctx* ctx, const tpamd_field* E, const tpamd_field* B, int precision,
                 KParams& p)
{
    if (!E || !B) return fail(ctx, TPAMD_ERR_ARGUMENT, "null field");
    if (B->d.kind == kFieldUxB) return fail(ctx, TPAMD_ERR_ARGUMENT, "B cannot be a UxB field");
    p.B = B->d;
    p.E = E->d;
    p.e_is_uxb = 0;
    if (E->d.kind == kFieldUxB) {
        if (E->src != B)
            return fail(ctx, TPAMD_ERR_UNSUPPORTED, "UxB electric field must derive from the solve's B");
        p.e_is_uxb = 1;
    }
    // an affine field with a non-zero ft term and a snapshot sequence depend
    // on t; a UxB E follows B
    auto is_static = [](const DevField& f) {
        if (f.kind == kFieldTimeGrid) return f.nsnap == 1;
        return f.kind != kFieldAffine || (f.ft[0] == 0.0 && f.ft[1] == 0.0 && f.ft[2] == 0.0);
    };
    p.static_fields = is_static(B->d) && (p.e_is_uxb || is_static(E->d)) ? 1 : 0;
    auto is_uniform = [](const DevField& f) { return f.kind == kFieldZero || f.kind == kFieldUniform; };
    p.uniform_fields = is_uniform(B->d) && (p.e_is_uxb || is_uniform(E->d)) ? 1 : 0;
    for (int d = 0; d < 3; ++d) { p.frame_origin[d] = 0.0; p.frame_period[d] = 0.0; }
    p.frame_periodic = 0;
    p.span = 64;              // adaptive single solves set their own
    if (precision == TPAMD_FP32C) {
        const DevField* g = nullptr;
        if (is_gridded(B->d)) g = &B->d;
        if (!p.e_is_uxb && is_gridded(E->d)) {
            if (g && !same_geometry(*g, E->d))
                return fail(ctx, TPAMD_ERR_UNSUPPORTED,
                            "FP32C needs gridded E and B on the same grid");
            if (!g) g = &E->d;
        }
        if (g) {
            for (int d = 0; d < 3; ++d) p.frame_origin[d] = g->origin[d];
            if (g->extrap == kExtrapPeriodic) {
                for (int d = 0; d < 3; ++d) p.frame_period[d] = g->period[d];
                p.frame_periodic = 1;
            }
        }
    }
    return TPAMD_OK;
}

// ---------------------------------------------------------------------------
// Policy selection: the specialised gridded-B kernels where they apply, the
// generic ones otherwise.  The specialisations cover: B a grid whose bytes
// fit 32-bit offsets, E zero, uniform or UxB, the Boris update (fixed or
// adaptive) or the rational multistep update with n <= 2 (fixed step), FP32C
// on float storage.
// ---------------------------------------------------------------------------
// -1 when no specialisation applies, else the subcycle class NS (0 = Boris,
// n = rational multistep, kRefBase + n = reference-form multistep)
int spec_class(const tpamd_solve_desc* d, const KParams& p, int* emode)
{
    if (d->flags & TPAMD_FLAG_GENERIC_KERNELS) return -1;
    if (TPAMD_RECORD_WIDTH(d) > 6) return -1;       // E, B, work columns: generic REC kernels
    if (p.B.kind != kFieldGrid) return -1;
    if (p.e_is_uxb) *emode = kEUxB;
    else if (p.E.kind == kFieldZero || p.E.kind == kFieldUniform) *emode = kEUniform;
    else return -1;
    int ns;
    if (d->alg == TPAMD_ALG_BORIS) ns = 0;
    else if (d->nsub == 1 || d->nsub == 2) ns = d->form == TPAMD_FORM_RATIONAL ? d->nsub : kRefBase + d->nsub;
    else return -1;
    if (d->adaptive && ns != 0) return -1;
    const double es = p.B.dtype == 1 ? 4.0 : 8.0;
    if (3.0 * es * (double)p.B.n[0] * (double)p.B.n[1] * (double)p.B.n[2] >= 4294967296.0)
        return -1;
    if (d->precision == TPAMD_FP32C && p.B.dtype != 1) return -1;
    return ns;
}

void select_policy(const tpamd_solve_desc* d, const KParams& p, bool allow_spec, int fk,
                   const LaunchReq& r)
{
    int emode = 0;
    const int ns = allow_spec ? spec_class(d, p, &emode) : -1;
    if (ns >= 0) {
        const bool per = p.B.extrap == kExtrapPeriodic;
        const bool cells = p.B.cells != nullptr;
        if (ns > kRefBase) {
            if (d->precision == TPAMD_FP32C) launch_spec32r(r, emode, ns, per, cells);
            else launch_spec64r(r, emode, ns, per, p.B.dtype == 1, cells);
        } else if (d->precision == TPAMD_FP32C) {
            launch_spec32(r, emode, ns, per, cells);
        } else {
            launch_spec64(r, emode, ns, per, p.B.dtype == 1, cells);
        }
        return;
    }
    // analytic fields have kernels of their own for uniform fields and for
    // the dipole (fixed-step Boris), unless the generic ones are asked for
    const bool special = !(d->flags & TPAMD_FLAG_GENERIC_KERNELS);
    if (d->precision == TPAMD_FP64) {
        if (fk == kFkMesh) launch_generic64_mesh(r, p.alg);
        else if (fk == kFkGrid) launch_generic64(r, p.alg);
        else launch_generic64_analytic(r, p.alg, special);
    } else {
        if (fk == kFkMesh) launch_generic32_mesh(r, p.alg);
        else if (fk == kFkGrid) launch_generic32(r, p.alg);
        else launch_generic32_analytic(r, p.alg, special);
    }
}

// Validates the descriptor and fills the kernel parameter block (everything
// but the launch).  Shared by single and batched solves.
int fill_params(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
                const tpamd_field* B, const double* d_xv0, int64_t N, double* d_out_u,
                double* d_out_t, int32_t* d_len, int32_t* d_ret, KParams& p, int64_t& nt,
                int64_t& nout)
{
    nt = 0;
    nout = 0;
    int rc = check_desc(ctx, desc, &nt, &nout);
    if (rc) return rc;
    if (N < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "N must be >= 0");
    if (!desc->adaptive && desc->nout_cap < nout)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "nout_cap (%lld) < records per particle (%lld)",
                    (long long)desc->nout_cap, (long long)nout);
    if (desc->nout_cap < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "nout_cap must be >= 0");
    std::memset(&p, 0, sizeof p);
    rc = setup_fields(ctx, E, B, desc->precision, p);
    if (rc) return rc;
    if (desc->alg == TPAMD_ALG_MULTISTEP && desc->form == TPAMD_FORM_RATIONAL) {
        fill_poly(desc->nsub, p.poly);
        for (int j = 0; j <= kRationalNMax; ++j) {
            p.f32.D[j] = (float)p.poly.D[j];
            p.f32.P1[j] = (float)p.poly.P1[j];
        }
        for (int j = 0; j < kRationalNMax; ++j) {
            p.f32.P2[j] = (float)p.poly.P2[j];
            p.f32.P3[j] = (float)p.poly.P3[j];
            p.f32.P6[j] = (float)p.poly.P6[j];
        }
    }
    for (int c = 0; c < 3; ++c) {
        p.f32.u[c] = (float)E->d.v[c];     // UxB drift (kind UxB) or uniform E value
        p.f32.e0[c] = (float)E->d.v[c];
    }
    p.alg = desc->alg;
    p.order_N = desc->order_N;
    p.nsub = desc->alg == TPAMD_ALG_MULTISTEP ? desc->nsub : 1;
    p.q2m_2n = desc->q2m / (2.0 * p.nsub);
    p.form = desc->form;
    p.save_start = desc->save_start != 0;
    p.save_end = desc->save_end != 0;
    p.save_every = desc->save_everystep != 0;
    p.bnd_flags = desc->bnd_flags;
    p.q2m = desc->q2m;
    p.t0 = desc->t0;
    p.t1 = desc->t1;
    p.dt = desc->dt;
    {
        // C = 2 pi safety sign(t1 - t0) / |q2m|, src/boris/adaptive_boris.jl:164
        const double tt = desc->t1 - desc->t0;
        const double sg = tt > 0 ? 1.0 : (tt < 0 ? -1.0 : 0.0);
        p.C = ((2.0 * M_PI) * desc->safety * sg) / std::fabs(desc->q2m);
    }
    p.nt = nt;
    p.nout = nout;
    p.cap = desc->nout_cap;
    p.K = desc->savestepinterval;
    p.save_fields = desc->save_fields != 0;
    p.save_work = desc->save_work != 0;
    p.rec_w = TPAMD_RECORD_WIDTH(desc);
    p.mass = desc->mass;
    p.max_steps = desc->max_steps;
    p.N = N;
    p.r_in = desc->r_in;
    p.r_out = desc->r_out;
    // sqrt is correctly rounded and monotone: |x|^2 < (r (1 - 1e-12))^2 gives
    // sqrt(|x|^2) < r, and |x|^2 > (r (1 + 1e-12))^2 gives sqrt(|x|^2) > r
    p.rin2_lo = (p.r_in * (1.0 - 1e-12)) * (p.r_in * (1.0 - 1e-12));
    p.rin2_hi = (p.r_in * (1.0 + 1e-12)) * (p.r_in * (1.0 + 1e-12));
    p.rout2_lo = (p.r_out * (1.0 - 1e-12)) * (p.r_out * (1.0 - 1e-12));
    p.rout2_hi = (p.r_out * (1.0 + 1e-12)) * (p.r_out * (1.0 + 1e-12));
    // FP32C pre-test: a float |x|^2 from the float position is within 1e-6
    // of the exact one, so a margin of 1e-5 settles the sphere tests away
    // from the radii (inf / NaN thresholds leave every step to the exact test)
    p.f32.rin2_skip = std::nextafter((float)(p.rin2_hi * (1.0 + 1e-5)), std::numeric_limits<float>::infinity());
    p.f32.rout2_skip = std::nextafter((float)(p.rout2_lo * (1.0 - 1e-5)), -std::numeric_limits<float>::infinity());
    for (int d = 0; d < 3; ++d) { p.box_lo[d] = desc->box_lo[d]; p.box_hi[d] = desc->box_hi[d]; }
    if (desc->bnd_flags & TPAMD_BND_GRID) {
        // open boundary at the faces of B's grid: the node span becomes the box
        const bool mesh = B->d.kind == kFieldMesh;
        if (!is_gridded(B->d) && !(mesh && B->mesh_host.coord == kMeshCartesian))
            return fail(ctx, TPAMD_ERR_ARGUMENT, "TPAMD_BND_GRID needs a Cartesian gridded B field");
        for (int d = 0; d < 3; ++d) {
            const MeshAxis& ma = B->mesh_host.ax[d];
            const double inf = std::numeric_limits<double>::infinity();
            const double lo = mesh ? (ma.n == 1 ? -inf : ma.lo) : B->d.origin[d];
            const double hi = mesh ? (ma.n == 1 ? inf : ma.hi)
                                   : B->d.origin[d] + (double)(B->d.n[d] - 1) / B->d.inv_h[d];
            if (!(desc->bnd_flags & TPAMD_BND_BOX)) { p.box_lo[d] = lo; p.box_hi[d] = hi; }
            else { p.box_lo[d] = std::fmax(p.box_lo[d], lo); p.box_hi[d] = std::fmin(p.box_hi[d], hi); }
        }
        p.bnd_flags = (p.bnd_flags & ~TPAMD_BND_GRID) | TPAMD_BND_BOX;
    }
    p.xv0 = d_xv0;
    p.dt_pp = desc->dt_pp;
    p.out_u = d_out_u;
    p.out_t = d_out_t;
    p.out_len = d_len;
    p.retcode = d_ret;
    p.steps = ctx->d_steps;
    return TPAMD_OK;
}

int solve_impl(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
               const tpamd_field* B, const double* d_xv0, int64_t N, double* d_out_u,
               double* d_out_t, int32_t* d_len, int32_t* d_ret, hipStream_t st,
               tpamd_stats* stats)
{
    KParams p;
    int64_t nt = 0, nout = 0;
    int rc = fill_params(ctx, desc, E, B, d_xv0, N, d_out_u, d_out_t, d_len, d_ret, p, nt, nout);
    if (rc) return rc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_steps, 0, sizeof(unsigned long long), st));
    const int bs = desc->block_size > 0 ? desc->block_size : 256;
    if (bs % 64 != 0 || bs > 256)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "block_size must be a multiple of 64 up to 256");
    // adaptive: a wave takes p.span particles.  More particles per lane let
    // lanes refill when lives are uneven but leave fewer waves; by default a
    // lane takes one particle until the grid holds 4,096 waves (16 per CU),
    // and at most 8
    const int ppl_req = desc->particles_per_lane;
    if (ppl_req < 0 || ppl_req > 1024)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "particles_per_lane must be in [0, 1024]");
    if (ppl_req > 0 && !desc->adaptive)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "particles_per_lane applies to adaptive solves only");
    const long long ppl = ppl_req > 0 ? ppl_req : std::max(1LL, std::min(8LL, (long long)N / (64LL * 4096)));
    p.span = (int)(64 * ppl);
    float ms = 0.f;
    if (N > 0) {
        const long long threads = desc->adaptive ? (N + p.span - 1) / p.span * 64 : N;
        const dim3 block(bs), grid((unsigned)((threads + bs - 1) / bs));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
        LaunchReq r{};
        r.batch = false;
        r.adaptive = desc->adaptive;
        r.g = grid;
        r.b = block;
        r.st = st;
        r.p = &p;
        select_policy(desc, p, true, solve_class(p), r);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    }
    unsigned long long steps = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&steps, ctx->d_steps, sizeof steps, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (N > 0) HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->particle_steps = (int64_t)steps;
        stats->nt = nt;
        stats->nout = nout;
    }
    return TPAMD_OK;
}

int64_t max_len_host(const int32_t* len, int64_t N)
{
    int64_t m = 0;
    for (int64_t i = 0; i < N; ++i) m = len[i] > m ? len[i] : m;
    return m;
}

}  // namespace

struct tpamd_batch {
    tpamd_ctx* ctx;
    tpamd_solve_desc desc;
    int n_inst;
    int64_t N;
    std::vector<KParams> table;
    std::vector<int64_t> offsets;
    KParams* d_table;
    Chunk* d_chunks;
    int nchunks;
    size_t chunks_cap;
    unsigned* d_counter;
    unsigned long long* d_steps;
    bool allow_spec;
    int fk;                  // field-kind class covering every instance
    int block_size;
    int ppt;
    int n_cu;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool have_graph;
};

namespace {

// the work chunks of a batch: 64 ppt consecutive particles of one instance,
// every stride-th of them
std::vector<Chunk> make_chunks(const std::vector<int64_t>& offsets, int n_inst, int ppt, int stride)
{
    std::vector<Chunk> ch;
    const int64_t csz = 64 * (int64_t)ppt;
    int64_t k = 0;
    for (int i = 0; i < n_inst; ++i)
        for (int64_t s = offsets[i]; s < offsets[i + 1]; s += csz, ++k)
            if (k % stride == 0) {
                Chunk c;
                c.inst = i;
                c.pad = 0;
                c.begin = s;
                c.end = std::min(s + csz, offsets[i + 1]);
                ch.push_back(c);
            }
    return ch;
}

int batch_chunks(tpamd_batch* b, int ppt, int stride)
{
    const std::vector<Chunk> ch = make_chunks(b->offsets, b->n_inst, ppt, stride);
    tpamd_ctx* ctx = b->ctx;
    if (ch.size() > b->chunks_cap) {
        if (b->d_chunks) HIP_TRY(ctx, hipFree(b->d_chunks));
        b->d_chunks = nullptr;
        b->chunks_cap = 0;
        HIP_TRY(ctx, hipMalloc((void**)&b->d_chunks, sizeof(Chunk) * std::max<size_t>(ch.size(), 1)));
        b->chunks_cap = ch.size();
    }
    if (!ch.empty())
        HIP_TRY(ctx, hipMemcpy(b->d_chunks, ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice));
    b->nchunks = (int)ch.size();
    return TPAMD_OK;
}

// enqueue reset + persistent launch on `st`
int batch_enqueue(tpamd_batch* b, hipStream_t st)
{
    tpamd_ctx* ctx = b->ctx;
    const dim3 block(b->block_size);
    int occ = 0;
    const bool rec = TPAMD_RECORD_WIDTH(&b->desc) > 6;
    LaunchReq q{};
    q.batch = true;
    q.adaptive = b->desc.adaptive;
    q.g = dim3(1);
    q.b = block;
    q.st = st;
    q.table = b->d_table;
    q.chunks = b->d_chunks;
    q.nchunks = b->nchunks;
    q.counter = b->d_counter;
    q.steps = b->d_steps;
    q.occ = &occ;
    q.rec = rec;
    select_policy(&b->desc, b->table[0], b->allow_spec, b->fk, q);
    occ = std::max(occ, 1);
    const long long waves_needed = b->nchunks;
    // The occupancy query counts 256 VGPRs per lane, half of what gfx950
    // gives a SIMD (e.g. 3 instead of 6 waves per SIMD at 80 VGPRs), so the
    // persistent grid is doubled: blocks that find no free slot start when
    // others leave, see an empty queue and exit.
    long long blocks = (long long)b->n_cu * occ * 2;
    const long long waves_per_block = b->block_size / 64;
    blocks = std::min(blocks, (waves_needed + waves_per_block - 1) / waves_per_block);
    blocks = std::max(blocks, 1LL);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, b->d_counter, b->d_steps);
    LaunchReq l = q;
    l.g = dim3((unsigned)blocks);
    l.occ = nullptr;
    select_policy(&b->desc, b->table[0], b->allow_spec, b->fk, l);
    HIP_TRY(ctx, hipGetLastError());
    return TPAMD_OK;
}

int batch_capture(tpamd_batch* b)
{
    tpamd_ctx* ctx = b->ctx;
    if (b->have_graph) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->have_graph = false;
    }
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue(b, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(ctx, TPAMD_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->have_graph = true;
    return TPAMD_OK;
}

int batch_time_once(tpamd_batch* b, hipStream_t st, float* ms)
{
    tpamd_ctx* ctx = b->ctx;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    int rc = batch_enqueue(b, st);
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return TPAMD_OK;
}


// ---------------------------------------------------------------------------
// Scratch of the allocation-free batched solve (tpamd_solve_batch_noalloc):
// the kQueues claim counters and the step counter, the per-instance
// parameter table and the chunk list, in one file-scope device array of
// fixed size instead of memory the library allocates.  Each GPU holds its
// own copy; tpamd_init looks up its address on the context's device.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxInst = 256;
constexpr long long kNoAllocMaxChunks = 131072;
constexpr size_t kNoAllocCounterBytes = 4096;
constexpr size_t kNoAllocTableBytes = (sizeof(KParams) * kNoAllocMaxInst + 255) & ~(size_t)255;
constexpr size_t kNoAllocBytes =
    kNoAllocCounterBytes + kNoAllocTableBytes + sizeof(Chunk) * (size_t)kNoAllocMaxChunks;
static_assert(sizeof(unsigned) * kQueues * kQueueStride + sizeof(unsigned long long) <= kNoAllocCounterBytes,
              "claim counters and the step counter fit their region");

}  // namespace

// 3.4 MiB: counters, a 256-instance table, 131,072 chunks
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

extern "C" {

int32_t tpamd_abi_version(void) { return TPAMD_ABI_VERSION; }

int32_t tpamd_init(int32_t device, tpamd_ctx** out)
{
    if (!out) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, TPAMD_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n)
        return fail(nullptr, TPAMD_ERR_ARGUMENT, "device %d out of range (have %d)", device, n);
    tpamd_ctx* c = new tpamd_ctx();
    c->device = device;
    c->d_buf = nullptr;
    c->d_buf_bytes = 0;
    c->scratch = nullptr;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipMalloc((void**)&c->d_steps, sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return fail(nullptr, TPAMD_ERR_DEVICE, "HIP context set-up failed on device %d", device);
    }
    // this device's copy of the allocation-free entry point's scratch array
    // (with the device current); the name is a host-side stand-in, only the
    // looked-up address may reach a kernel or a copy
    void* sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess) c->scratch = (unsigned char*)sp;
    else (void)hipGetLastError();
    *out = c;
    return TPAMD_OK;
}

int32_t tpamd_finalize(tpamd_ctx* ctx)
{
    if (!ctx) return TPAMD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_buf) (void)hipFree(ctx->d_buf);
    (void)hipFree(ctx->d_steps);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TPAMD_OK;
}

const char* tpamd_last_error(const tpamd_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_init_err.c_str();
}

static int32_t new_field(tpamd_ctx* ctx, tpamd_field** out, int kind)
{
    if (!ctx || !out) return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    tpamd_field* f = new tpamd_field();
    std::memset(&f->d, 0, sizeof f->d);
    f->d.kind = kind;
    f->dev_data = nullptr;
    f->cells_data = nullptr;
    f->src = nullptr;
    f->bytes = 0;
    *out = f;
    return TPAMD_OK;
}

int32_t tpamd_field_zero(tpamd_ctx* ctx, tpamd_field** out) { return new_field(ctx, out, kFieldZero); }

int32_t tpamd_field_uniform(tpamd_ctx* ctx, const double value[3], tpamd_field** out)
{
    if (!value) return fail(ctx, TPAMD_ERR_ARGUMENT, "null value");
    int rc = new_field(ctx, out, kFieldUniform);
    if (rc) return rc;
    for (int c = 0; c < 3; ++c) (*out)->d.v[c] = value[c];
    return TPAMD_OK;
}

int32_t tpamd_field_affine(tpamd_ctx* ctx, const double f0[3], const double jac[9],
                           const double ft[3], tpamd_field** out)
{
    if (!f0) return fail(ctx, TPAMD_ERR_ARGUMENT, "null f0");
    int rc = new_field(ctx, out, kFieldAffine);
    if (rc) return rc;
    for (int c = 0; c < 3; ++c) (*out)->d.v[c] = f0[c];
    for (int c = 0; c < 9; ++c) (*out)->d.jac[c] = jac ? jac[c] : 0.0;
    for (int c = 0; c < 3; ++c) (*out)->d.ft[c] = ft ? ft[c] : 0.0;
    return TPAMD_OK;
}

int32_t tpamd_field_dipole(tpamd_ctx* ctx, double b0, double radius, tpamd_field** out)
{
    if (!(radius > 0)) return fail(ctx, TPAMD_ERR_ARGUMENT, "dipole radius must be > 0");
    int rc = new_field(ctx, out, kFieldDipole);
    if (rc) return rc;
    (*out)->d.b0 = b0;
    (*out)->d.radius = radius;
    return TPAMD_OK;
}

namespace {

int check_grid_args(tpamd_ctx* ctx, int32_t dtype, int64_t nx, int64_t ny, int64_t nz,
                    const double origin[3], const double spacing[3], int32_t extrap,
                    int32_t storage_dtype)
{
    if (nx < 2 || ny < 2 || nz < 2)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "grid needs at least 2 nodes per dimension");
    if (nx > (1 << 30) || ny > (1 << 30) || nz > (1 << 30))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "grid dimension too large");
    if ((dtype != TPAMD_F64 && dtype != TPAMD_F32) ||
        (storage_dtype != TPAMD_F64 && storage_dtype != TPAMD_F32))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "dtype must be TPAMD_F64 or TPAMD_F32");
    if (extrap < 0 || extrap > 2) return fail(ctx, TPAMD_ERR_ARGUMENT, "unknown extrapolation %d", extrap);
    for (int d = 0; d < 3; ++d)
        if (!(spacing[d] > 0) || !std::isfinite(origin[d]))
            return fail(ctx, TPAMD_ERR_ARGUMENT, "grid spacing must be > 0 and origin finite");
    return TPAMD_OK;
}

// copies nn host values of `dtype` to dev as `storage_dtype` (converted on the
// host in chunks when the types differ: a one-time set-up cost)
hipError_t upload_nodes(void* dev, const void* data, int32_t dtype, size_t nn, int32_t storage_dtype)
{
    const size_t es = storage_dtype == TPAMD_F64 ? 8 : 4;
    if (dtype == storage_dtype) return hipMemcpy(dev, data, nn * es, hipMemcpyHostToDevice);
    const size_t chunk = 1 << 22;
    char* tmp = (char*)std::malloc(chunk * es);
    hipError_t e = tmp ? hipSuccess : hipErrorOutOfMemory;
    for (size_t off = 0; e == hipSuccess && off < nn; off += chunk) {
        const size_t m = (nn - off) < chunk ? (nn - off) : chunk;
        for (size_t j = 0; j < m; ++j) {
            const double val = dtype == TPAMD_F64 ? ((const double*)data)[off + j]
                                                  : (double)((const float*)data)[off + j];
            if (es == 8) ((double*)tmp)[j] = val;
            else ((float*)tmp)[j] = (float)val;
        }
        e = hipMemcpy((char*)dev + off * es, tmp, m * es, hipMemcpyHostToDevice);
    }
    std::free(tmp);
    return e;
}

void set_grid_geometry(DevField& f, int64_t nx, int64_t ny, int64_t nz, const double origin[3],
                       const double spacing[3], int32_t extrap, int32_t storage_dtype)
{
    f.extrap = extrap;
    f.dtype = storage_dtype == TPAMD_F64 ? 0 : 1;
    f.n[0] = (int)nx; f.n[1] = (int)ny; f.n[2] = (int)nz;
    for (int d = 0; d < 3; ++d) {
        f.origin[d] = origin[d];
        f.inv_h[d] = 1.0 / spacing[d];
        f.inv_hf[d] = (float)f.inv_h[d];
        f.period[d] = (double)f.n[d] * spacing[d];
    }
}

} // namespace

int32_t tpamd_field_grid(tpamd_ctx* ctx, const void* data, int32_t dtype, int64_t nx, int64_t ny,
                         int64_t nz, const double origin[3], const double spacing[3],
                         int32_t extrap, int32_t storage_dtype, tpamd_field** out)
{
    if (!ctx || !out || !data || !origin || !spacing)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    int rc = check_grid_args(ctx, dtype, nx, ny, nz, origin, spacing, extrap, storage_dtype);
    if (rc) return rc;
    const size_t nn = (size_t)nx * (size_t)ny * (size_t)nz * 3;
    const size_t es = storage_dtype == TPAMD_F64 ? 8 : 4;
    void* dev = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc(&dev, nn * es));
    const hipError_t e = upload_nodes(dev, data, dtype, nn, storage_dtype);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return fail(ctx, TPAMD_ERR_DEVICE, "grid upload failed: %s", hipGetErrorString(e));
    }
    rc = new_field(ctx, out, kFieldGrid);
    if (rc) {
        (void)hipFree(dev);
        return rc;
    }
    set_grid_geometry((*out)->d, nx, ny, nz, origin, spacing, extrap, storage_dtype);
    (*out)->d.data = dev;
    (*out)->dev_data = dev;
    (*out)->bytes = nn * es;
    return TPAMD_OK;
}

int32_t tpamd_field_time_grid(tpamd_ctx* ctx, const void* const* snapshots, int32_t dtype,
                              int64_t nx, int64_t ny, int64_t nz, int32_t nsnap,
                              const double* times, const double origin[3],
                              const double spacing[3], int32_t extrap, int32_t storage_dtype,
                              tpamd_field** out)
{
    if (!ctx || !out || !snapshots || !times || !origin || !spacing)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (nsnap < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "a time-dependent grid needs >= 1 snapshot");
    for (int k = 0; k < nsnap; ++k) {
        if (!snapshots[k]) return fail(ctx, TPAMD_ERR_ARGUMENT, "null snapshot %d", k);
        if (!std::isfinite(times[k]))
            return fail(ctx, TPAMD_ERR_ARGUMENT, "snapshot times must be finite");
        if (k > 0 && !(times[k] > times[k - 1]))
            return fail(ctx, TPAMD_ERR_ARGUMENT, "snapshot times must be strictly increasing");
    }
    int rc = check_grid_args(ctx, dtype, nx, ny, nz, origin, spacing, extrap, storage_dtype);
    if (rc) return rc;
    const size_t nn = (size_t)nx * (size_t)ny * (size_t)nz * 3;
    const size_t es = storage_dtype == TPAMD_F64 ? 8 : 4;
    // one allocation: the times, padded to 256 B, then the snapshots, each
    // padded to 256 B so every snapshot starts on its own cache-line boundary
    const size_t head = ((size_t)nsnap * 8 + 255) / 256 * 256;
    const size_t snap = (nn * es + 255) / 256 * 256;
    void* dev = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc(&dev, head + snap * (size_t)nsnap));
    hipError_t e = hipMemcpy(dev, times, (size_t)nsnap * 8, hipMemcpyHostToDevice);
    for (int k = 0; e == hipSuccess && k < nsnap; ++k)
        e = upload_nodes((char*)dev + head + (size_t)k * snap, snapshots[k], dtype, nn, storage_dtype);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return fail(ctx, TPAMD_ERR_DEVICE, "grid upload failed: %s", hipGetErrorString(e));
    }
    rc = new_field(ctx, out, kFieldTimeGrid);
    if (rc) {
        (void)hipFree(dev);
        return rc;
    }
    DevField& f = (*out)->d;
    set_grid_geometry(f, nx, ny, nz, origin, spacing, extrap, storage_dtype);
    f.times = (const double*)dev;
    f.data = (const char*)dev + head;
    f.snap_bytes = (long long)snap;
    f.nsnap = nsnap;
    (*out)->dev_data = dev;
    (*out)->bytes = head + snap * (size_t)nsnap;
    return TPAMD_OK;
}

int32_t tpamd_field_mesh(tpamd_ctx* ctx, const void* data, int32_t dtype, const int64_t n[3],
                         const double* const axes[3], int32_t coord, int32_t order,
                         const int32_t extrap[3], int32_t storage_dtype, tpamd_field** out)
{
    if (!ctx || !out || !data || !n || !axes || !extrap)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (coord != kMeshCartesian && coord != kMeshSpherical)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "unknown mesh coordinates %d", coord);
    if (order != 1 && order != 3)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "interpolation order %d is not provided (1 or 3)", order);
    if ((dtype != TPAMD_F64 && dtype != TPAMD_F32) ||
        (storage_dtype != TPAMD_F64 && storage_dtype != TPAMD_F32))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "dtype must be TPAMD_F64 or TPAMD_F32");
    MeshGrid g;
    std::memset(&g, 0, sizeof g);
    std::vector<std::vector<int>> buckets(3);
    int active = 0;
    for (int d = 0; d < 3; ++d) {
        MeshAxis& a = g.ax[d];
        if (n[d] < 1 || n[d] > (1 << 30)) return fail(ctx, TPAMD_ERR_ARGUMENT, "mesh axis %d: bad size", d);
        if (!axes[d]) return fail(ctx, TPAMD_ERR_ARGUMENT, "mesh axis %d: null coordinates", d);
        if (extrap[d] < 0 || extrap[d] > kExtrapPeriodicIncl)
            return fail(ctx, TPAMD_ERR_ARGUMENT, "mesh axis %d: unknown boundary %d", d, extrap[d]);
        const double* c = axes[d];
        a.n = (int)n[d];
        a.bc = extrap[d];
        a.lo = c[0];
        a.hi = c[a.n - 1];
        for (int i = 0; i < a.n; ++i) {
            if (!std::isfinite(c[i])) return fail(ctx, TPAMD_ERR_ARGUMENT, "mesh axis %d: coordinates must be finite", d);
            if (i > 0 && !(c[i] > c[i - 1]))
                return fail(ctx, TPAMD_ERR_ARGUMENT, "Grid vectors must be sorted (mesh axis %d).", d);
        }
        if (a.n == 1) continue;
        ++active;
        const double h = (a.hi - a.lo) / (double)(a.n - 1);
        bool uni = true;
        double hmin = a.hi - a.lo;
        for (int i = 1; i < a.n; ++i) {
            hmin = std::fmin(hmin, c[i] - c[i - 1]);
            if (std::fabs(c[i] - (a.lo + (double)i * h)) > 1e-9 * (a.hi - a.lo)) uni = false;
        }
        a.uniform = uni ? 1 : 0;
        a.inv_h = uni ? 1.0 / h : 0.0;
        if (a.bc == kExtrapPeriodic) {
            if (!uni)
                return fail(ctx, TPAMD_ERR_UNSUPPORTED, "mesh axis %d: exclusive periodic wrap needs a uniform axis", d);
            a.period = (double)a.n * h;
        } else if (a.bc == kExtrapPeriodicIncl) {
            a.period = a.hi - a.lo;
        }
        if (order == 3 && !uni)
            return fail(ctx, TPAMD_ERR_UNSUPPORTED, "order 3 needs uniform axes (mesh axis %d is not)", d);
        if (!uni) {
            // bucket b covers [lo + b bw, lo + (b+1) bw): its entry is the cell
            // holding the bucket's left end; the device walks from there
            const double span = a.hi - a.lo;
            double want = std::ceil(span / hmin);
            const double cap = 64.0 * (double)(a.n - 1);
            want = want < (double)(a.n - 1) ? (double)(a.n - 1) : want;
            want = want > cap ? cap : want;
            a.nb = (int)want;
            a.inv_bw = (double)a.nb / span;
            std::vector<int>& bk = buckets[d];
            bk.resize(a.nb);
            int i = 0;
            for (int b = 0; b < a.nb; ++b) {
                const double xb = a.lo + (double)b * (span / (double)a.nb);
                while (i < a.n - 2 && c[i + 1] <= xb) ++i;
                bk[b] = i;
            }
        }
    }
    if (active == 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "a mesh needs an axis with at least 2 nodes");
    if (coord == kMeshSpherical) {
        // build_interpolator(StructuredGrid, ...) (interpolation.jl:260-284)
        if (g.ax[0].n < 2 || g.ax[1].n < 2 || g.ax[2].n < 2)
            return fail(ctx, TPAMD_ERR_ARGUMENT, "a spherical mesh needs 2 or more nodes per axis");
        if (!(g.ax[0].lo >= 0.0)) return fail(ctx, TPAMD_ERR_ARGUMENT, "r must be non-negative.");
        const double pi = 3.141592653589793;   // the double nearest pi
        if (!(g.ax[1].lo >= 0.0 && g.ax[1].hi <= pi))
            return fail(ctx, TPAMD_ERR_ARGUMENT, "θ must be within [0, π].");
        if (!(g.ax[2].lo >= 0.0 && g.ax[2].hi <= 2.0 * pi))
            return fail(ctx, TPAMD_ERR_ARGUMENT, "ϕ must be within [0, 2π].");
    }
    g.coord = coord;
    g.order = order;
    g.dtype = storage_dtype == TPAMD_F64 ? 0 : 1;
    g.sy = n[0];
    g.sz = n[0] * n[1];
    const size_t nn = (size_t)n[0] * (size_t)n[1] * (size_t)n[2] * 3;
    const size_t es = storage_dtype == TPAMD_F64 ? 8 : 4;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t off = up(sizeof(MeshGrid));
    size_t off_c[3], off_b[3];
    for (int d = 0; d < 3; ++d) {
        off_c[d] = off;
        off += up((size_t)n[d] * 8);
        off_b[d] = off;
        off += up(buckets[d].size() * sizeof(int));
    }
    const size_t off_data = off;
    const size_t total = off_data + nn * es;
    char* dev = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc((void**)&dev, total));
    hipError_t e = hipSuccess;
    for (int d = 0; d < 3 && e == hipSuccess; ++d) {
        g.ax[d].c = (const double*)(dev + off_c[d]);
        g.ax[d].bucket = buckets[d].empty() ? nullptr : (const int*)(dev + off_b[d]);
        e = hipMemcpy(dev + off_c[d], axes[d], (size_t)n[d] * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess && !buckets[d].empty())
            e = hipMemcpy(dev + off_b[d], buckets[d].data(), buckets[d].size() * sizeof(int),
                          hipMemcpyHostToDevice);
    }
    g.data = dev + off_data;
    if (e == hipSuccess) e = upload_nodes(dev + off_data, data, dtype, nn, storage_dtype);
    if (e == hipSuccess) e = hipMemcpy(dev, &g, sizeof g, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return fail(ctx, TPAMD_ERR_DEVICE, "mesh upload failed: %s", hipGetErrorString(e));
    }
    int rc = new_field(ctx, out, kFieldMesh);
    if (rc) {
        (void)hipFree(dev);
        return rc;
    }
    DevField& f = (*out)->d;
    f.mesh = (const MeshGrid*)dev;
    f.data = dev + off_data;
    f.dtype = g.dtype;
    for (int d = 0; d < 3; ++d) f.n[d] = (int)n[d];
    (*out)->dev_data = dev;
    (*out)->bytes = total;
    (*out)->mesh_host = g;
    return TPAMD_OK;
}

int32_t tpamd_field_uxb(tpamd_ctx* ctx, const double u[3], const tpamd_field* B, tpamd_field** out)
{
    if (!u || !B) return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (B->d.kind == kFieldUxB) return fail(ctx, TPAMD_ERR_ARGUMENT, "B cannot be a UxB field");
    int rc = new_field(ctx, out, kFieldUxB);
    if (rc) return rc;
    for (int c = 0; c < 3; ++c) (*out)->d.v[c] = u[c];
    (*out)->src = B;
    return TPAMD_OK;
}

int32_t tpamd_field_free(tpamd_ctx* ctx, tpamd_field* f)
{
    if (!f) return TPAMD_OK;
    if (ctx) (void)hipSetDevice(ctx->device);
    if (f->dev_data) (void)hipFree(f->dev_data);
    if (f->cells_data) (void)hipFree(f->cells_data);
    delete f;
    return TPAMD_OK;
}

int32_t tpamd_field_grid_cells(tpamd_ctx* ctx, tpamd_field* f, int32_t enable)
{
    if (!ctx || !f) return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (f->d.kind != kFieldGrid)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "the cell layout needs a static gridded field");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!enable) {
        if (f->cells_data) HIP_TRY(ctx, hipFree(f->cells_data));
        f->cells_data = nullptr;
        f->d.cells = nullptr;
        return TPAMD_OK;
    }
    if (f->d.cells) return TPAMD_OK;
    const bool per = f->d.extrap == kExtrapPeriodic;
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = per ? f->d.n[d] : f->d.n[d] - 1;
    const bool f32 = f->d.dtype == 1;
    const unsigned stride = f32 ? 128u : 192u;
    const double total = (double)nc[0] * nc[1] * nc[2];
    if (total * stride >= 4294967296.0)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "cell layout needs fewer than 4 GiB of cells");
    void* cells = nullptr;
    HIP_TRY(ctx, hipMalloc(&cells, (size_t)(total * stride)));
    const long long ncell = (long long)total;
    const unsigned blocks = (unsigned)((ncell + 255) / 256);
    if (f32)
        hipLaunchKernelGGL(k_build_cells<float>, dim3(blocks), dim3(256), 0, ctx->stream,
                           (const float*)f->d.data, f->d.n[0], f->d.n[1], f->d.n[2], (int)per, nc[0],
                           nc[1], nc[2], stride / 4u, (float*)cells);
    else
        hipLaunchKernelGGL(k_build_cells<double>, dim3(blocks), dim3(256), 0, ctx->stream,
                           (const double*)f->d.data, f->d.n[0], f->d.n[1], f->d.n[2], (int)per, nc[0],
                           nc[1], nc[2], stride / 8u, (double*)cells);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(cells);
        return fail(ctx, TPAMD_ERR_DEVICE, "cell layout build failed: %s", hipGetErrorString(e));
    }
    f->cells_data = cells;
    f->d.cells = cells;
    for (int d = 0; d < 3; ++d) f->d.nc[d] = nc[d];
    f->d.cell_stride = stride;
    return TPAMD_OK;
}

int32_t tpamd_field_eval(tpamd_ctx* ctx, const tpamd_field* f, const tpamd_field* B,
                         const double* x, int64_t n, double t, int32_t precision, double* out)
{
    if (!ctx || !f || !x || !out || n < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (precision != TPAMD_FP64 && precision != TPAMD_FP32C)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "unknown precision");
    // a UxB field is evaluated in the E slot next to its B; any other field
    // in the B slot
    const bool uxb = f->d.kind == kFieldUxB;
    if (uxb && (!B || f->src != B))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "UxB field needs the B field it was built from");
    tpamd_field zero;
    std::memset(&zero, 0, sizeof zero);
    KParams p;
    std::memset(&p, 0, sizeof p);
    int rc = uxb ? setup_fields(ctx, f, B, precision, p) : setup_fields(ctx, &zero, f, precision, p);
    if (rc) return rc;
    const int which = uxb ? 1 : 0;
    if (n == 0) return TPAMD_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* dx = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&dx, sizeof(double) * 6 * n));
    double* dout = dx + 3 * n;
    hipError_t e = hipMemcpyAsync(dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_field_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           p, dx, (long long)n, t, which, precision, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dx);
    if (e != hipSuccess) return fail(ctx, TPAMD_ERR_DEVICE, "field eval failed: %s", hipGetErrorString(e));
    return TPAMD_OK;
}

int32_t tpamd_record_eval(tpamd_ctx* ctx, const tpamd_field* E, const tpamd_field* B, double q2m,
                          double mass, const double* xv, const double* t, int64_t n,
                          int32_t want_fields, int32_t want_work, double* out)
{
    if (!ctx || !E || !B || n < 0 || (n > 0 && (!xv || !t || !out)))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (!want_fields && !want_work) return fail(ctx, TPAMD_ERR_ARGUMENT, "nothing asked for");
    KParams p;
    std::memset(&p, 0, sizeof p);
    int rc = setup_fields(ctx, E, B, TPAMD_FP64, p);
    if (rc) return rc;
    p.q2m = q2m;
    p.mass = mass;
    p.save_fields = want_fields != 0;
    p.save_work = want_work != 0;
    p.rec_w = 6 + (p.save_fields ? 6 : 0) + (p.save_work ? 4 : 0);
    if (n == 0) return TPAMD_OK;
    const int w = p.rec_w - 6;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* dxv = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&dxv, sizeof(double) * (7 + (size_t)w) * n));
    double* dt = dxv + 6 * n;
    double* dout = dt + n;
    hipError_t e = hipMemcpyAsync(dxv, xv, sizeof(double) * 6 * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, t, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_record_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p,
                           dxv, dt, (long long)n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(double) * w * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dxv);
    if (e != hipSuccess) return fail(ctx, TPAMD_ERR_DEVICE, "record eval failed: %s", hipGetErrorString(e));
    return TPAMD_OK;
}

int32_t tpamd_plan(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int64_t* nt, int64_t* nout)
{
    return check_desc(ctx, desc, nt, nout);
}

int32_t tpamd_solve_device(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
                           const tpamd_field* B, const double* d_xv0, int64_t N, double* d_out_u,
                           double* d_out_t, int32_t* d_out_len, int32_t* d_retcode, void* stream,
                           tpamd_stats* stats)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (N > 0 && (!d_xv0 || !d_out_len || !d_retcode || (desc && desc->nout_cap > 0 && (!d_out_u || !d_out_t))))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null buffer");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    int rc = solve_impl(ctx, desc, E, B, d_xv0, N, d_out_u, d_out_t, d_out_len, d_retcode, st, stats);
    if (rc) return rc;
    if (stats) {
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->max_len = -1;   // not read back for device buffers
    }
    return TPAMD_OK;
}

int32_t tpamd_solve(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
                    const tpamd_field* B, const double* xv0, int64_t N, double* out_u,
                    double* out_t, int32_t* out_len, int32_t* retcode, tpamd_stats* stats)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (!desc) return fail(ctx, TPAMD_ERR_ARGUMENT, "null descriptor");
    if (N < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "N must be >= 0");
    if (N > 0 && (!xv0 || !out_len || !retcode || (desc->nout_cap > 0 && (!out_u || !out_t))))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null buffer");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t cap = desc->nout_cap;
    const size_t b_xv = sizeof(double) * 6 * (size_t)N;
    const size_t b_u = sizeof(double) * (size_t)TPAMD_RECORD_WIDTH(desc) * (size_t)cap * (size_t)N;
    const size_t b_t = sizeof(double) * (size_t)cap * (size_t)N;
    const size_t b_i = sizeof(int32_t) * (size_t)N;
    const size_t b_dt = desc->dt_pp ? sizeof(double) * (size_t)N : 0;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t need = al(b_xv) + al(b_u) + al(b_t) + 2 * al(b_i) + al(b_dt);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (need > ctx->d_buf_bytes) {
        if (ctx->d_buf) HIP_TRY(ctx, hipFree(ctx->d_buf));
        ctx->d_buf = nullptr;
        ctx->d_buf_bytes = 0;
        HIP_TRY(ctx, hipMalloc(&ctx->d_buf, need));
        ctx->d_buf_bytes = need;
    }
    char* b = (char*)ctx->d_buf;
    double* d_xv = (double*)b; b += al(b_xv);
    double* d_u = (double*)b; b += al(b_u);
    double* d_t = (double*)b; b += al(b_t);
    int32_t* d_len = (int32_t*)b; b += al(b_i);
    int32_t* d_ret = (int32_t*)b; b += al(b_i);
    double* d_dt = (double*)b;
    hipStream_t st = ctx->stream;
    if (N > 0) HIP_TRY(ctx, hipMemcpyAsync(d_xv, xv0, b_xv, hipMemcpyHostToDevice, st));
    tpamd_solve_desc dd = *desc;
    if (desc->dt_pp && N > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(d_dt, desc->dt_pp, b_dt, hipMemcpyHostToDevice, st));
        dd.dt_pp = d_dt;
    }
    int rc = solve_impl(ctx, &dd, E, B, d_xv, N, d_u, d_t, d_len, d_ret, st, stats);
    if (rc) return rc;
    if (N > 0) {
        if (b_u) HIP_TRY(ctx, hipMemcpyAsync(out_u, d_u, b_u, hipMemcpyDeviceToHost, st));
        if (b_t) HIP_TRY(ctx, hipMemcpyAsync(out_t, d_t, b_t, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(out_len, d_len, b_i, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(retcode, d_ret, b_i, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
    }
    const int64_t ml = N > 0 ? max_len_host(out_len, N) : 0;
    if (stats) {
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->max_len = ml;
    }
    if (desc->adaptive && ml > cap)
        return fail(ctx, TPAMD_ERR_CAPACITY, "adaptive solve produced %lld records for a capacity of %lld",
                    (long long)ml, (long long)cap);
    return TPAMD_OK;
}

int32_t tpamd_adjoint(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* B,
                      const double* xv0, int64_t N, const double* targets, int32_t ckpt_interval,
                      double* loss, double* dL_dB, double* dL_dxv0, tpamd_stats* stats)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (!desc || !B || !loss) return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    if (N < 1 || !xv0 || !targets || !dL_dB || !dL_dxv0)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "adjoint needs N >= 1 and all buffers");
    if (desc->alg != TPAMD_ALG_BORIS || desc->adaptive)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the adjoint covers the fixed-step standard Boris pusher");
    if (B->d.kind != kFieldGrid) return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the adjoint needs a gridded B");
    if (desc->bnd_flags) return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the adjoint takes no boundary");
    if (desc->dt_pp) return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the adjoint takes one dt");
    if (ckpt_interval < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "checkpoint interval must be >= 1");
    tpamd_field zero;
    std::memset(&zero, 0, sizeof zero);
    zero.d.kind = kFieldZero;
    tpamd_solve_desc d = *desc;
    d.save_start = 0;
    d.save_end = 1;
    d.save_everystep = 0;
    d.nout_cap = 1;
    KParams p;
    int64_t nt = 0, nout = 0;
    int rc = fill_params(ctx, &d, &zero, B, nullptr, N, nullptr, nullptr, nullptr, nullptr, p, nt, nout);
    if (rc) return rc;
    if (nt < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "the adjoint needs at least one step");
    const bool f32 = desc->precision == TPAMD_FP32C;
    const bool bf32 = B->d.dtype == 1;
    if (f32 && !bf32)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the FP32C adjoint needs float (f32) B storage");
    // the forward pass runs the specialised pushers, which index the node array
    // with 32-bit byte offsets
    if (3.0 * (bf32 ? 4.0 : 8.0) * (double)B->d.n[0] * (double)B->d.n[1] * (double)B->d.n[2] >= 4294967296.0)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "the adjoint needs a node array below 4 GiB");
    const bool per = B->d.extrap == kExtrapPeriodic;
    const bool cells = B->d.cells != nullptr;      // gathers from the cell-replicated copy
    return adjoint_launch(ctx, p, B, xv0, N, targets, ckpt_interval, nt, loss, dL_dB, dL_dxv0, stats,
                          f32, bf32, per, cells);
}

int32_t tpamd_fo_segment(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
                         const tpamd_field* B, int64_t N, double dtmin, double dtmax,
                         int64_t check_interval, double* state, int64_t* iters, int32_t* flags,
                         double* sync, double* out_u, double* out_t, int32_t* out_len,
                         tpamd_stats* stats)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (!desc) return fail(ctx, TPAMD_ERR_ARGUMENT, "null descriptor");
    if (N < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "N must be >= 0");
    const int64_t cap = desc->nout_cap;
    if (cap < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "nout_cap must be >= 0");
    if (N > 0 && (!state || !iters || !flags || !sync || !out_len || (cap > 0 && (!out_u || !out_t))))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null buffer");
    if (desc->alg != TPAMD_ALG_BORIS || desc->adaptive || desc->precision != TPAMD_FP64)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "full-orbit segments run the FP64 Boris update (src/hybrid.jl:419)");
    if (desc->save_fields || desc->save_work)
        return fail(ctx, TPAMD_ERR_UNSUPPORTED, "full-orbit segments save (x, v) records");
    if (check_interval < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "check_interval must be positive.");
    if (desc->savestepinterval < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "savestepinterval must be >= 1");
    if (!(std::fabs(desc->q2m) > 0.0) || !std::isfinite(desc->t1))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "full-orbit segments need a nonzero q/m and a finite t_end");
    const auto t0 = std::chrono::steady_clock::now();
    // the parameter block of a one-step solve ending at t_end, then the
    // segment's own settings
    tpamd_solve_desc dd = *desc;
    dd.t0 = desc->t1 - 1.0;
    dd.dt = 1.0;
    dd.dt_pp = nullptr;
    dd.maxiters = 1;
    dd.save_start = dd.save_end = dd.save_everystep = 0;
    dd.nout_cap = 0;
    KParams p;
    int64_t nt = 0, nout = 0;
    int rc = fill_params(ctx, &dd, E, B, nullptr, N, nullptr, nullptr, nullptr, nullptr, p, nt, nout);
    if (rc) return rc;
    p.t1 = desc->t1;
    p.cap = cap;
    p.K = desc->savestepinterval;
    p.save_every = desc->save_everystep != 0;
    p.rec_w = 6;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_st = sizeof(double) * 8 * (size_t)N, b_it = sizeof(int64_t) * 2 * (size_t)N,
                 b_i = sizeof(int32_t) * (size_t)N, b_sy = sizeof(double) * 6 * (size_t)N,
                 b_u = sizeof(double) * 6 * (size_t)cap * (size_t)N, b_t = sizeof(double) * (size_t)cap * (size_t)N;
    const size_t need = al(b_st) + al(b_it) + 2 * al(b_i) + al(b_sy) + al(b_u) + al(b_t);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (need > ctx->d_buf_bytes) {
        if (ctx->d_buf) HIP_TRY(ctx, hipFree(ctx->d_buf));
        ctx->d_buf = nullptr;
        ctx->d_buf_bytes = 0;
        HIP_TRY(ctx, hipMalloc(&ctx->d_buf, need));
        ctx->d_buf_bytes = need;
    }
    char* b = (char*)ctx->d_buf;
    FoArgs a;
    a.state = (double*)b; b += al(b_st);
    a.iters = (long long*)b; b += al(b_it);
    a.flags = (int*)b; b += al(b_i);
    int32_t* d_len = (int32_t*)b; b += al(b_i);
    a.sync = (double*)b; b += al(b_sy);
    double* d_u = (double*)b; b += al(b_u);
    double* d_t = (double*)b;
    a.dtmin = dtmin;
    a.dtmax = dtmax;
    a.safety = desc->safety;
    a.check_interval = check_interval;
    a.maxiters = desc->maxiters;
    p.out_u = d_u;
    p.out_t = d_t;
    p.out_len = d_len;
    hipStream_t st = ctx->stream;
    float ms = 0.f;
    int64_t ml = 0;
    if (N > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(a.state, state, b_st, hipMemcpyHostToDevice, st));
        HIP_TRY(ctx, hipMemcpyAsync(a.iters, iters, b_it, hipMemcpyHostToDevice, st));
        HIP_TRY(ctx, hipMemcpyAsync(a.flags, flags, b_i, hipMemcpyHostToDevice, st));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
        launch_fo_segment(p, a, solve_class(p), st);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
        HIP_TRY(ctx, hipMemcpyAsync(state, a.state, b_st, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(iters, a.iters, b_it, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(flags, a.flags, b_i, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(sync, a.sync, b_sy, hipMemcpyDeviceToHost, st));
        if (b_u) HIP_TRY(ctx, hipMemcpyAsync(out_u, d_u, b_u, hipMemcpyDeviceToHost, st));
        if (b_t) HIP_TRY(ctx, hipMemcpyAsync(out_t, d_t, b_t, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(out_len, d_len, b_i, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ml = max_len_host(out_len, N);
    }
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->max_len = ml;
        stats->particle_steps = 0;
        stats->nt = 0;
        stats->nout = 0;
        stats->aux_ms = 0.0;
    }
    if (ml > cap)
        return fail(ctx, TPAMD_ERR_CAPACITY, "a full-orbit segment saved %lld records for a capacity of %lld",
                    (long long)ml, (long long)cap);
    return TPAMD_OK;
}

int32_t tpamd_batch_free(tpamd_batch* b)
{
    if (!b) return TPAMD_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->have_graph) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
    }
    if (b->d_table) (void)hipFree(b->d_table);
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    if (b->d_counter) (void)hipFree(b->d_counter);
    delete b;
    return TPAMD_OK;
}

int32_t tpamd_batch_create(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int32_t n_inst,
                           const tpamd_field* const* E, const tpamd_field* const* B,
                           const int64_t* offsets, const double* d_xv0, double* d_out_u,
                           double* d_out_t, int32_t* d_out_len, int32_t* d_retcode,
                           tpamd_batch** out)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (!desc || !E || !B || !offsets || !out || n_inst < 1)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "batch needs a descriptor, n_inst >= 1 fields and offsets");
    if (desc->dt_pp) return fail(ctx, TPAMD_ERR_UNSUPPORTED, "per-particle dt is not supported in batches");
    *out = nullptr;
    if (offsets[0] != 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "offsets[0] must be 0");
    for (int i = 0; i < n_inst; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(ctx, TPAMD_ERR_ARGUMENT, "offsets must be non-decreasing");
    const int64_t N = offsets[n_inst];
    if (N > 0 && (!d_xv0 || !d_out_len || !d_retcode || (desc->nout_cap > 0 && (!d_out_u || !d_out_t))))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null buffer");
    tpamd_batch* b = new tpamd_batch();
    b->ctx = ctx;
    b->desc = *desc;
    b->n_inst = n_inst;
    b->N = N;
    b->offsets.assign(offsets, offsets + n_inst + 1);
    b->table.resize(n_inst);
    b->d_table = nullptr;
    b->d_chunks = nullptr;
    b->nchunks = 0;
    b->chunks_cap = 0;
    b->d_counter = nullptr;
    b->d_steps = nullptr;
    b->block_size = desc->block_size > 0 ? desc->block_size : 256;
    b->ppt = desc->particles_per_lane > 0 ? desc->particles_per_lane : 4;
    b->have_graph = false;
    if (b->block_size % 64 != 0 || b->block_size > 256) {
        tpamd_batch_free(b);
        return fail(ctx, TPAMD_ERR_ARGUMENT, "block_size must be a multiple of 64 up to 256");
    }
    if (desc->particles_per_lane < 0 || desc->particles_per_lane > 1024) {
        tpamd_batch_free(b);
        return fail(ctx, TPAMD_ERR_ARGUMENT, "particles_per_lane must be in [0, 1024]");
    }
    int64_t nt, nout;
    for (int i = 0; i < n_inst; ++i) {
        const int rc = fill_params(ctx, desc, E[i], B[i], d_xv0, N, d_out_u, d_out_t, d_out_len,
                                   d_retcode, b->table[i], nt, nout);
        if (rc) {
            tpamd_batch_free(b);
            return rc;
        }
    }
    // a specialised policy only if every instance maps to the same one
    int em0 = 0;
    const int ns0 = spec_class(desc, b->table[0], &em0);
    b->allow_spec = ns0 >= 0;
    b->fk = kFkAnalytic;
    for (int i = 0; i < n_inst; ++i) b->fk = std::max(b->fk, solve_class(b->table[i]));
    for (int i = 1; i < n_inst && b->allow_spec; ++i) {
        int em = 0;
        const int ns = spec_class(desc, b->table[i], &em);
        b->allow_spec = ns == ns0 && em == em0 && b->table[i].B.dtype == b->table[0].B.dtype &&
                        (b->table[i].B.extrap == kExtrapPeriodic) == (b->table[0].B.extrap == kExtrapPeriodic) &&
                        (b->table[i].B.cells != nullptr) == (b->table[0].B.cells != nullptr);
    }
    hipDeviceProp_t prop;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, ctx->device);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_table, sizeof(KParams) * n_inst);
    // kQueues claim counters 128 B apart, then the step counter
    const size_t ctr_bytes = sizeof(unsigned) * kQueues * kQueueStride;
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_counter, ctr_bytes + 128);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_table, b->table.data(), sizeof(KParams) * n_inst, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        tpamd_batch_free(b);
        return fail(ctx, TPAMD_ERR_DEVICE, "batch set-up failed: %s", hipGetErrorString(e));
    }
    b->d_steps = (unsigned long long*)((char*)b->d_counter + ctr_bytes);
    b->n_cu = prop.multiProcessorCount;
    int rc = batch_chunks(b, b->ppt, 1);
    if (!rc) rc = batch_capture(b);
    if (rc) {
        tpamd_batch_free(b);
        return rc;
    }
    *out = b;
    return TPAMD_OK;
}

int32_t tpamd_solve_batch_noalloc(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int32_t n_inst,
                                  const tpamd_field* const* E, const tpamd_field* const* B,
                                  const int64_t* offsets, const double* d_xv0, double* d_out_u,
                                  double* d_out_t, int32_t* d_out_len, int32_t* d_retcode,
                                  void* stream, tpamd_stats* stats)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (!desc || !E || !B || !offsets || n_inst < 1)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "batch needs a descriptor, n_inst >= 1 fields and offsets");
    if (n_inst > kNoAllocMaxInst)
        return fail(ctx, TPAMD_ERR_CAPACITY, "the allocation-free batch takes at most %d instances (got %d)",
                    kNoAllocMaxInst, n_inst);
    if (desc->dt_pp) return fail(ctx, TPAMD_ERR_UNSUPPORTED, "per-particle dt is not supported in batches");
    if (offsets[0] != 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "offsets[0] must be 0");
    for (int i = 0; i < n_inst; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(ctx, TPAMD_ERR_ARGUMENT, "offsets must be non-decreasing");
    const int64_t N = offsets[n_inst];
    if (N > 0 && (!d_xv0 || !d_out_len || !d_retcode || (desc->nout_cap > 0 && (!d_out_u || !d_out_t))))
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null buffer");
    if (!ctx->scratch)
        return fail(ctx, TPAMD_ERR_DEVICE, "the scratch array's device address was not found at tpamd_init");
    const int bs = desc->block_size > 0 ? desc->block_size : 256;
    if (bs % 64 != 0 || bs > 256) return fail(ctx, TPAMD_ERR_ARGUMENT, "block_size must be a multiple of 64 up to 256");
    const int ppt = desc->particles_per_lane > 0 ? desc->particles_per_lane : 8;
    if (ppt > 1024) return fail(ctx, TPAMD_ERR_ARGUMENT, "particles_per_lane must be in [0, 1024]");
    // host-side set-up, as tpamd_batch_create does, on a batch object whose
    // device pointers are regions of the scratch array
    tpamd_batch b;
    b.ctx = ctx;
    b.desc = *desc;
    b.n_inst = n_inst;
    b.N = N;
    b.offsets.assign(offsets, offsets + n_inst + 1);
    b.table.resize(n_inst);
    b.block_size = bs;
    b.ppt = ppt;
    b.have_graph = false;
    int64_t nt, nout;
    for (int i = 0; i < n_inst; ++i) {
        const int rc = fill_params(ctx, desc, E[i], B[i], d_xv0, N, d_out_u, d_out_t, d_out_len, d_retcode,
                                   b.table[i], nt, nout);
        if (rc) return rc;
    }
    int em0 = 0;
    const int ns0 = spec_class(desc, b.table[0], &em0);
    b.allow_spec = ns0 >= 0;
    b.fk = kFkAnalytic;
    for (int i = 0; i < n_inst; ++i) b.fk = std::max(b.fk, solve_class(b.table[i]));
    for (int i = 1; i < n_inst && b.allow_spec; ++i) {
        int em = 0;
        const int ns = spec_class(desc, b.table[i], &em);
        b.allow_spec = ns == ns0 && em == em0 && b.table[i].B.dtype == b.table[0].B.dtype &&
                       (b.table[i].B.extrap == kExtrapPeriodic) == (b.table[0].B.extrap == kExtrapPeriodic) &&
                       (b.table[i].B.cells != nullptr) == (b.table[0].B.cells != nullptr);
    }
    const std::vector<Chunk> ch = make_chunks(b.offsets, n_inst, ppt, 1);
    if ((long long)ch.size() > kNoAllocMaxChunks)
        return fail(ctx, TPAMD_ERR_CAPACITY,
                    "the allocation-free batch takes at most %lld chunks of 64 x particles_per_lane particles "
                    "(got %lld at particles_per_lane %d): raise particles_per_lane or split the batch",
                    kNoAllocMaxChunks, (long long)ch.size(), ppt);
    b.nchunks = (int)ch.size();
    b.chunks_cap = 0;
    unsigned char* sp = ctx->scratch;
    b.d_counter = (unsigned*)sp;
    b.d_steps = (unsigned long long*)(sp + sizeof(unsigned) * kQueues * kQueueStride);
    b.d_table = (KParams*)(sp + kNoAllocCounterBytes);
    b.d_chunks = (Chunk*)(sp + kNoAllocCounterBytes + kNoAllocTableBytes);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int n_cu = 0;
    HIP_TRY(ctx, hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    b.n_cu = n_cu;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    // the table and the chunks into the scratch array; the counters are
    // cleared by the launch's own reset kernel (no memset node)
    HIP_TRY(ctx, hipMemcpyAsync(b.d_table, b.table.data(), sizeof(KParams) * n_inst, hipMemcpyHostToDevice, st));
    if (!ch.empty())
        HIP_TRY(ctx, hipMemcpyAsync(b.d_chunks, ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    const int rc = batch_enqueue(&b, st);
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    unsigned long long steps = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&steps, b.d_steps, sizeof steps, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->particle_steps = (int64_t)steps;
        stats->nt = b.table[0].nt;
        stats->nout = b.table[0].nout;
        stats->max_len = -1;
        stats->aux_ms = 0.0;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return TPAMD_OK;
}

int32_t tpamd_batch_run(tpamd_batch* b, void* stream, tpamd_stats* stats)
{
    if (!b) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null batch");
    tpamd_ctx* ctx = b->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    HIP_TRY(ctx, hipGraphLaunch(b->exec, st));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    unsigned long long steps = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&steps, b->d_steps, sizeof steps, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->particle_steps = (int64_t)steps;
        stats->nt = b->table[0].nt;
        stats->nout = b->table[0].nout;
        stats->max_len = -1;
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return TPAMD_OK;
}

int32_t tpamd_batch_autotune(tpamd_batch* b, void* stream, int32_t* block_size, int32_t* ppt)
{
    if (!b) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null batch");
    tpamd_ctx* ctx = b->ctx;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    // time each candidate on the full chunk list: a sample of the chunks
    // under-fills the GPU at large ppt (fewer chunks than resident waves) and
    // so biases the choice toward small ppt
    static const int kBlocks[] = {64, 128, 256};
    static const int kPpt[] = {1, 2, 4, 8, 16, 32};
    float best = 3.0e38f;
    int best_bs = b->block_size, best_ppt = b->ppt;
    for (int bs : kBlocks) {
        for (int pp : kPpt) {
            b->block_size = bs;
            int rc = batch_chunks(b, pp, 1);
            float ms = 0.f;
            if (!rc) rc = batch_time_once(b, st, &ms);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                best_bs = bs;
                best_ppt = pp;
            }
        }
    }
    b->block_size = best_bs;
    b->ppt = best_ppt;
    int rc = batch_chunks(b, best_ppt, 1);
    if (!rc) rc = batch_capture(b);
    if (rc) return rc;
    if (block_size) *block_size = best_bs;
    if (ppt) *ppt = best_ppt;
    return TPAMD_OK;
}

int32_t tpamd_solve_batch(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int32_t n_inst,
                          const tpamd_field* const* E, const tpamd_field* const* B,
                          const int64_t* offsets, const double* xv0, double* out_u,
                          double* out_t, int32_t* out_len, int32_t* retcode, tpamd_stats* stats)
{
    if (!ctx) return fail(nullptr, TPAMD_ERR_ARGUMENT, "null context");
    if (!desc || !offsets || n_inst < 1) return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
    const int64_t N = offsets[n_inst];
    if (N < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "bad offsets");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t cap = desc->nout_cap;
    const size_t b_xv = sizeof(double) * 6 * (size_t)N;
    const size_t b_u = sizeof(double) * (size_t)TPAMD_RECORD_WIDTH(desc) * (size_t)cap * (size_t)N;
    const size_t b_t = sizeof(double) * (size_t)cap * (size_t)N;
    const size_t b_i = sizeof(int32_t) * (size_t)N;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t need = al(b_xv) + al(b_u) + al(b_t) + 2 * al(b_i);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (need > ctx->d_buf_bytes) {
        if (ctx->d_buf) HIP_TRY(ctx, hipFree(ctx->d_buf));
        ctx->d_buf = nullptr;
        ctx->d_buf_bytes = 0;
        HIP_TRY(ctx, hipMalloc(&ctx->d_buf, need));
        ctx->d_buf_bytes = need;
    }
    char* p = (char*)ctx->d_buf;
    double* d_xv = (double*)p; p += al(b_xv);
    double* d_u = (double*)p; p += al(b_u);
    double* d_t = (double*)p; p += al(b_t);
    int32_t* d_len = (int32_t*)p; p += al(b_i);
    int32_t* d_ret = (int32_t*)p;
    hipStream_t st = ctx->stream;
    if (N > 0) HIP_TRY(ctx, hipMemcpyAsync(d_xv, xv0, b_xv, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    tpamd_batch* b = nullptr;
    int rc = tpamd_batch_create(ctx, desc, n_inst, E, B, offsets, d_xv, d_u, d_t, d_len, d_ret, &b);
    if (rc) return rc;
    rc = tpamd_batch_run(b, st, stats);
    tpamd_batch_free(b);
    if (rc) return rc;
    if (N > 0) {
        if (b_u) HIP_TRY(ctx, hipMemcpyAsync(out_u, d_u, b_u, hipMemcpyDeviceToHost, st));
        if (b_t) HIP_TRY(ctx, hipMemcpyAsync(out_t, d_t, b_t, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(out_len, d_len, b_i, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(retcode, d_ret, b_i, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
    }
    const int64_t ml = N > 0 ? max_len_host(out_len, N) : 0;
    if (stats) {
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->max_len = ml;
    }
    if (desc->adaptive && ml > cap)
        return fail(ctx, TPAMD_ERR_CAPACITY, "adaptive solve produced %lld records for a capacity of %lld",
                    (long long)ml, (long long)cap);
    return TPAMD_OK;
}

}  // extern "C"

// tpamd.hip -- kernels and C ABI of the MI355X ensemble Boris pusher.
//
// One thread per particle; the whole integration (all nt steps, every save)
// runs inside one launch, with the particle state in registers.  This
// replaces the reference's accelerator path, which launches one kernel per
// step, synchronises, and copies the full 6 x N state to the host for every
// saved record (src/boris/boris_kernel.jl:130-236).
#include "tpamd_device.h"
#include "tpamd_grid.h"
#include "tpamd.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

using namespace tpamd;

namespace {

// ---------------------------------------------------------------------------
// Precision policies
// ---------------------------------------------------------------------------
struct S64 {
    using R = double;
    struct Pos { double x[3]; };
    Pos pos;
    double v[3];
    double dt_cache;

    __device__ void load(const KParams& p, long long i)
    {
        const double* s = p.xv0 + 6 * i;
#pragma unroll
        for (int c = 0; c < 3; ++c) { pos.x[c] = s[c]; v[c] = s[3 + c]; }
    }
    __device__ void set_dt(double dt) { dt_cache = dt; }
    __device__ void fields(const KParams& p, double t, double E[3], double B[3]) const
    {
        field_f64(p.B, pos.x, t, B);
        if (p.e_is_uxb) cross3(B, p.E.v, E);       // E = -u x B = B x u
        else field_f64(p.E, pos.x, t, E);
    }
    // update_velocity (boris.jl:38-45) / MultistepUpdater (multistep_boris.jl:7-9)
    template <int ALG>
    __device__ void U(const KParams& p, const double vin[3], double h, double t, double out[3]) const
    {
        double E[3], B[3];
        fields(p, t, E, B);
        if (ALG == 0) boris_f64(vin, E, B, (p.q2m * 0.5) * h, out);
        else multistep_f64(p, vin, E, B, h, out);
    }
    __device__ double bmag(const KParams& p, double t) const
    {
        double B[3];
        field_f64(p.B, pos.x, t, B);
        return sqrt(dot3(B, B));
    }
    __device__ void step(const double vv[3], double dt, Pos& xn) const
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) xn.x[c] = pos.x[c] + vv[c] * dt;
    }
    __device__ void abs_pos(const KParams&, const Pos& q, double o[3]) const
    {
        o[0] = q.x[0]; o[1] = q.x[1]; o[2] = q.x[2];
    }
    __device__ void vel_d(const double vv[3], double o[3]) const
    {
        o[0] = vv[0]; o[1] = vv[1]; o[2] = vv[2];
    }
};

struct S32C {
    using R = float;
    using Pos = Pos32;
    Pos pos;
    float v[3];
    float dt_f, dt_hi, dt_lo;

    __device__ void load(const KParams& p, long long i) { load_from(p, p.xv0 + 6 * i); }
    __device__ void load_from(const KParams& p, const double* s)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double r = s[c] - p.frame_origin[c];
            int w = 0;
            if (p.frame_period[c] > 0.0) {
                const double q = floor(r / p.frame_period[c]);
                w = (int)q;
                r = r - q * p.frame_period[c];
            }
            pos.hi[c] = (float)r;
            pos.lo[c] = (float)(r - (double)pos.hi[c]);
            pos.wrap[c] = w;
            v[c] = (float)s[3 + c];
        }
    }
    __device__ void set_dt(double dt)
    {
        dt_hi = (float)dt;
        dt_lo = (float)(dt - (double)dt_hi);
        dt_f = dt_hi;
    }
    __device__ void fields(const KParams& p, const Pos& q, double t, float E[3], float B[3]) const
    {
        field_f32(p, p.B, q, t, B);
        if (p.e_is_uxb) {
            const float u[3] = {(float)p.E.v[0], (float)p.E.v[1], (float)p.E.v[2]};
            cross3f(B, u, E);
        } else {
            field_f32(p, p.E, q, t, E);
        }
    }
    template <int ALG>
    __device__ void U(const KParams& p, const float vin[3], double h, double t, float out[3]) const
    {
        float E[3], B[3];
        fields(p, pos, t, E, B);
        if (ALG == 0) boris_f32(vin, E, B, (float)((p.q2m * 0.5) * h), out);
        else multistep_f32(p, vin, E, B, (float)h, out);
    }
    __device__ double bmag(const KParams& p, double t) const
    {
        float B[3];
        field_f32(p, p.B, pos, t, B);
        return (double)sqrtf(dot3f(B, B));
    }
    __device__ void step_impl(const KParams& p, const float vv[3], Pos& xn) const
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float d = fmaf(vv[c], dt_hi, vv[c] * dt_lo);
            float s, e;
            two_sum(pos.hi[c], d + pos.lo[c], s, e);
            int w = pos.wrap[c];
            if (p.frame_periodic) {
                const float P = (float)p.frame_period[c];
                if (P > 0.0f && (s >= P || s < 0.0f)) {
                    const float Plo = (float)(p.frame_period[c] - (double)P);
                    const float sgn = s >= P ? 1.0f : -1.0f;
                    float s2, e2;
                    two_sum(s, -sgn * P, s2, e2);
                    two_sum(s2, fmaf(-sgn, Plo, e + e2), s, e);
                    w += (int)sgn;
                }
            }
            xn.hi[c] = s;
            xn.lo[c] = e;
            xn.wrap[c] = w;
        }
    }
    __device__ void abs_pos(const KParams& p, const Pos& q, double o[3]) const { pos_abs(p, q, o); }
    __device__ void vel_d(const float vv[3], double o[3]) const
    {
        o[0] = (double)vv[0]; o[1] = (double)vv[1]; o[2] = (double)vv[2];
    }
};

template <class S>
__device__ __forceinline__ bool outside(const KParams& p, const S& s, const typename S::Pos& xn,
                                        const typename S::R vv[3])
{
    const int f = p.bnd_flags;
    if (!f) return false;
    double x[3], v[3];
    s.abs_pos(p, xn, x);
    s.vel_d(vv, v);
    if (f & kBndNaN) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (x[c] != x[c] || v[c] != v[c]) return true;
    }
    if (f & (kBndSphereIn | kBndSphereOut)) {
        const double r = sqrt(dot3(x, x));
        if ((f & kBndSphereIn) && r < p.r_in) return true;
        if ((f & kBndSphereOut) && r > p.r_out) return true;
    }
    if (f & kBndBox) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (x[c] < p.box_lo[c] || x[c] > p.box_hi[c]) return true;
    }
    return false;
}

template <class S>
__device__ __forceinline__ void record(const KParams& p, const S& s, long long i, long long k,
                                       const typename S::R vv[3], double t)
{
    double x[3], v[3];
    s.abs_pos(p, s.pos, x);
    s.vel_d(vv, v);
    double* r = p.out_u + (i * p.cap + k) * 6;
    r[0] = x[0]; r[1] = x[1]; r[2] = x[2];
    r[3] = v[0]; r[4] = v[1]; r[5] = v[2];
    p.out_t[i * p.cap + k] = t;
}

template <class S>
__device__ __forceinline__ void do_step(const KParams& p, const S& s, const typename S::R vv[3],
                                        double dt, typename S::Pos& xn)
{
    if constexpr (sizeof(typename S::R) == 8) s.step(vv, dt, xn);
    else s.step_impl(p, vv, xn);
}

// _boris_single + _boris_loop!, src/boris/boris.jl:445-557.  The save test
// `(it - 1) % K == 0` is carried as a countdown to the next save step.
template <class S, int ALG>
__device__ long long run_fixed(const KParams& p, long long i)
{
    using R = typename S::R;
    S s;
    s.load(p, i);
    const double dt = p.dt_pp ? p.dt_pp[i] : p.dt;
    const double t0 = p.t0;
    const double t1 = p.dt_pp ? t0 + (double)p.nt * dt : p.t1;
    s.set_dt(dt);
    long long iout = 0;
    double tlast = 0.0;
    if (p.save_start) {                                       // :507-511
        record(p, s, i, iout, s.v, t0);
        tlast = t0;
        iout += 1;
    }
    R vt[3];
    s.template U<ALG>(p, s.v, -0.5 * dt, t0, vt);             // :514
#pragma unroll
    for (int c = 0; c < 3; ++c) s.v[c] = vt[c];

    const long long nt = p.nt;
    long long it = 1;
    long long next_save = p.save_every ? p.K + 1 : nt + 2;    // first it with (it-1) % K == 0, it-1 > 0
    double t = fma(-0.5, dt, t0);                             // :451
    for (; it <= nt; ++it) {
        R vprev[3] = {s.v[0], s.v[1], s.v[2]};
        t += dt;                                              // :454
        s.template U<ALG>(p, s.v, dt, t, vt);                 // :455
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
        typename S::Pos xn;
        do_step(p, s, s.v, dt, xn);                           // :457
        if (outside(p, s, xn, s.v)) {                         // :459-461
#pragma unroll
            for (int c = 0; c < 3; ++c) s.v[c] = vprev[c];
            break;
        }
        if (it == next_save) {                                // :463-472
            next_save += p.K;
            iout += 1;
            if (iout <= p.nout) {
                const double tc = fma(-0.5, dt, t);
                R vs[3];
                s.template U<ALG>(p, vprev, 0.5 * dt, tc, vs);
                record(p, s, i, iout - 1, vs, tc);
                tlast = tc;
            }
        }
        s.pos = xn;                                           // :474
    }
    const long long final_step = it - 1;
    bool should = p.save_end ||
                  (p.save_every && final_step > 0 && final_step % p.K == 0);     // :522-528
    if (iout < p.nout && should) {                                            // :530-543
        const double tf = (final_step == nt) ? t1 : fma((double)final_step, dt, t0);
        if (iout == 0 || tlast < tf) {
            const double dtf = tf - fma((double)final_step - 0.5, dt, t0);
            R vf[3];
            s.template U<ALG>(p, s.v, dtf, tf, vf);
            record(p, s, i, iout, vf, tf);
            iout += 1;
        }
    }
    p.out_len[i] = (int)iout;                                                 // :545-551
    p.retcode[i] = (iout < p.nout) ? kRetTerminated : kRetSuccess;
    return final_step;
}

// _adaptive_boris_single, src/boris/adaptive_boris.jl:145-266
template <class S, int ALG>
__device__ long long run_adaptive(const KParams& p, long long i)
{
    using R = typename S::R;
    S s;
    s.load(p, i);
    const double t0 = p.t0, t1 = p.t1;
    const double C = p.C;                                     // :164
    const double ttotal = t1 - t0;
    double t = t0;
    long long nrec = 0;
    double tlast = 0.0;
    R vt[3];
    auto push = [&](const R vv[3], double tt) {
        if (nrec < p.cap) record(p, s, i, nrec, vv, tt);
        tlast = tt;
        nrec += 1;
    };
    if (p.save_start) push(s.v, t);                           // :179-186
    double dt = C / s.bmag(p, t);                             // :188-189
    s.set_dt(dt);
    s.template U<ALG>(p, s.v, -0.5 * dt, t, vt);              // :192
#pragma unroll
    for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
    long long it = 1;
    long long next_save = p.save_every ? p.K + 1 : -1;
    bool should = p.save_end;
    int ret = kRetSuccess;
    // every wave must leave the loop: a zero step (|B| = inf) or an
    // unbounded run stops with MaxIters instead of spinning
    const long long cap_steps = p.max_steps > 0 ? p.max_steps : kAdaptiveHardCap;
    while (fabs(t - t0) < fabs(ttotal)) {                     // :197
        if (dt == 0.0) {
            ret = kRetMaxIters;
            break;
        }
        if (fabs(t + dt - t0) > fabs(ttotal)) {               // :198-203
            const double dts = t1 - t;
            s.template U<ALG>(p, s.v, 0.5 * dt, t, vt);
            s.template U<ALG>(p, vt, -0.5 * dts, t, s.v);
            dt = dts;
            s.set_dt(dt);
        }
        if (it == next_save) {                                // :205-218
            next_save += p.K;
            s.template U<ALG>(p, s.v, 0.5 * dt, t, vt);
            push(vt, t);
        }
        const double tmid = t + 0.5 * dt;                     // :220
        R vn[3];
        s.template U<ALG>(p, s.v, dt, tmid, vn);              // :221
        typename S::Pos xn;
        do_step(p, s, vn, dt, xn);                            // :223
        const double tn = t + dt;
        if (outside(p, s, xn, vn)) {                          // :227-231
            should = true;
            ret = kRetTerminated;
            break;
        }
        s.pos = xn;                                           // :233-235
        t = tn;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vn[c];
        const double dtn = C / s.bmag(p, t);                  // :237-238
        s.template U<ALG>(p, s.v, 0.5 * dt, t, vt);           // :241-242
        s.template U<ALG>(p, vt, -0.5 * dtn, t, s.v);
        dt = dtn;
        s.set_dt(dt);
        it += 1;
        if (it == next_save) should = true;                   // :245-247, (it-1) % K == 0
        if (it - 1 >= cap_steps) {
            ret = kRetMaxIters;
            break;
        }
    }
    if (should && (nrec == 0 || tlast != t)) {                // :250-259
        s.template U<ALG>(p, s.v, 0.5 * dt, t, vt);
        push(vt, t);
    }
    p.out_len[i] = (int)nrec;
    p.retcode[i] = ret;
    return it - 1;
}

__device__ __forceinline__ void add_steps(const KParams& p, long long steps)
{
    // wave-level sum, then one atomic per wave
    unsigned long long v = (unsigned long long)steps;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(p.steps, v);
}

template <class S, int ALG>
__global__ void __launch_bounds__(256) k_fixed(const KParams p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long steps = 0;
    if (i < p.N) steps = run_fixed<S, ALG>(p, i);
    add_steps(p, steps);
}

template <class S, int ALG>
__global__ void __launch_bounds__(256) k_adaptive(const KParams p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long steps = 0;
    if (i < p.N) steps = run_adaptive<S, ALG>(p, i);
    add_steps(p, steps);
}

// ---------------------------------------------------------------------------
// Specialised gridded-B policies (see tpamd_grid.h).  NS = 0 is the Boris
// update, NS >= 1 the rational multistep update with n = NS subcycles.
// ---------------------------------------------------------------------------
template <int EMODE, int NS, bool PER, typename T>
struct S64G : S64 {
    __device__ void bfield(const KParams& p, double B[3]) const
    {
        int i0[3], i1[3];
        double w[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double s = (pos.x[d] - p.B.origin[d]) * p.B.inv_h[d];
            if (!locate_f64<PER>(p.B, d, s, i0[d], i1[d], w[d])) {
                B[0] = B[1] = B[2] = __builtin_nan("");
                return;
            }
        }
        unsigned ox[2], oy[2], oz[2];
        offsets(p.B, i0, i1, sizeof(T), ox, oy, oz);
        trilinear<T, double>(p.B.data, ox, oy, oz, w, B);
    }
    template <int ALG>
    __device__ void U(const KParams& p, const double vin[3], double h, double t, double out[3]) const
    {
        double E[3], B[3];
        bfield(p, B);
        if (EMODE == kEUxB) cross3(B, p.E.v, E);
        else { E[0] = p.E.v[0]; E[1] = p.E.v[1]; E[2] = p.E.v[2]; }
        if (NS == 0) boris_f64(vin, E, B, (p.q2m * 0.5) * h, out);
        else multistep_f64_n<NS == 0 ? 1 : NS>(p, vin, E, B, h, out);
    }
    __device__ double bmag(const KParams& p, double) const
    {
        double B[3];
        bfield(p, B);
        return sqrt(dot3(B, B));
    }
};

template <int EMODE, int NS, bool PER>
struct S32G : S32C {
    __device__ void bfield(const KParams& p, float B[3]) const
    {
        int i0[3], i1[3];
        float w[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float s = fmaf(pos.hi[d], p.B.inv_hf[d], pos.lo[d] * p.B.inv_hf[d]);
            if (!locate_f32<PER>(p.B, d, s, i0[d], i1[d], w[d])) {
                B[0] = B[1] = B[2] = __builtin_nanf("");
                return;
            }
        }
        unsigned ox[2], oy[2], oz[2];
        offsets(p.B, i0, i1, 4u, ox, oy, oz);
        trilinear<float, float>(p.B.data, ox, oy, oz, w, B);
    }
    template <int ALG>
    __device__ void U(const KParams& p, const float vin[3], double h, double t, float out[3]) const
    {
        float E[3], B[3];
        bfield(p, B);
        if (EMODE == kEUxB) cross3f(B, p.f32.u, E);
        else { E[0] = p.f32.e0[0]; E[1] = p.f32.e0[1]; E[2] = p.f32.e0[2]; }
        if (NS == 0) boris_f32_fast(vin, E, B, (float)((p.q2m * 0.5) * h), out);
        else multistep_f32_n<NS == 0 ? 1 : NS>(p, vin, E, B, (float)(p.q2m_2n * h), out);
    }
    __device__ double bmag(const KParams& p, double) const
    {
        float B[3];
        bfield(p, B);
        return (double)sqrtf(dot3f(B, B));
    }
};

__global__ void __launch_bounds__(256) k_field_eval(const KParams p, const double* __restrict__ x,
                                                    long long n, double t, int which, int prec,
                                                    double* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (prec == 0) {
        S64 s;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.pos.x[c] = x[3 * i + c];
        double E[3], B[3];
        s.fields(p, t, E, B);
        const double* o = which ? E : B;
        for (int c = 0; c < 3; ++c) out[3 * i + c] = o[c];
    } else {
        S32C s;
        const double tmp[6] = {x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0, 0.0, 0.0};
        s.load_from(p, tmp);
        float E[3], B[3];
        s.fields(p, s.pos, t, E, B);
        const float* o = which ? E : B;
        for (int c = 0; c < 3; ++c) out[3 * i + c] = (double)o[c];
    }
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
struct tpamd_field {
    DevField d;
    void* dev_data;          // owned device buffer (grids)
    const tpamd_field* src;  // UxB: the B field it was built from
    size_t bytes;
};

struct tpamd_ctx {
    int device;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::string err;
    // staging buffers for tpamd_solve (grown on demand)
    void* d_buf;
    size_t d_buf_bytes;
    unsigned long long* d_steps;
};

namespace {

thread_local std::string g_init_err;

int fail(tpamd_ctx* ctx, int code, const char* fmt, ...)
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

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((ctx), e_ == hipErrorOutOfMemory ? TPAMD_ERR_OOM : TPAMD_ERR_DEVICE, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                     \
    } while (0)

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
    for (int d = 0; d < 3; ++d) { p.frame_origin[d] = 0.0; p.frame_period[d] = 0.0; }
    p.frame_periodic = 0;
    if (precision == TPAMD_FP32C) {
        const DevField* g = nullptr;
        if (B->d.kind == kFieldGrid) g = &B->d;
        if (!p.e_is_uxb && E->d.kind == kFieldGrid) {
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

template <class S>
void launch_solve(const KParams& p, int adaptive, int alg, dim3 grid, dim3 block, hipStream_t st)
{
    if (adaptive) {
        if (alg == 0) hipLaunchKernelGGL((k_adaptive<S, 0>), grid, block, 0, st, p);
        else hipLaunchKernelGGL((k_adaptive<S, 1>), grid, block, 0, st, p);
    } else {
        if (alg == 0) hipLaunchKernelGGL((k_fixed<S, 0>), grid, block, 0, st, p);
        else hipLaunchKernelGGL((k_fixed<S, 1>), grid, block, 0, st, p);
    }
}

// one specialised policy: fixed step for every NS, adaptive for Boris (NS = 0)
template <class S, int NS>
void launch_spec(const KParams& p, int adaptive, dim3 g, dim3 b, hipStream_t st)
{
    if constexpr (NS == 0) {
        if (adaptive) {
            hipLaunchKernelGGL((k_adaptive<S, 0>), g, b, 0, st, p);
            return;
        }
    }
    hipLaunchKernelGGL((k_fixed<S, NS == 0 ? 0 : 1>), g, b, 0, st, p);
}

template <int EMODE, int NS>
void launch_g32(const KParams& p, bool per, int adaptive, dim3 g, dim3 b, hipStream_t st)
{
    if (per) launch_spec<S32G<EMODE, NS, true>, NS>(p, adaptive, g, b, st);
    else launch_spec<S32G<EMODE, NS, false>, NS>(p, adaptive, g, b, st);
}

template <int EMODE, int NS, typename T>
void launch_g64(const KParams& p, bool per, int adaptive, dim3 g, dim3 b, hipStream_t st)
{
    if (per) launch_spec<S64G<EMODE, NS, true, T>, NS>(p, adaptive, g, b, st);
    else launch_spec<S64G<EMODE, NS, false, T>, NS>(p, adaptive, g, b, st);
}

template <int EMODE, typename T>
void launch_g64_ns(const KParams& p, int ns, bool per, int adaptive, dim3 g, dim3 b, hipStream_t st)
{
    if (ns == 0) launch_g64<EMODE, 0, T>(p, per, adaptive, g, b, st);
    else if (ns == 1) launch_g64<EMODE, 1, T>(p, per, adaptive, g, b, st);
    else launch_g64<EMODE, 2, T>(p, per, adaptive, g, b, st);
}

template <int EMODE>
void launch_g32_ns(const KParams& p, int ns, bool per, int adaptive, dim3 g, dim3 b, hipStream_t st)
{
    if (ns == 0) launch_g32<EMODE, 0>(p, per, adaptive, g, b, st);
    else if (ns == 1) launch_g32<EMODE, 1>(p, per, adaptive, g, b, st);
    else launch_g32<EMODE, 2>(p, per, adaptive, g, b, st);
}

// Launches a specialised kernel when the solve is a gridded-B one it covers;
// returns false (nothing launched) otherwise.
bool launch_specialised(const KParams& p, const tpamd_solve_desc* d, const tpamd_field* E,
                        const tpamd_field* B, dim3 g, dim3 b, hipStream_t st)
{
    if (d->flags & TPAMD_FLAG_GENERIC_KERNELS) return false;
    if (B->d.kind != kFieldGrid) return false;
    int emode;
    if (E->d.kind == kFieldUxB) emode = kEUxB;        // setup_fields checked E->src == B
    else if (E->d.kind == kFieldZero || E->d.kind == kFieldUniform) emode = kEUniform;
    else return false;
    int ns;
    if (d->alg == TPAMD_ALG_BORIS) ns = 0;
    else if (d->form == TPAMD_FORM_RATIONAL && (d->nsub == 1 || d->nsub == 2)) ns = d->nsub;
    else return false;
    if (d->adaptive && ns != 0) return false;
    const double es = B->d.dtype == 1 ? 4.0 : 8.0;
    if (3.0 * es * (double)B->d.n[0] * (double)B->d.n[1] * (double)B->d.n[2] >= 4294967296.0)
        return false;                                  // 32-bit byte offsets
    const bool per = B->d.extrap == kExtrapPeriodic;
    if (d->precision == TPAMD_FP32C) {
        if (B->d.dtype != 1) return false;
        if (emode == kEUxB) launch_g32_ns<kEUxB>(p, ns, per, d->adaptive, g, b, st);
        else launch_g32_ns<kEUniform>(p, ns, per, d->adaptive, g, b, st);
    } else if (B->d.dtype == 1) {
        if (emode == kEUxB) launch_g64_ns<kEUxB, float>(p, ns, per, d->adaptive, g, b, st);
        else launch_g64_ns<kEUniform, float>(p, ns, per, d->adaptive, g, b, st);
    } else {
        if (emode == kEUxB) launch_g64_ns<kEUxB, double>(p, ns, per, d->adaptive, g, b, st);
        else launch_g64_ns<kEUniform, double>(p, ns, per, d->adaptive, g, b, st);
    }
    return true;
}

int solve_impl(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
               const tpamd_field* B, const double* d_xv0, int64_t N, double* d_out_u,
               double* d_out_t, int32_t* d_len, int32_t* d_ret, hipStream_t st,
               tpamd_stats* stats)
{
    int64_t nt = 0, nout = 0;
    int rc = check_desc(ctx, desc, &nt, &nout);
    if (rc) return rc;
    if (N < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "N must be >= 0");
    if (!desc->adaptive && desc->nout_cap < nout)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "nout_cap (%lld) < records per particle (%lld)",
                    (long long)desc->nout_cap, (long long)nout);
    if (desc->nout_cap < 0) return fail(ctx, TPAMD_ERR_ARGUMENT, "nout_cap must be >= 0");
    KParams p;
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
    p.max_steps = desc->max_steps;
    p.N = N;
    p.r_in = desc->r_in;
    p.r_out = desc->r_out;
    for (int d = 0; d < 3; ++d) { p.box_lo[d] = desc->box_lo[d]; p.box_hi[d] = desc->box_hi[d]; }
    p.xv0 = d_xv0;
    p.dt_pp = desc->dt_pp;
    p.out_u = d_out_u;
    p.out_t = d_out_t;
    p.out_len = d_len;
    p.retcode = d_ret;
    p.steps = ctx->d_steps;

    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_steps, 0, sizeof(unsigned long long), st));
    const int bs = desc->block_size > 0 ? desc->block_size : 256;
    if (bs % 64 != 0 || bs > 256)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "block_size must be a multiple of 64 up to 256");
    float ms = 0.f;
    if (N > 0) {
        const dim3 block(bs), grid((unsigned)((N + bs - 1) / bs));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
        if (!launch_specialised(p, desc, E, B, grid, block, st)) {
            if (desc->precision == TPAMD_FP64) launch_solve<S64>(p, desc->adaptive, p.alg, grid, block, st);
            else launch_solve<S32C>(p, desc->adaptive, p.alg, grid, block, st);
        }
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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipMalloc((void**)&c->d_steps, sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return fail(nullptr, TPAMD_ERR_DEVICE, "HIP context set-up failed on device %d", device);
    }
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

int32_t tpamd_field_grid(tpamd_ctx* ctx, const void* data, int32_t dtype, int64_t nx, int64_t ny,
                         int64_t nz, const double origin[3], const double spacing[3],
                         int32_t extrap, int32_t storage_dtype, tpamd_field** out)
{
    if (!ctx || !out || !data || !origin || !spacing)
        return fail(ctx, TPAMD_ERR_ARGUMENT, "null argument");
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
    const size_t nn = (size_t)nx * (size_t)ny * (size_t)nz * 3;
    const size_t es = storage_dtype == TPAMD_F64 ? 8 : 4;
    void* dev = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc(&dev, nn * es));
    hipError_t e;
    if (dtype == storage_dtype) {
        e = hipMemcpy(dev, data, nn * es, hipMemcpyHostToDevice);
    } else {
        // convert on the host in chunks (one-time set-up cost)
        const size_t chunk = 1 << 22;
        char* tmp = (char*)std::malloc(chunk * es);
        e = tmp ? hipSuccess : hipErrorOutOfMemory;
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
    }
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return fail(ctx, TPAMD_ERR_DEVICE, "grid upload failed: %s", hipGetErrorString(e));
    }
    int rc = new_field(ctx, out, kFieldGrid);
    if (rc) {
        (void)hipFree(dev);
        return rc;
    }
    DevField& f = (*out)->d;
    f.extrap = extrap;
    f.dtype = storage_dtype == TPAMD_F64 ? 0 : 1;
    f.n[0] = (int)nx; f.n[1] = (int)ny; f.n[2] = (int)nz;
    for (int d = 0; d < 3; ++d) {
        f.origin[d] = origin[d];
        f.inv_h[d] = 1.0 / spacing[d];
        f.inv_hf[d] = (float)f.inv_h[d];
        f.period[d] = (double)f.n[d] * spacing[d];
    }
    f.data = dev;
    (*out)->dev_data = dev;
    (*out)->bytes = nn * es;
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
    if (f->dev_data) {
        if (ctx) (void)hipSetDevice(ctx->device);
        (void)hipFree(f->dev_data);
    }
    delete f;
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
    const size_t b_u = sizeof(double) * 6 * (size_t)cap * (size_t)N;
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

}  // extern "C"

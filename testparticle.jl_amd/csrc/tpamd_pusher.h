// tpamd_pusher.h -- the solve kernels: precision policies, the fixed-step
// and adaptive loops, the specialised gridded-B policies, the batched
// chunk-queue kernel and the launchers.  Each kernel family's translation
// unit includes this header and instantiates only its own kernels.
//
// One thread per particle; the whole integration (all nt steps, every save)
// runs inside one launch, with the particle state in registers.  This
// replaces the reference's accelerator path, which launches one kernel per
// step, synchronises, and copies the full 6 x N state to the host for every
// saved record (src/boris/boris_kernel.jl:130-236).
#pragma once

#include "tpamd_device.h"
#include "tpamd_grid.h"
#include "tpamd_internal.h"

#include <hip/hip_runtime.h>

#include <type_traits>

using namespace tpamd;

namespace {


// ---------------------------------------------------------------------------
// Precision policies
// ---------------------------------------------------------------------------
template <int FK>
struct S64T {
    static constexpr int kNS = -1;   // generic: any algorithm, adaptive or not
    static constexpr int kFK = FK;   // field-kind class (kFk*)
    static constexpr int kWaves = 1; // waves per SIMD the fixed-step kernel must fit (1: no bound)
    static constexpr bool kHoist = false;
    using R = double;
    struct Pos { double x[3]; };
    Pos pos;
    double v[3];
    double dt_cache;

    __device__ __forceinline__ void load(const KParams& p, long long i)
    {
        const auto s = gbl(p.xv0) + 6 * i;
#pragma unroll
        for (int c = 0; c < 3; ++c) { pos.x[c] = s[c]; v[c] = s[3 + c]; }
    }
    __device__ __forceinline__ void set_dt(double dt) { dt_cache = dt; }
    __device__ __forceinline__ void fields(const KParams& p, double t, double E[3], double B[3]) const
    {
        field_f64<FK>(p.B, pos.x, t, B);
        if (p.e_is_uxb) cross3(B, p.E.v, E);       // E = -u x B = B x u
        else field_f64<FK>(p.E, pos.x, t, E);
    }
    // fields at the current position, then the velocity update from given
    // fields: update_velocity (boris.jl:38-45) / MultistepUpdater
    // (multistep_boris.jl:7-9) is U = eval + Uf
    __device__ __forceinline__ void eval(const KParams& p, double t, double E[3], double B[3]) const
    {
        fields(p, t, E, B);
    }
    template <int ALG>
    __device__ __forceinline__ void Uf(const KParams& p, const double E[3], const double B[3], const double vin[3],
                       double h, double out[3]) const
    {
        if (ALG == 0) boris_f64(vin, E, B, (p.q2m * 0.5) * h, out);
        else multistep_f64(p, vin, E, B, h, out);
    }
    template <int ALG>
    __device__ __forceinline__ void U(const KParams& p, const double vin[3], double h, double t, double out[3]) const
    {
        double E[3], B[3];
        eval(p, t, E, B);
        Uf<ALG>(p, E, B, vin, h, out);
    }
    __device__ __forceinline__ double bmag_of(const double B[3]) const { return sqrt(dot3(B, B)); }
    __device__ __forceinline__ void step(const double vv[3], double dt, Pos& xn) const
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) xn.x[c] = pos.x[c] + vv[c] * dt;
    }
    __device__ __forceinline__ void abs_pos(const KParams&, const Pos& q, double o[3]) const
    {
        o[0] = q.x[0]; o[1] = q.x[1]; o[2] = q.x[2];
    }
    __device__ __forceinline__ void vel_d(const double vv[3], double o[3]) const
    {
        o[0] = vv[0]; o[1] = vv[1]; o[2] = vv[2];
    }
};

template <int FK>
struct S32CT {
    static constexpr int kNS = -1;
    static constexpr int kFK = FK;
    static constexpr int kWaves = 1;
    static constexpr bool kHoist = false;
    using R = float;
    using Pos = Pos32;
    Pos pos;
    float v[3];
    float dt_f, dt_hi, dt_lo;

    __device__ __forceinline__ void load(const KParams& p, long long i)
    {
        const auto g = gbl(p.xv0) + 6 * i;
        double s[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) s[c] = g[c];
        load_from(p, s);
    }
    // the state from x, v values (positions relative to the frame, wrapped)
    __device__ __forceinline__ void load_from(const KParams& p, const double s[6])
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
    __device__ __forceinline__ void set_dt(double dt)
    {
        dt_hi = (float)dt;
        dt_lo = (float)(dt - (double)dt_hi);
        dt_f = dt_hi;
    }
    __device__ __forceinline__ void fields(const KParams& p, const Pos& q, double t, float E[3], float B[3]) const
    {
        field_f32<FK>(p, p.B, q, t, B);
        if (p.e_is_uxb) {
            const float u[3] = {(float)p.E.v[0], (float)p.E.v[1], (float)p.E.v[2]};
            cross3f(B, u, E);
        } else {
            field_f32<FK>(p, p.E, q, t, E);
        }
    }
    __device__ __forceinline__ void eval(const KParams& p, double t, float E[3], float B[3]) const
    {
        fields(p, pos, t, E, B);
    }
    template <int ALG>
    __device__ __forceinline__ void Uf(const KParams& p, const float E[3], const float B[3], const float vin[3],
                       double h, float out[3]) const
    {
        if (ALG == 0) boris_f32(vin, E, B, (float)((p.q2m * 0.5) * h), out);
        else multistep_f32(p, vin, E, B, (float)h, out);
    }
    template <int ALG>
    __device__ __forceinline__ void U(const KParams& p, const float vin[3], double h, double t, float out[3]) const
    {
        float E[3], B[3];
        eval(p, t, E, B);
        Uf<ALG>(p, E, B, vin, h, out);
    }
    __device__ __forceinline__ double bmag_of(const float B[3]) const { return (double)sqrtf(dot3f(B, B)); }
    __device__ __forceinline__ void step_impl(const KParams& p, const float vv[3], Pos& xn) const
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float d = fmaf(vv[c], dt_hi, vv[c] * dt_lo);
            float s, e;
            two_sum(pos.hi[c], d + pos.lo[c], s, e);
            int w = pos.wrap[c];
            // analytic fields have no frame: no grid, no period (setup_fields)
            if (FK != kFkAnalytic && p.frame_periodic) {
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
    __device__ __forceinline__ void abs_pos(const KParams& p, const Pos& q, double o[3]) const { pos_abs_t<FK>(p, q, o); }
    __device__ __forceinline__ void vel_d(const float vv[3], double o[3]) const
    {
        o[0] = (double)vv[0]; o[1] = (double)vv[1]; o[2] = (double)vv[2];
    }
};

// generic policies per field-kind class: analytic fields, + grids, + meshes
using S64A = S64T<kFkAnalytic>;
using S64 = S64T<kFkGrid>;
using S64M = S64T<kFkMesh>;
using S32CA = S32CT<kFkAnalytic>;
using S32C = S32CT<kFkGrid>;
using S32CM = S32CT<kFkMesh>;

// a policy whose fixed-step Boris kernel forms the update's rotation vectors
// once per particle (run_fixed_impl's HOIST): for solves whose fields are the
// same at every x and t; a kernel of its own, so that the analytic-class
// kernels keep their register counts
template <class Base>
struct Hoisted : Base {
    static constexpr bool kHoist = true;
};

// a policy whose fields are fixed when the kernel is compiled: B the dipole,
// E zero (trapped particles in a planetary field).  Its kernel carries no
// field-kind switch; B is computed by the dipole case of field_f64 /
// field_f32 itself, so results equal the generic kernel's.
template <class Base>
struct DipoleOnly : Base {
    using R = typename Base::R;
    __device__ __forceinline__ void eval(const KParams& p, double, R E[3], R B[3]) const
    {
        E[0] = E[1] = E[2] = R(0);
        if constexpr (sizeof(R) == 8) dipole_f64(p.B, this->pos.x, B);
        else dipole_f32<Base::kFK>(p, p.B, this->pos, B);
    }
    template <int ALG>
    __device__ __forceinline__ void Uf(const KParams& p, const R E[3], const R B[3], const R vin[3], double h,
                                       R out[3]) const
    {
        if constexpr (sizeof(R) == 8 && ALG == 0) boris_f64_sd(vin, E, B, (p.q2m * 0.5) * h, out);
        else Base::template Uf<ALG>(p, E, B, vin, h, out);
    }
    template <int ALG>
    __device__ __forceinline__ void U(const KParams& p, const R vin[3], double h, double t, R out[3]) const
    {
        R E[3], B[3];
        eval(p, t, E, B);
        Uf<ALG>(p, E, B, vin, h, out);
    }
};

template <class S>
__device__ __forceinline__ bool outside(const KParams& p, const S& s, const typename S::Pos& xn,
                                        const typename S::R vv[3])
{
    const int f = p.bnd_flags;
    if (!f) return false;
    if constexpr (sizeof(typename S::R) == 4 && S::kFK == kFkAnalytic) {
        // FP32C without a frame, sphere tests only: decided in float away
        // from the radii (same outcome as the exact test below)
        if (!(f & ~(kBndSphereIn | kBndSphereOut))) {
            float xf[3];
            pos_f32<kFkAnalytic>(p, xn, xf);
            const float df = (xf[0] * xf[0] + xf[1] * xf[1]) + xf[2] * xf[2];
            const bool in_ok = !(f & kBndSphereIn) || df > p.f32.rin2_skip;
            const bool out_ok = !(f & kBndSphereOut) || df < p.f32.rout2_skip;
            if (in_ok && out_ok) return false;
        }
    }
    double x[3], v[3];
    s.abs_pos(p, xn, x);
    s.vel_d(vv, v);
    if (f & kBndNaN) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (x[c] != x[c] || v[c] != v[c]) return true;
    }
    if (f & (kBndSphereIn | kBndSphereOut)) {
        // r = sqrt(|x|^2) compared with the radius; away from the sphere the
        // squared norm alone decides it, near it the exact test runs
        const double d = dot3(x, x);
        if (f & kBndSphereIn) {
            if (d < p.rin2_lo) return true;
            if (d <= p.rin2_hi && sqrt(d) < p.r_in) return true;
        }
        if (f & kBndSphereOut) {
            if (d > p.rout2_hi) return true;
            if (d >= p.rout2_lo && sqrt(d) > p.r_out) return true;
        }
    }
    if (f & kBndBox) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (x[c] < p.box_lo[c] || x[c] > p.box_hi[c]) return true;
    }
    return false;
}

// extra record columns (E, B and/or work rates) at (x, v, t), in FP64.  Only
// the REC instantiations of the generic kernels contain it, so the pushers
// that save plain records keep their register budget.
template <int FK, class RP>
__device__ __forceinline__ void record_extra(const KParams& p, const double x[3], const double v[3],
                                          double t, RP r)
{
    double E[3], B[3], J[3][3], dBdt[3];
    field_jac_f64<FK>(p.B, x, t, B, J, dBdt);
    if (p.e_is_uxb) cross3(B, p.E.v, E);                  // E = -u x B = B x u
    else field_f64<FK>(p.E, x, t, E);
    int c = 6;
    if (p.save_fields) {
        r[6] = E[0]; r[7] = E[1]; r[8] = E[2];
        r[9] = B[0]; r[10] = B[1]; r[11] = B[2];
        c = 12;
    }
    if (p.save_work) {
        double w[4];
        work_rates_f64(p.q2m, p.mass, v, E, B, J, dBdt, w);
        r[c] = w[0]; r[c + 1] = w[1]; r[c + 2] = w[2]; r[c + 3] = w[3];
    }
}

// one record (_prepare_saved_data, boris.jl:405-440): x, v, then E and B,
// then the work rates, as the descriptor asks (REC kernels only)
template <bool REC, class S>
__device__ __forceinline__ void record(const KParams& p, const S& s, long long i, long long k,
                                       const typename S::R vv[3], double t)
{
    double x[3], v[3];
    s.abs_pos(p, s.pos, x);
    s.vel_d(vv, v);
    const auto r = gbl(p.out_u) + (i * p.cap + k) * p.rec_w;
    r[0] = x[0]; r[1] = x[1]; r[2] = x[2];
    r[3] = v[0]; r[4] = v[1]; r[5] = v[2];
    if constexpr (REC) record_extra<S::kFK>(p, x, v, t, r);
    gbl(p.out_t)[i * p.cap + k] = t;
}

template <class S>
__device__ __forceinline__ void do_step(const KParams& p, const S& s, const typename S::R vv[3],
                                        double dt, typename S::Pos& xn)
{
    if constexpr (sizeof(typename S::R) == 8) s.step(vv, dt, xn);
    else s.step_impl(p, vv, xn);
}

// the per-step update of fields that are the same everywhere, formed once
template <class S, bool F64 = sizeof(typename S::R) == 8>
struct BorisConst : BorisConst64 {
    __device__ __forceinline__ void setup(const KParams& p, const double E[3], const double B[3], double dt)
    {
        init(E, B, (p.q2m * 0.5) * dt);                      // as S64T::Uf forms h
    }
};
template <class S>
struct BorisConst<S, false> : BorisConst32 {
    __device__ __forceinline__ void setup(const KParams& p, const float E[3], const float B[3], double dt)
    {
        init(E, B, (float)((p.q2m * 0.5) * dt));             // as S32CT::Uf forms h
    }
};

// _boris_single + _boris_loop!, src/boris/boris.jl:445-557.  The save test
// `(it - 1) % K == 0` is carried as a countdown to the next save step.
// HOIST (Boris, fields the same at every x and t): the fields are evaluated
// once and the rotation vectors of the step's update formed once per
// particle; every step computes exactly what the plain update computes.
template <class S, int ALG, bool REC, bool HOIST>
__device__ __forceinline__ long long run_fixed_impl(const KParams& p, long long i)
{
    using R = typename S::R;
    S s;
    s.load(p, i);
    const double dt = p.dt_pp ? gbl(p.dt_pp)[i] : p.dt;
    const double t0 = p.t0;
    const double t1 = p.dt_pp ? t0 + (double)p.nt * dt : p.t1;
    s.set_dt(dt);
    long long iout = 0;
    double tlast = 0.0;
    if (p.save_start) {                                       // :507-511
        record<REC>(p, s, i, iout, s.v, t0);
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
    R Ec[3], Bc[3];
    BorisConst<S> bc;
    if constexpr (HOIST) {
        s.eval(p, t0, Ec, Bc);
        bc.setup(p, Ec, Bc, dt);
    }
    for (; it <= nt; ++it) {
        if constexpr (HOIST) {
            // with no boundary, the steps before the next save step need
            // neither the outside test nor the save test: the same update
            // and position step, without them
            if (!p.bnd_flags) {
                const long long stop = next_save - 1 < nt ? next_save - 1 : nt;
                for (; it <= stop; ++it) {
                    t += dt;
                    bc.apply(s.v, vt);
#pragma unroll
                    for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
                    typename S::Pos xn;
                    do_step(p, s, s.v, dt, xn);
                    s.pos = xn;
                }
                if (it > nt) break;
            }
        }
        R vprev[3] = {s.v[0], s.v[1], s.v[2]};
        t += dt;                                              // :454
        if constexpr (HOIST) {
            bc.apply(s.v, vt);                                // :455
        } else {
            s.eval(p, t, Ec, Bc);
            s.template Uf<ALG>(p, Ec, Bc, s.v, dt, vt);       // :455
        }
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
                // same position as the step's evaluation: static fields are re-used
                if (p.static_fields) s.template Uf<ALG>(p, Ec, Bc, vprev, 0.5 * dt, vs);
                else s.template U<ALG>(p, vprev, 0.5 * dt, tc, vs);
                record<REC>(p, s, i, iout - 1, vs, tc);
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
            record<REC>(p, s, i, iout, vf, tf);
            iout += 1;
        }
    }
    gbl(p.out_len)[i] = (int)iout;                                            // :545-551
    gbl(p.retcode)[i] = (iout < p.nout) ? kRetTerminated : kRetSuccess;
    return final_step;
}

template <class S, int ALG, bool REC>
__device__ __forceinline__ long long run_fixed(const KParams& p, long long i)
{
    if constexpr (S::kHoist && ALG == 0) {
        if (p.uniform_fields) return run_fixed_impl<S, ALG, REC, true>(p, i);
    }
    return run_fixed_impl<S, ALG, REC, false>(p, i);
}

// _adaptive_boris_single, src/boris/adaptive_boris.jl:145-266, as a
// resumable particle in three parts: begin (the prologue up to the first field
// evaluation, :160-186), then per loop pass settle (the fields at the current
// (x, t) and the step size they set: the rest of the prologue, :164,
// :188-192, for a new particle, the end of the previous pass, :237-247,
// otherwise) and advance (the pass up to the move, :197-235), and finish (the
// closing record, :250-259).  The kernels interleave many particles on one
// wave, handing a lane a new particle as soon as its previous one is done
// (adaptive_chunk); a new particle's first field evaluation and a running
// particle's resynchronisation are the same code, so a wave issues one
// gather and one dt computation per pass whatever mix of new and running
// particles it holds.  Each particle still sees the same operations in the
// same order as when it runs alone.
template <class S, int ALG, bool REC>
struct AdaptiveRun {
    using R = typename S::R;
    S s;
    long long i, it, next_save, nrec, cap_steps;
    double t, dt, tlast;
    bool should, fresh;
    int ret;
    // the fields at the current (x, t).  The reference evaluates there up to
    // four times per step (|B| for dt, both halves of the resynchronisation,
    // a save); one evaluation serves them all, and with static fields the
    // mid-step update as well.
    R Ec[3], Bc[3];

    __device__ __forceinline__ void push(const KParams& p, const R vv[3], double tt)
    {
        if (nrec < p.cap) record<REC>(p, s, i, nrec, vv, tt);
        tlast = tt;
        nrec += 1;
    }

    __device__ __forceinline__ void begin(const KParams& p, long long idx)
    {
        s.load(p, idx);
        start(p, idx);
    }
    // the prologue with the state already loaded into s
    __device__ __forceinline__ void start(const KParams& p, long long idx)
    {
        i = idx;
        t = p.t0;
        nrec = 0;
        tlast = 0.0;
        dt = 0.0;
        if (p.save_start) push(p, s.v, t);                   // :179-186
        it = 1;
        next_save = p.save_every ? p.K + 1 : -1;
        should = p.save_end;
        ret = kRetSuccess;
        // every particle must stop: a zero step (|B| = inf) or an unbounded
        // run ends with MaxIters instead of spinning
        cap_steps = p.max_steps > 0 ? p.max_steps : kAdaptiveHardCap;
        fresh = true;
    }

    // false once the particle is done
    __device__ __forceinline__ bool settle(const KParams& p)
    {
        s.eval(p, t, Ec, Bc);
        const double dtn = p.C / s.bmag_of(Bc);               // :164, :188-189 / :237-238
        R vt[3];
        s.template Uf<ALG>(p, Ec, Bc, s.v, 0.5 * dt, vt);     // :241-242 (running particles)
        if (fresh) {
#pragma unroll
            for (int c = 0; c < 3; ++c) vt[c] = s.v[c];
        }
        s.template Uf<ALG>(p, Ec, Bc, vt, -0.5 * dtn, s.v);   // :192 / :242
        dt = dtn;
        s.set_dt(dt);
        if (fresh) {
            fresh = false;
            return true;
        }
        it += 1;
        if (it == next_save) should = true;                   // :245-247, (it-1) % K == 0
        if (it - 1 >= cap_steps) {
            ret = kRetMaxIters;
            return false;
        }
        return true;
    }

    __device__ __forceinline__ bool advance(const KParams& p)
    {
        const double t0 = p.t0, t1 = p.t1, ttotal = t1 - t0;
        if (!(fabs(t - t0) < fabs(ttotal))) return false;    // :197
        if (dt == 0.0) {
            ret = kRetMaxIters;
            return false;
        }
        R vt[3];
        if (fabs(t + dt - t0) > fabs(ttotal)) {               // :198-203
            const double dts = t1 - t;
            s.template Uf<ALG>(p, Ec, Bc, s.v, 0.5 * dt, vt);
            s.template Uf<ALG>(p, Ec, Bc, vt, -0.5 * dts, s.v);
            dt = dts;
            s.set_dt(dt);
        }
        if (it == next_save) {                                // :205-218
            next_save += p.K;
            s.template Uf<ALG>(p, Ec, Bc, s.v, 0.5 * dt, vt);
            push(p, vt, t);
        }
        const double tmid = t + 0.5 * dt;                     // :220
        R vn[3];
        if (p.static_fields) s.template Uf<ALG>(p, Ec, Bc, s.v, dt, vn);   // :221
        else s.template U<ALG>(p, s.v, dt, tmid, vn);
        typename S::Pos xn;
        do_step(p, s, vn, dt, xn);                            // :223
        const double tn = t + dt;
        if (outside(p, s, xn, vn)) {                          // :227-231
            should = true;
            ret = kRetTerminated;
            return false;
        }
        s.pos = xn;                                           // :233-235
        t = tn;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vn[c];
        return true;
    }

    __device__ __forceinline__ long long finish(const KParams& p)
    {
        if (should && (nrec == 0 || tlast != t)) {            // :250-259
            R vt[3];
            s.template Uf<ALG>(p, Ec, Bc, s.v, 0.5 * dt, vt);
            push(p, vt, t);
        }
        gbl(p.out_len)[i] = (int)nrec;
        gbl(p.retcode)[i] = ret;
        return it - 1;
    }
};

__device__ __forceinline__ void add_steps(const KParams& p, long long steps)
{
    // wave-level sum, then one atomic per wave
    unsigned long long v = (unsigned long long)steps;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(p.steps, v);
}

template <class S, int ALG, bool REC = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S::kWaves))) k_fixed(const KParams p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long steps = 0;
    if (i < p.N) steps = run_fixed<S, ALG, REC>(p, i);
    add_steps(p, steps);
}

// ---------------------------------------------------------------------------
// Fixed-step Boris in fields that are the same everywhere, with no boundary,
// in FP64, one particle per quad of lanes: lane c of the quad holds
// component c of x and v (lane 3 repeats component 0 and stores nothing).
// A solve of few particles is one chain of dependent instructions per wave;
// split by component, each lane issues about half the instructions a step,
// and the other components of a vector come from the quad's other lanes by
// DPP quad permutations.  Every component is formed by the operations of
// boris_f64 and cross3 in their order, so the results are those of
// run_fixed_impl bit for bit.  Control flow depends on particle values only,
// so the four lanes of a quad always execute together.
// ---------------------------------------------------------------------------
namespace quad {

// every lane of a quad_perm reads a lane of its own quad, so no lane keeps
// an old value and none is needed
template <int CTRL>
__device__ __forceinline__ double perm(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
// components c+1 and c+2 (mod 3) of a vector held by the quad's lanes 0-2:
// quad_perm [1,2,0,3] and [2,0,1,3]
__device__ __forceinline__ double next1(double x) { return perm<0xC9>(x); }
__device__ __forceinline__ double next2(double x) { return perm<0xD2>(x); }
__device__ __forceinline__ double pick(const double a[3], int c) { return c == 0 ? a[0] : (c == 1 ? a[1] : a[2]); }

// boris_f64 with rotation factor h in uniform fields, for component c
struct Rot {
    double t1, t2, s1, s2, hE;    // t_{c+1}, t_{c+2}, s_{c+1}, s_{c+2}, h E_c
    __device__ __forceinline__ void init(const double E[3], const double B[3], double h, int c)
    {
        BorisConst64 k;
        k.init(E, B, h);
        const int c1 = c == 2 ? 0 : c + 1, c2 = c == 0 ? 2 : c - 1;
        t1 = pick(k.t, c1);
        t2 = pick(k.t, c2);
        s1 = pick(k.s, c1);
        s2 = pick(k.s, c2);
        hE = pick(k.hE, c);
    }
    __device__ __forceinline__ double apply(double v) const
    {
        const double vm = v + hE;
        const double vmt = next1(vm) * t2 - next2(vm) * t1;   // cross3(vm, t)_c
        const double vp = vm + vmt;
        const double vps = next1(vp) * s2 - next2(vp) * s1;   // cross3(vp, s)_c
        return (vm + vps) + hE;
    }
};

// the fields of S64T::fields for zero / uniform kinds
__device__ __forceinline__ void uniform_fields(const KParams& p, double E[3], double B[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) B[k] = p.B.kind == kFieldUniform ? p.B.v[k] : 0.0;
    if (p.e_is_uxb) {
        cross3(B, p.E.v, E);                                  // E = -u x B = B x u
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) E[k] = p.E.kind == kFieldUniform ? p.E.v[k] : 0.0;
    }
}

__device__ __forceinline__ void put(const KParams& p, long long i, long long k, int c, bool owner, bool lead,
                                    double x, double v, double t)
{
    const auto r = gbl(p.out_u) + (i * p.cap + k) * 6;
    if (owner) {
        r[c] = x;
        r[3 + c] = v;
    }
    if (lead) gbl(p.out_t)[i * p.cap + k] = t;
}

// run_fixed_impl<S64T, Boris, plain records, HOIST> with bnd_flags == 0
__device__ __forceinline__ long long run(const KParams& p, long long i, int c, bool owner, bool lead)
{
    const auto xv = gbl(p.xv0) + 6 * i;
    double x = xv[c], v = xv[3 + c];
    const double dt = p.dt_pp ? gbl(p.dt_pp)[i] : p.dt;
    const double t0 = p.t0;
    const double t1 = p.dt_pp ? t0 + (double)p.nt * dt : p.t1;
    double E[3], B[3];
    uniform_fields(p, E, B);
    long long iout = 0;
    double tlast = 0.0;
    if (p.save_start) {                                       // boris.jl:507-511
        put(p, i, 0, c, owner, lead, x, v, t0);
        tlast = t0;
        iout = 1;
    }
    Rot r;
    r.init(E, B, (p.q2m * 0.5) * (-0.5 * dt), c);             // :514
    v = r.apply(v);
    const long long nt = p.nt;
    long long it = 1;
    long long next_save = p.save_every ? p.K + 1 : nt + 2;
    double t = fma(-0.5, dt, t0);                             // :451
    Rot step, half;
    step.init(E, B, (p.q2m * 0.5) * dt, c);
    half.init(E, B, (p.q2m * 0.5) * (0.5 * dt), c);
    while (it <= nt) {
        const long long stop = next_save - 1 < nt ? next_save - 1 : nt;
        for (; it <= stop; ++it) {                            // :454-457, no save due
            t += dt;
            v = step.apply(v);
            x = x + v * dt;
        }
        if (it > nt) break;
        const double vprev = v;                               // it == next_save
        t += dt;
        v = step.apply(v);
        const double xn = x + v * dt;
        next_save += p.K;
        iout += 1;
        if (iout <= p.nout) {                                 // :463-472
            const double tc = fma(-0.5, dt, t);
            put(p, i, iout - 1, c, owner, lead, x, half.apply(vprev), tc);
            tlast = tc;
        }
        x = xn;
        ++it;
    }
    const long long final_step = it - 1;
    const bool should = p.save_end || (p.save_every && final_step > 0 && final_step % p.K == 0);
    if (iout < p.nout && should) {                            // :522-543
        const double tf = (final_step == nt) ? t1 : fma((double)final_step, dt, t0);
        if (iout == 0 || tlast < tf) {
            const double dtf = tf - fma((double)final_step - 0.5, dt, t0);
            r.init(E, B, (p.q2m * 0.5) * dtf, c);
            put(p, i, iout, c, owner, lead, x, r.apply(v), tf);
            iout += 1;
        }
    }
    if (lead) {                                               // :545-551
        gbl(p.out_len)[i] = (int)iout;
        gbl(p.retcode)[i] = (iout < p.nout) ? kRetTerminated : kRetSuccess;
    }
    return lead ? final_step : 0;
}

// FP32C: boris_f32 / BorisConst32 for component c, and the compensated
// position step of S32CT::step_impl for analytic fields (no frame)
__device__ __forceinline__ float permf(float x, int which)
{
    const int b = __float_as_int(x);
    return __int_as_float(which == 1 ? __builtin_amdgcn_mov_dpp(b, 0xC9, 0xF, 0xF, true)
                                     : __builtin_amdgcn_mov_dpp(b, 0xD2, 0xF, 0xF, true));
}
__device__ __forceinline__ float pickf(const float a[3], int c) { return c == 0 ? a[0] : (c == 1 ? a[1] : a[2]); }

struct Rot32 {
    float t1, t2, s1, s2, e, h;   // t_{c+1}, t_{c+2}, s_{c+1}, s_{c+2}, E_c, h
    __device__ __forceinline__ void init(const float E[3], const float B[3], float hv, int c)
    {
        BorisConst32 k;
        k.init(E, B, hv);
        const int c1 = c == 2 ? 0 : c + 1, c2 = c == 0 ? 2 : c - 1;
        t1 = pickf(k.t, c1);
        t2 = pickf(k.t, c2);
        s1 = pickf(k.s, c1);
        s2 = pickf(k.s, c2);
        e = pickf(k.E, c);
        h = k.h;
    }
    __device__ __forceinline__ float apply(float v) const
    {
        const float vm = fmaf(h, e, v);
        const float vmt = fmaf(permf(vm, 1), t2, -(permf(vm, 2) * t1));   // cross3f(vm, t)_c
        const float vp = vm + vmt;
        const float vps = fmaf(permf(vp, 1), s2, -(permf(vp, 2) * s1));   // cross3f(vp, s)_c
        return fmaf(h, e, vm + vps);
    }
};

// S32CT::fields for zero / uniform kinds, in float
__device__ __forceinline__ void uniform_fields32(const KParams& p, float E[3], float B[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) B[k] = p.B.kind == kFieldUniform ? (float)p.B.v[k] : 0.0f;
    if (p.e_is_uxb) {
        const float u[3] = {(float)p.E.v[0], (float)p.E.v[1], (float)p.E.v[2]};
        cross3f(B, u, E);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) E[k] = p.E.kind == kFieldUniform ? (float)p.E.v[k] : 0.0f;
    }
}

// record values of S32CT with no frame: 0 + (hi + lo) in double, (double) v
__device__ __forceinline__ void put32(const KParams& p, long long i, long long k, int c, bool owner, bool lead,
                                      float hi, float lo, float v, double t)
{
    put(p, i, k, c, owner, lead, 0.0 + ((double)hi + (double)lo), (double)v, t);
}

// run_fixed_impl<S32CT<analytic>, Boris, plain records, HOIST> with
// bnd_flags == 0
__device__ __forceinline__ long long run32(const KParams& p, long long i, int c, bool owner, bool lead)
{
    const auto xv = gbl(p.xv0) + 6 * i;
    const double r0 = xv[c] - p.frame_origin[c];              // S32CT::load_from (no period)
    float hi = (float)r0;
    float lo = (float)(r0 - (double)hi);
    float v = (float)xv[3 + c];
    const double dt = p.dt_pp ? gbl(p.dt_pp)[i] : p.dt;
    const float dt_hi = (float)dt;                            // S32CT::set_dt
    const float dt_lo = (float)(dt - (double)dt_hi);
    const double t0 = p.t0;
    const double t1 = p.dt_pp ? t0 + (double)p.nt * dt : p.t1;
    float E[3], B[3];
    uniform_fields32(p, E, B);
    long long iout = 0;
    double tlast = 0.0;
    if (p.save_start) {
        put32(p, i, 0, c, owner, lead, hi, lo, v, t0);
        tlast = t0;
        iout = 1;
    }
    Rot32 r;
    r.init(E, B, (float)((p.q2m * 0.5) * (-0.5 * dt)), c);
    v = r.apply(v);
    const long long nt = p.nt;
    long long it = 1;
    long long next_save = p.save_every ? p.K + 1 : nt + 2;
    double t = fma(-0.5, dt, t0);
    Rot32 step, half;
    step.init(E, B, (float)((p.q2m * 0.5) * dt), c);
    half.init(E, B, (float)((p.q2m * 0.5) * (0.5 * dt)), c);
    auto move = [&](float vv, float& h2, float& l2) {          // S32CT::step_impl, no frame
        const float d = fmaf(vv, dt_hi, vv * dt_lo);
        two_sum(hi, d + lo, h2, l2);
    };
    while (it <= nt) {
        const long long stop = next_save - 1 < nt ? next_save - 1 : nt;
        for (; it <= stop; ++it) {
            t += dt;
            v = step.apply(v);
            move(v, hi, lo);
        }
        if (it > nt) break;
        const float vprev = v;
        t += dt;
        v = step.apply(v);
        float hn, ln;
        move(v, hn, ln);
        next_save += p.K;
        iout += 1;
        if (iout <= p.nout) {
            const double tc = fma(-0.5, dt, t);
            put32(p, i, iout - 1, c, owner, lead, hi, lo, half.apply(vprev), tc);
            tlast = tc;
        }
        hi = hn;
        lo = ln;
        ++it;
    }
    const long long final_step = it - 1;
    const bool should = p.save_end || (p.save_every && final_step > 0 && final_step % p.K == 0);
    if (iout < p.nout && should) {
        const double tf = (final_step == nt) ? t1 : fma((double)final_step, dt, t0);
        if (iout == 0 || tlast < tf) {
            const double dtf = tf - fma((double)final_step - 0.5, dt, t0);
            r.init(E, B, (float)((p.q2m * 0.5) * dtf), c);
            put32(p, i, iout, c, owner, lead, hi, lo, r.apply(v), tf);
            iout += 1;
        }
    }
    if (lead) {
        gbl(p.out_len)[i] = (int)iout;
        gbl(p.retcode)[i] = (iout < p.nout) ? kRetTerminated : kRetSuccess;
    }
    return lead ? final_step : 0;
}

}  // namespace quad

// four threads per particle: threadIdx 4 i + c
__global__ void __launch_bounds__(256) k_fixed_quad(const KParams p)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long i = g >> 2;
    const int lc = (int)(g & 3);
    long long steps = 0;
    if (i < p.N) steps = quad::run(p, i, lc == 3 ? 0 : lc, lc < 3, lc == 0);
    add_steps(p, steps);
}

__global__ void __launch_bounds__(256) k_fixed_quad32(const KParams p)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long i = g >> 2;
    const int lc = (int)(g & 3);
    long long steps = 0;
    if (i < p.N) steps = quad::run32(p, i, lc == 3 ? 0 : lc, lc < 3, lc == 0);
    add_steps(p, steps);
}

// Adaptive particles [begin, end) of one problem on one wave: lives are
// uneven (a particle may leave after one step or run to the step cap), so
// instead of giving each lane a fixed list, which holds the wave until the
// longest life of every round ends, a lane takes the range's next particle
// as soon as its own is done.  The hand-out is a wave-uniform
// cursor advanced by a ballot: no atomics.
template <class S, int ALG, bool REC>
__device__ __forceinline__ long long adaptive_chunk(const KParams& p, long long begin, long long end,
                                                    int lane)
{
    long long steps = 0;
    AdaptiveRun<S, ALG, REC> r;
    bool active = begin + lane < end;
    if (active) r.begin(p, begin + lane);
    long long next = begin + 64;                          // wave-uniform cursor
    const unsigned long long below = (1ull << lane) - 1ull;
    while (true) {
        if (active && !(r.settle(p) && r.advance(p))) {
            steps += r.finish(p);
            active = false;
        }
        const unsigned long long idle = __ballot(!active);
        if (next >= end) {
            if (idle == ~0ull) break;                     // nothing left to hand out, all done
            continue;
        }
        if (!active) {
            const long long idx = next + (long long)__popcll(idle & below);
            if (idx < end) {
                r.begin(p, idx);
                active = true;
            }
        }
        next += (long long)__popcll(idle);
    }
    return steps;
}

// one wave per p.span consecutive particles, refilled lane by lane
template <class S, int ALG, bool REC = false>
__global__ void __launch_bounds__(256) k_adaptive(const KParams p)
{
    const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long begin = w * (long long)p.span;
    long long steps = 0;
    if (begin < p.N)
        steps = adaptive_chunk<S, ALG, REC>(p, begin, begin + p.span < p.N ? begin + p.span : p.N,
                                            (int)(threadIdx.x & 63));
    add_steps(p, steps);
}

// ---------------------------------------------------------------------------
// Specialised gridded-B policies (see tpamd_grid.h).  NS = 0 is the Boris
// update, 1 <= NS < kRefBase the rational multistep update with n = NS
// subcycles, NS = kRefBase + n the reference-form multistep update.
// ---------------------------------------------------------------------------
template <int EMODE, int NS, bool PER, typename T, bool CELLS>
struct S64G : S64 {
    static constexpr int kNS = NS;
    // FP64 periodic cells with E = -u x B: an occupancy target for the
    // compiler (amdgpu_waves_per_eu).  Measured at 256^3, MultistepBoris4(n=2):
    // the rational form at 5 waves spills 56 B per lane (2.36e10
    // particle-steps/s), at 4 waves it fits 120 VGPRs with no spill (2.58e10);
    // with no target the same 120 VGPRs ran at 2.27e10.  The reference form
    // (atan, sincos, three divisions) holds more live values: at 5 waves it
    // spilled 200 B (1.80e10), at 4 waves 96 B (2.12e10), at 3 it takes 163
    // VGPRs and no spill (2.61e10); with no target, 2.27e10.  With a stored
    // uniform E a 5-wave bound spilled 100-136 B and lost 20 %, and the
    // non-periodic variants would spill 76-84 B: no target there.
#ifndef TPAMD_RAT_WAVES
#define TPAMD_RAT_WAVES 4
#endif
#ifndef TPAMD_REF_WAVES
#define TPAMD_REF_WAVES 3
#endif
    static constexpr int kWaves = (EMODE == kEUxB && PER && CELLS && sizeof(T) == 8)
                                      ? (NS < kRefBase ? TPAMD_RAT_WAVES : TPAMD_REF_WAVES) : 1;
    __device__ __forceinline__ void bfield(const KParams& p, double B[3]) const
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
        if constexpr (CELLS) {
            trilinear_cell<T, double>(p.B.cells, cell_offset(p.B, i0), w, B);
        } else {
            unsigned ox[2], oy[2], oz[2];
            offsets(p.B, i0, i1, sizeof(T), ox, oy, oz);
            trilinear<T, double>(p.B.data, ox, oy, oz, w, B);
        }
    }
    __device__ __forceinline__ void eval(const KParams& p, double, double E[3], double B[3]) const
    {
        bfield(p, B);
        if (EMODE == kEUxB) cross3(B, p.E.v, E);
        else { E[0] = p.E.v[0]; E[1] = p.E.v[1]; E[2] = p.E.v[2]; }
    }
    template <int ALG>
    __device__ __forceinline__ void Uf(const KParams& p, const double E[3], const double B[3], const double vin[3],
                       double h, double out[3]) const
    {
        if (NS == 0) boris_f64(vin, E, B, (p.q2m * 0.5) * h, out);
        else multistep_f64_n<NS == 0 ? 1 : NS>(p, vin, E, B, h, out);
    }
    template <int ALG>
    __device__ __forceinline__ void U(const KParams& p, const double vin[3], double h, double t, double out[3]) const
    {
        double E[3], B[3];
        eval(p, t, E, B);
        Uf<ALG>(p, E, B, vin, h, out);
    }
};

template <int EMODE, int NS, bool PER, bool CELLS>
struct S32G : S32C {
    static constexpr int kNS = NS;
    __device__ __forceinline__ void bfield(const KParams& p, float B[3]) const
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
        if constexpr (CELLS) {
            trilinear_cell<float, float>(p.B.cells, cell_offset(p.B, i0), w, B);
        } else {
            unsigned ox[2], oy[2], oz[2];
            offsets(p.B, i0, i1, 4u, ox, oy, oz);
            trilinear<float, float>(p.B.data, ox, oy, oz, w, B);
        }
    }
    __device__ __forceinline__ void eval(const KParams& p, double, float E[3], float B[3]) const
    {
        bfield(p, B);
        if (EMODE == kEUxB) cross3f(B, p.f32.u, E);
        else { E[0] = p.f32.e0[0]; E[1] = p.f32.e0[1]; E[2] = p.f32.e0[2]; }
    }
    template <int ALG>
    __device__ __forceinline__ void Uf(const KParams& p, const float E[3], const float B[3], const float vin[3],
                       double h, float out[3]) const
    {
        if (NS == 0) boris_f32_fast(vin, E, B, (float)((p.q2m * 0.5) * h), out);
        else multistep_f32_n<NS == 0 ? 1 : NS>(p, vin, E, B, (float)(p.q2m_2n * h), out);
    }
    template <int ALG>
    __device__ __forceinline__ void U(const KParams& p, const float vin[3], double h, double t, float out[3]) const
    {
        float E[3], B[3];
        eval(p, t, E, B);
        Uf<ALG>(p, E, B, vin, h, out);
    }
};

// The batch kernel's work source: the chunks of the kQueues slices, claimed
// one at a time by a wave.  The chunk index is made wave-uniform
// (readfirstlane), so the chunk and its instance's parameter block are read
// with scalar loads.  The index of the chunk after the current one is
// claimed when the current one is taken (reading the claim later, to overlap
// the atomic with work, kept a value live through the particle loop and cost
// 15 VGPRs: 6 waves per SIMD down to 5).  A wave starts on slice
// blockIdx % kQueues and moves to the next slice once its own is drained.
struct ChunkSource {
    const Chunk* chunks;
    unsigned* counter;
    int nchunks, lane, j;
    unsigned q0, q, nq, c, pending;

    __device__ __forceinline__ unsigned claim() const
    {
        unsigned v = 0;
        if (lane == 0) v = atomicAdd(counter + q * kQueueStride, 1u);
        return (unsigned)__builtin_amdgcn_readfirstlane((int)v);
    }
    __device__ __forceinline__ void open_slice()
    {
        q = (q0 + (unsigned)j) % kQueues;
        nq = (unsigned)nchunks > q ? ((unsigned)nchunks - q + kQueues - 1) / kQueues : 0u;
        c = claim();
        pending = c < nq ? claim() : nq;
    }
    __device__ __forceinline__ void init(const Chunk* ch, unsigned* ctr, int n, int ln)
    {
        chunks = ch;
        counter = ctr;
        nchunks = n;
        lane = ln;
        q0 = blockIdx.x % kQueues;
        j = 0;
        open_slice();
    }
    // the next chunk of this wave, false once every slice is drained
    __device__ __forceinline__ bool next(Chunk& out)
    {
        while (c >= nq) {
            if (++j >= kQueues) return false;
            open_slice();
        }
        out = chunks[q + c * kQueues];
        c = pending;
        pending = c < nq ? claim() : nq;
        return true;
    }
};

// Adaptive particles of every chunk the wave claims: a lane takes the next
// particle as soon as its own is done (see adaptive_chunk).  Lives are
// uneven (in a turbulent box most particles leave within a step or two,
// the longest of a 512-particle chunk lives tens of steps), so when a chunk's
// last particle has been handed out the wave claims its next chunk at once
// and keeps feeding idle lanes from it while the old chunk's long-lived
// particles finish, instead of idling until they do.  That needs the new
// chunk to be of the same instance (the lanes share one parameter block);
// a chunk of another instance waits until every lane is done.
template <class S, int ALG, bool REC>
__device__ __forceinline__ long long batch_adaptive(const KParams* __restrict__ table, ChunkSource& src,
                                                    int lane)
{
    long long steps = 0;
    AdaptiveRun<S, ALG, REC> r;
    const unsigned long long below = (1ull << lane) - 1ull;
    Chunk ch, nch;
    bool have = src.next(ch);
    while (have) {
        // one instance at a time: its parameter block stays in scalar
        // registers (the index is made wave-uniform explicitly)
        const int inst = __builtin_amdgcn_readfirstlane(ch.inst);
        const KParams& p = table[inst];
        long long next = ch.begin, end = ch.end;          // wave-uniform cursor
        bool held = false;                                // nch: claimed, other instance
        bool drained = false;                             // no chunk left to claim
        bool active = false;
        while (true) {
            if (active && !(r.settle(p) && r.advance(p))) {
                steps += r.finish(p);
                active = false;
            }
            const unsigned long long idle = __ballot(!active);
            while (next >= end && !held && !drained) {
                if (!src.next(nch)) {
                    drained = true;
                } else if (__builtin_amdgcn_readfirstlane(nch.inst) == inst) {
                    next = nch.begin;
                    end = nch.end;
                } else {
                    held = true;
                }
            }
            if (next >= end) {
                if (idle != ~0ull) continue;              // running particles still to finish
                break;
            }
            if (!active) {
                const long long idx = next + (long long)__popcll(idle & below);
                if (idx < end) {
                    r.begin(p, idx);
                    active = true;
                }
            }
            next += (long long)__popcll(idle);
        }
        ch = nch;
        have = held;
    }
    return steps;
}

template <class S, int ALG, bool ADAPT, bool REC = false>
__global__ void __launch_bounds__(256) k_batch(const KParams* __restrict__ table, const Chunk* __restrict__ chunks, int nchunks,
        unsigned* __restrict__ counter, unsigned long long* __restrict__ steps_out)
{
    const int lane = threadIdx.x & 63;
    long long steps = 0;
    ChunkSource src;
    src.init(chunks, counter, nchunks, lane);
    if constexpr (ADAPT) {
        steps = batch_adaptive<S, ALG, REC>(table, src, lane);
    } else {
        Chunk ch;
        while (src.next(ch)) {
            const KParams& p = table[ch.inst];
            for (long long i = ch.begin + lane; i < ch.end; i += 64) steps += run_fixed<S, ALG, REC>(p, i);
        }
    }
    unsigned long long v = (unsigned long long)steps;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) atomicAdd(steps_out, v);
}

// ---------------------------------------------------------------------------
// Policy selection.  select_policy calls l.go<S, ALG>() with the specialised
// gridded-B policy that matches the solve, or with the generic policy.  The
// specialisations cover: B a grid whose bytes fit 32-bit offsets, E zero,
// uniform or UxB, the Boris update (fixed or adaptive) or the rational
// multistep update with n <= 2 (fixed step), FP32C on float storage.
// ---------------------------------------------------------------------------
template <class L, int EMODE, int NS, bool CELLS>
void sel_g32(const L& l, bool per)
{
    if (per) l.template go<S32G<EMODE, NS, true, CELLS>, NS == 0 ? 0 : 1>();
    else l.template go<S32G<EMODE, NS, false, CELLS>, NS == 0 ? 0 : 1>();
}

template <class L, int EMODE, int NS, typename T, bool CELLS>
void sel_g64(const L& l, bool per)
{
    if (per) l.template go<S64G<EMODE, NS, true, T, CELLS>, NS == 0 ? 0 : 1>();
    else l.template go<S64G<EMODE, NS, false, T, CELLS>, NS == 0 ? 0 : 1>();
}

template <class L, int EMODE, bool CELLS>
void sel_g32_ns(const L& l, int ns, bool per)
{
    if (ns == 0) sel_g32<L, EMODE, 0, CELLS>(l, per);
    else if (ns == 1) sel_g32<L, EMODE, 1, CELLS>(l, per);
    else sel_g32<L, EMODE, 2, CELLS>(l, per);
}

template <class L, int EMODE, typename T, bool CELLS>
void sel_g64_ns(const L& l, int ns, bool per)
{
    if (ns == 0) sel_g64<L, EMODE, 0, T, CELLS>(l, per);
    else if (ns == 1) sel_g64<L, EMODE, 1, T, CELLS>(l, per);
    else sel_g64<L, EMODE, 2, T, CELLS>(l, per);
}

// the reference-form classes kRefBase + 1 and kRefBase + 2
template <class L, int EMODE, bool CELLS>
void sel_g32_nsr(const L& l, int ns, bool per)
{
    if (ns == kRefBase + 1) sel_g32<L, EMODE, kRefBase + 1, CELLS>(l, per);
    else sel_g32<L, EMODE, kRefBase + 2, CELLS>(l, per);
}

template <class L, int EMODE, typename T, bool CELLS>
void sel_g64_nsr(const L& l, int ns, bool per)
{
    if (ns == kRefBase + 1) sel_g64<L, EMODE, kRefBase + 1, T, CELLS>(l, per);
    else sel_g64<L, EMODE, kRefBase + 2, T, CELLS>(l, per);
}

// one solve: a thread per particle
struct SoloLaunch {
    const KParams& p;
    int adaptive;
    dim3 g, b;
    hipStream_t st;
    template <class S, int ALG>
    void go() const
    {
        // the hoisted policies are launched for fixed-step Boris with plain
        // records only, so no other kernel of theirs is instantiated
        if constexpr (S::kHoist) {
            hipLaunchKernelGGL((k_fixed<S, ALG>), g, b, 0, st, p);
        } else {
            // records with E, B or work columns: generic policies only
            // (spec_class declines them), in the REC instantiations
            if constexpr (S::kNS < 0) {
                if (p.rec_w > 6) {
                    if (adaptive) hipLaunchKernelGGL((k_adaptive<S, ALG, true>), g, b, 0, st, p);
                    else hipLaunchKernelGGL((k_fixed<S, ALG, true>), g, b, 0, st, p);
                    return;
                }
            }
            if constexpr (S::kNS <= 0) {
                if (adaptive) {
                    hipLaunchKernelGGL((k_adaptive<S, ALG>), g, b, 0, st, p);
                    return;
                }
            }
            hipLaunchKernelGGL((k_fixed<S, ALG>), g, b, 0, st, p);
        }
    }
};

// batched solve: persistent waves over the chunk queue.  With `occ` set it
// only reports the kernel's resident blocks per CU for block size b.x.
struct BatchLaunch {
    const KParams* table;
    const Chunk* chunks;
    int nchunks;
    unsigned* counter;
    unsigned long long* steps;
    int adaptive;
    dim3 g, b;
    hipStream_t st;
    int* occ;
    bool rec;                 // records carry E, B or work columns
    template <class K>
    void run(K kern) const
    {
        if (occ) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, (int)b.x, 0) != hipSuccess)
                *occ = 1;
            return;
        }
        hipLaunchKernelGGL(kern, g, b, 0, st, table, chunks, nchunks, counter, steps);
    }
    template <class S, int ALG>
    void go() const
    {
        if constexpr (S::kNS < 0) {
            if (rec) {
                if (adaptive) run(k_batch<S, ALG, true, true>);
                else run(k_batch<S, ALG, false, true>);
                return;
            }
        }
        if constexpr (S::kNS <= 0) {
            if (adaptive) {
                run(k_batch<S, ALG, true>);
                return;
            }
        }
        run(k_batch<S, ALG, false>);
    }
};

template <class F>
void with_launcher(const LaunchReq& r, const F& f)
{
    if (r.batch)
        f(BatchLaunch{r.table, r.chunks, r.nchunks, r.counter, r.steps, r.adaptive, r.g, r.b, r.st,
                      r.occ, r.rec});
    else
        f(SoloLaunch{*r.p, r.adaptive, r.g, r.b, r.st});
}

}  // namespace

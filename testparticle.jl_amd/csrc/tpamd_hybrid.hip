// tpamd_hybrid.hip -- the full-orbit segments of the hybrid guiding-centre /
// full-orbit solver (_hybrid_adaptive!, src/hybrid.jl:257-452), FP64, one
// thread per particle.  A particle runs from its current state to its next
// adiabaticity check point, t_end, the step cap or its isoutside; the caller
// makes every mode decision.
#include "tpamd_pusher.h"

namespace {

enum { kFoCheck = TPAMD_FO_CHECK, kFoEnd = TPAMD_FO_END, kFoOutside = TPAMD_FO_OUTSIDE };

// clamp(x, lo, hi) as Julia defines it: x > hi ? hi : (x < lo ? lo : x)
__device__ __forceinline__ double jl_clamp(double x, double lo, double hi)
{
    return x > hi ? hi : (x < lo ? lo : x);
}

template <class S>
__global__ void __launch_bounds__(256) k_fo_segment(const KParams p, const FoArgs a)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.N) return;
    S s;
    double* st = a.state + 8 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.pos.x[c] = st[c];
        s.v[c] = st[3 + c];
    }
    double t = st[6], dt = st[7];
    long long it = a.iters[2 * i], steps = a.iters[2 * i + 1];
    const int fl = a.flags[i];
    bool checked = (fl & TPAMD_FO_CHECKED) != 0;
    double E[3], B[3], vt[3];
    if (fl & TPAMD_FO_ENTER) {
        // _fo_dt (:244-249) and the back half-step at t (:265 / :302), both at (x, t)
        s.eval(p, t, E, B);
        const double om = fabs(p.q2m) * sqrt(dot3(B, B));
        dt = jl_clamp((2.0 * M_PI) * a.safety / om, a.dtmin, a.dtmax);
        s.template Uf<0>(p, E, B, s.v, -0.5 * dt, vt);
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
    }
    int status;
    long long nrec = 0;
    for (;;) {
        if (!(t < p.t1 && steps < a.maxiters)) {             // :339
            status = kFoEnd;
            break;
        }
        if (it % a.check_interval == 0 && !checked) {       // :359
            status = kFoCheck;
            break;
        }
        checked = false;
        if (t + dt > p.t1) dt = p.t1 - t;                    // :412-414
        const double vprev[3] = {s.v[0], s.v[1], s.v[2]};   // :418
        s.eval(p, t + 0.5 * dt, E, B);                       // :419
        s.template Uf<0>(p, E, B, vprev, dt, vt);
#pragma unroll
        for (int c = 0; c < 3; ++c) s.v[c] = vt[c];
        typename S::Pos xn;
        s.step(s.v, dt, xn);                                  // :420
        const double tn = t + dt;                             // :421
        if (outside(p, s, xn, s.v)) {                         // :422-424
            status = kFoOutside;
            break;
        }
        if (p.save_every && it - 1 > 0 && (it - 1) % p.K == 0) {   // :426-430
            s.template U<0>(p, vprev, 0.5 * dt, t, vt);
            if (nrec < p.cap) record<false>(p, s, i, nrec, vt, t);
            nrec += 1;
        }
        s.pos = xn;                                           // :432-435
        t = tn;
        it += 1;
        steps += 1;
    }
    // the synchronised state at t (:361-364, :446-449)
    s.template U<0>(p, s.v, 0.5 * dt, t, vt);
    double* sy = a.sync + 6 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        sy[c] = s.pos.x[c];
        sy[3 + c] = vt[c];
        st[c] = s.pos.x[c];
        st[3 + c] = s.v[c];
    }
    st[6] = t;
    st[7] = dt;
    a.iters[2 * i] = it;
    a.iters[2 * i + 1] = steps;
    a.flags[i] = status;
    p.out_len[i] = (int)nrec;
}

}  // namespace

namespace tpamd {

void launch_fo_segment(const KParams& p, const FoArgs& a, int fk, hipStream_t st)
{
    const dim3 block(256), grid((unsigned)((p.N + 255) / 256));
    if (fk == kFkMesh) hipLaunchKernelGGL((k_fo_segment<S64M>), grid, block, 0, st, p, a);
    else if (fk == kFkGrid) hipLaunchKernelGGL((k_fo_segment<S64>), grid, block, 0, st, p, a);
    else hipLaunchKernelGGL((k_fo_segment<S64A>), grid, block, 0, st, p, a);
}

}  // namespace tpamd

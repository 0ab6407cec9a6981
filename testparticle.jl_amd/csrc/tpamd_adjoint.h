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
#pragma once

#include "tpamd_grid.h"

namespace tpamd {

struct AdjParams {
    const double* targets;    // [N][3]
    double* ckpt;             // [nseg][9][N]: frame position (3), wrap (3), velocity (3)
    void* seg;                // [K][9 or 6][N] recompute buffer (float or double)
    void* gradB;              // node gradient [k][j][i][3]: float (FP32C) or double (FP64)
    double* grad_xv0;         // [N][6]
    double* loss;             // scalar sum
    double* xfinal;           // [N][3] final positions
    int K;                    // checkpoint interval
    int nseg;
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

// scatter Bbar to the corner nodes and return (dB/dx)^T Bbar
template <typename R>
__device__ __forceinline__ void adj_scatter(const AdjParams& a, const DevField& f, const R c[8][3],
                                            const unsigned node[8], const R w[3], const R Bb[3],
                                            R xb[3])
{
    R g[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const R ax = (q & 1) ? w[0] : (R)1 - w[0];
        const R ay = ((q >> 1) & 1) ? w[1] : (R)1 - w[1];
        const R az = (q >> 2) ? w[2] : (R)1 - w[2];
        const R W = ax * ay * az;
        R* dst = (R*)a.gradB + 3u * node[q];
        unsafeAtomicAdd(dst, W * Bb[0]);
        unsafeAtomicAdd(dst + 1, W * Bb[1]);
        unsafeAtomicAdd(dst + 2, W * Bb[2]);
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

}  // namespace tpamd

// tpamd_grid.h -- compile-time specialised policies for gridded magnetic
// fields: B on a uniform grid, E uniform (or zero) or E = -u x B.
//
// Everything the generic policies decide per call (field kinds, storage
// type, wrap mode, multistep subcycle count and polynomial degree) is a
// template argument here, and node addresses are 32-bit byte offsets from
// the grid base.  The FP64 policy keeps the generic path's operation order
// exactly, so its trajectories stay bit-identical to the CPU restatement; the
// FP32C policy uses reciprocals and hand-placed fmaf.
#pragma once

#include "tpamd_device.h"

namespace tpamd {

enum { kEUniform = 0, kEUxB = 1 };

// ---------------------------------------------------------------------------
// Cell location: index pair and weight along one axis
// ---------------------------------------------------------------------------
template <bool PER>
__device__ __forceinline__ bool locate_f64(const DevField& f, int d, double s, int& i0, int& i1,
                                           double& w)
{
    const int nd = f.n[d];
    const double smax = (double)(nd - 1);
    if (s != s) return false;
    if (PER) {
        if (!(fabs(s) < 4.0e18)) return false;
        const double fl = floor(s);
        w = s - fl;
        int ii;
        if (fabs(fl) < 1.0e9) {
            ii = (int)fl % nd;
        } else {
            long long q = (long long)fl % (long long)nd;
            ii = (int)q;
        }
        if (ii < 0) ii += nd;
        i0 = ii;
        i1 = (ii + 1 == nd) ? 0 : ii + 1;
        return true;
    }
    if (f.extrap == kExtrapFillNaN) {
        if (!(s >= 0.0 && s <= smax)) return false;
    } else {
        s = s < 0.0 ? 0.0 : s;
        s = s > smax ? smax : s;
    }
    double fl = floor(s);
    if (fl > smax - 1.0) fl = smax - 1.0;
    w = s - fl;
    i0 = (int)fl;
    i1 = i0 + 1;
    return true;
}

template <bool PER>
__device__ __forceinline__ bool locate_f32(const DevField& f, int d, float s, int& i0, int& i1,
                                           float& w)
{
    const int nd = f.n[d];
    if (PER) {
        if (!(fabsf(s) < 1.0e9f)) return false;
        const float fl = floorf(s);
        w = s - fl;
        int ii = (int)fl;
        // positions live in [0, period) of the frame, so s is in [0, n]
        if ((unsigned)ii >= (unsigned)nd) {
            ii %= nd;
            if (ii < 0) ii += nd;
        }
        i0 = ii;
        i1 = (ii + 1 == nd) ? 0 : ii + 1;
        return true;
    }
    const float smax = (float)(nd - 1);
    if (f.extrap == kExtrapFillNaN) {
        if (!(s >= 0.0f && s <= smax)) return false;
    } else {
        s = s != s ? 0.0f : s;
        s = fminf(fmaxf(s, 0.0f), smax);
    }
    const float fl = fminf(floorf(s), smax - 1.0f);
    w = s - fl;
    i0 = (int)fl;
    i1 = i0 + 1;
    return true;
}

template <typename T>
__device__ __forceinline__ auto at(const void* base, unsigned off)
{
    return gbl((const T*)((const char*)base + off));
}

// Trilinear gather over byte offsets; OUT is float or double, lerp chosen to
// match the generic policies (lerp_f64 / lerp_f32).
template <typename T, typename OUT>
__device__ __forceinline__ void trilinear(const void* base, const unsigned ox[2], const unsigned oy[2],
                                          const unsigned oz[2], const OUT w[3], OUT out[3])
{
    OUT c[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const unsigned oyz = oy[q & 1] + oz[q >> 1];
        const auto a = at<T>(base, ox[0] + oyz);
        const auto b = at<T>(base, ox[1] + oyz);
        OUT av[3] = {(OUT)a[0], (OUT)a[1], (OUT)a[2]};
        OUT bv[3] = {(OUT)b[0], (OUT)b[1], (OUT)b[2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (sizeof(OUT) == 8) c[q][k] = lerp_f64(av[k], bv[k], w[0]);
            else c[q][k] = lerp_f32(av[k], bv[k], w[0]);
        }
    }
    // c[0] = (y0, z0), c[1] = (y1, z0), c[2] = (y0, z1), c[3] = (y1, z1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if constexpr (sizeof(OUT) == 8)
            out[k] = lerp_f64(lerp_f64(c[0][k], c[1][k], w[1]), lerp_f64(c[2][k], c[3][k], w[1]), w[2]);
        else
            out[k] = lerp_f32(lerp_f32(c[0][k], c[1][k], w[1]), lerp_f32(c[2][k], c[3][k], w[1]), w[2]);
    }
}

// Trilinear blend from one cell-replicated block (24 values, corner
// (dx, dy, dz) at 3 * (dx + 2 dy + 4 dz)); same lerp order as trilinear().
template <typename T, typename OUT>
__device__ __forceinline__ void trilinear_cell(const void* base, unsigned off, const OUT w[3],
                                               OUT out[3])
{
    T v[24];
    if constexpr (sizeof(T) == 4) {
        const auto q = gbl((const float4*)((const char*)base + off));
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const float4 a = q[j];
            v[4 * j] = a.x; v[4 * j + 1] = a.y; v[4 * j + 2] = a.z; v[4 * j + 3] = a.w;
        }
    } else {
        const auto q = gbl((const double2*)((const char*)base + off));
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            const double2 a = q[j];
            v[2 * j] = a.x; v[2 * j + 1] = a.y;
        }
    }
    OUT c[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const OUT a = (OUT)v[6 * q + k], b = (OUT)v[6 * q + 3 + k];
            if constexpr (sizeof(OUT) == 8) c[q][k] = lerp_f64(a, b, w[0]);
            else c[q][k] = lerp_f32(a, b, w[0]);
        }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if constexpr (sizeof(OUT) == 8)
            out[k] = lerp_f64(lerp_f64(c[0][k], c[1][k], w[1]), lerp_f64(c[2][k], c[3][k], w[1]), w[2]);
        else
            out[k] = lerp_f32(lerp_f32(c[0][k], c[1][k], w[1]), lerp_f32(c[2][k], c[3][k], w[1]), w[2]);
    }
}

__device__ __forceinline__ unsigned cell_offset(const DevField& f, const int i0[3])
{
    return f.cell_stride *
           ((unsigned)i0[0] + (unsigned)f.nc[0] * ((unsigned)i0[1] + (unsigned)f.nc[1] * (unsigned)i0[2]));
}

// builds the cell-replicated copy from the node array (one thread per cell)
template <typename T>
__global__ void k_build_cells(const T* __restrict__ nodes, int nx, int ny, int nz, int periodic,
                              int ncx, int ncy, int ncz, unsigned stride_elems, T* __restrict__ cells)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)ncx * ncy * ncz;
    if (c >= total) return;
    const int i = (int)(c % ncx);
    const int j = (int)((c / ncx) % ncy);
    const int k = (int)(c / ((long long)ncx * ncy));
    T* out = cells + c * stride_elems;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        int a = i + (q & 1), b = j + ((q >> 1) & 1), d = k + (q >> 2);
        if (periodic) {
            a = a == nx ? 0 : a;
            b = b == ny ? 0 : b;
            d = d == nz ? 0 : d;
        }
        const T* src = nodes + 3 * ((long long)a + (long long)nx * ((long long)b + (long long)ny * d));
        out[3 * q] = src[0];
        out[3 * q + 1] = src[1];
        out[3 * q + 2] = src[2];
    }
    for (unsigned e = 24; e < stride_elems; ++e) out[e] = (T)0;
}

__device__ __forceinline__ void offsets(const DevField& f, const int i0[3], const int i1[3],
                                        unsigned es, unsigned ox[2], unsigned oy[2], unsigned oz[2])
{
    const unsigned sx = 3u * es, sy = sx * (unsigned)f.n[0], sz = sy * (unsigned)f.n[1];
    ox[0] = sx * (unsigned)i0[0]; ox[1] = sx * (unsigned)i1[0];
    oy[0] = sy * (unsigned)i0[1]; oy[1] = sy * (unsigned)i1[1];
    oz[0] = sz * (unsigned)i0[2]; oz[1] = sz * (unsigned)i1[2];
}

// ---------------------------------------------------------------------------
// FP64 multistep with the polynomial degree fixed at compile time (same
// operation order as multistep_f64's rational branch)
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ double horner_fixed(const double* p, double u)
{
    double r = p[NS];
#pragma unroll
    for (int j = NS - 1; j >= 0; --j) r = fma(r, u, p[j]);
    return r;
}

// Subcycle classes of the specialised multistep kernels: NS = n for the
// rational coefficients, NS = kRefBase + n for the reference form
// (multistep_boris.jl:94-120: atan, sincos and the Taylor branch).
constexpr int kRefBase = 100;
__host__ __device__ constexpr int ns_sub(int NS) { return NS > kRefBase ? NS - kRefBase : NS; }

// reference-form coefficients c1, c2, c3, c6 (multistep_boris.jl:94-120),
// with the FDLIBM atan and sin/cos the oracle uses
__device__ __forceinline__ void multistep_ref_coeffs_f64(double tau2, int n, double& c1, double& c2,
                                                         double& c3, double& c6)
{
    const double tmag = sqrt(tau2);
    const double nn = (double)n;
    if (tmag < 1.0e-4) {
        c1 = fma(-(2.0 * nn * nn), tau2, 1.0);
        const double n1 = 2.0 * nn, n3 = 4.0 * nn * nn * nn;
        c2 = fma(-((n1 + n3) / 3.0), tau2, n1);
        c3 = fma(-((4.0 * nn * nn + 2.0 * nn * nn * nn * nn) / 3.0), tau2, 2.0 * nn * nn);
        c6 = (n1 + n3) / 3.0;
    } else {
        const double a = tpf_atan(tmag);
        double s, co;
        tpf_sincos(nn * a, &s, &co);
        const double s2 = 2.0 * s * co;
        c1 = fma(co, co, -(s * s));
        c2 = s2 / tmag;
        c3 = 2.0 * s * s / tau2;
        c6 = (2.0 * nn - c2) / tau2;
    }
}

template <int NS>
__device__ __forceinline__ void multistep_f64_n(const KParams& p, const double v[3], const double E[3],
                                                const double B[3], double dt, double out[3])
{
    constexpr int NSUB = ns_sub(NS);
    const double factor = p.q2m * dt / (2.0 * NSUB);
    double tn[3], en[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) tn[c] = factor * B[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) en[c] = factor * E[c];
    if (p.order_N != 2) {
        const double tm2 = dot3(tn, tn);
        double fN, ecf;
        if (p.order_N == 4) {
            fN = 1.0 + tm2 / 3.0;
            ecf = -1.0 / 3.0;
        } else {
            fN = (1.0 + tm2 / 3.0) + 2.0 * tm2 * tm2 / 15.0;
            ecf = -1.0 / 3.0 - 2.0 * tm2 / 15.0;
        }
        const double edt = dot3(en, tn);
#pragma unroll
        for (int c = 0; c < 3; ++c) en[c] = fN * en[c] + (ecf * edt) * tn[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) tn[c] = fN * tn[c];
    }
    const double tau2 = dot3(tn, tn);
    double c1, c2, c3, c6;
    if constexpr (NS > kRefBase) {
        multistep_ref_coeffs_f64(tau2, NSUB, c1, c2, c3, c6);
    } else {
        const double rD = 1.0 / horner_fixed<NSUB>(p.poly.D, tau2);
        c1 = horner_fixed<NSUB>(p.poly.P1, tau2) * rD;
        c2 = horner_fixed<NSUB - 1>(p.poly.P2, tau2) * rD;
        c3 = horner_fixed<NSUB - 1>(p.poly.P3, tau2) * rD;
        c6 = horner_fixed<NSUB - 1>(p.poly.P6, tau2) * rD;
    }
    const double vdt = dot3(v, tn), edt = dot3(en, tn);
    double vxt[3], ext[3];
    cross3(v, tn, vxt);
    cross3(en, tn, ext);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double r = c1 * v[c];
        r = r + c2 * vxt[c];
        r = r + (c3 * vdt) * tn[c];
        r = r + c2 * en[c];
        r = r + c3 * ext[c];
        r = r + (c6 * edt) * tn[c];
        out[c] = r;
    }
}

// ---------------------------------------------------------------------------
// FP32 multistep with compile-time degree; coefficients and q/m factors come
// pre-rounded in KParams::f32
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ float horner_fixed_f(const float* p, float u)
{
    float r = p[NS];
#pragma unroll
    for (int j = NS - 1; j >= 0; --j) r = fmaf(r, u, p[j]);
    return r;
}

template <int NS>
__device__ __forceinline__ void multistep_f32_n(const KParams& p, const float v[3], const float E[3],
                                                const float B[3], float factor, float out[3])
{
    float tn[3], en[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) tn[c] = factor * B[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) en[c] = factor * E[c];
    if (p.order_N != 2) {
        const float tm2 = dot3f(tn, tn);
        float fN, ecf;
        if (p.order_N == 4) {
            fN = fmaf(tm2, 1.0f / 3.0f, 1.0f);
            ecf = -1.0f / 3.0f;
        } else {
            fN = fmaf(tm2, fmaf(tm2, 2.0f / 15.0f, 1.0f / 3.0f), 1.0f);
            ecf = fmaf(tm2, -2.0f / 15.0f, -1.0f / 3.0f);
        }
        const float k = ecf * dot3f(en, tn);
#pragma unroll
        for (int c = 0; c < 3; ++c) en[c] = fmaf(fN, en[c], k * tn[c]);
#pragma unroll
        for (int c = 0; c < 3; ++c) tn[c] = fN * tn[c];
    }
    const float tau2 = dot3f(tn, tn);
    float c1, c2, c3, c6;
    if constexpr (NS > kRefBase) {
        // reference form in float (multistep_boris.jl:94-120)
        constexpr float nn = (float)ns_sub(NS);
        const float tmag = sqrtf(tau2);
        if (tmag < 1.0e-4f) {
            c1 = fmaf(-(2.0f * nn * nn), tau2, 1.0f);
            c2 = fmaf(-((2.0f * nn + 4.0f * nn * nn * nn) / 3.0f), tau2, 2.0f * nn);
            c3 = fmaf(-((4.0f * nn * nn + 2.0f * nn * nn * nn * nn) / 3.0f), tau2, 2.0f * nn * nn);
            c6 = (2.0f * nn + 4.0f * nn * nn * nn) / 3.0f;
        } else {
            float s, co;
            sincosf(nn * atanf(tmag), &s, &co);
            c1 = fmaf(co, co, -(s * s));
            c2 = 2.0f * s * co / tmag;
            c3 = 2.0f * s * s / tau2;
            c6 = (2.0f * nn - c2) / tau2;
        }
    } else {
        const float rD = __builtin_amdgcn_rcpf(horner_fixed_f<NS>(p.f32.D, tau2));
        c1 = horner_fixed_f<NS>(p.f32.P1, tau2) * rD;
        c2 = horner_fixed_f<NS - 1>(p.f32.P2, tau2) * rD;
        c3 = horner_fixed_f<NS - 1>(p.f32.P3, tau2) * rD;
        c6 = horner_fixed_f<NS - 1>(p.f32.P6, tau2) * rD;
    }
    const float vdt = dot3f(v, tn), edt = dot3f(en, tn);
    float vxt[3], ext[3];
    cross3f(v, tn, vxt);
    cross3f(en, tn, ext);
    const float a3 = c3 * vdt, a6 = c6 * edt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float r = c1 * v[c];
        r = fmaf(c2, vxt[c], r);
        r = fmaf(a3, tn[c], r);
        r = fmaf(c2, en[c], r);
        r = fmaf(c3, ext[c], r);
        out[c] = fmaf(a6, tn[c], r);
    }
}

__device__ __forceinline__ void boris_f32_fast(const float v[3], const float E[3], const float B[3],
                                               float h, float out[3])
{
    float t[3], s[3], vm[3], vmt[3], vp[3], vps[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = h * B[c];
    const float r = 2.0f * __builtin_amdgcn_rcpf(1.0f + dot3f(t, t));
#pragma unroll
    for (int c = 0; c < 3; ++c) s[c] = t[c] * r;
#pragma unroll
    for (int c = 0; c < 3; ++c) vm[c] = fmaf(h, E[c], v[c]);
    cross3f(vm, t, vmt);
#pragma unroll
    for (int c = 0; c < 3; ++c) vp[c] = vm[c] + vmt[c];
    cross3f(vp, s, vps);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = fmaf(h, E[c], vm[c] + vps[c]);
}

}  // namespace tpamd

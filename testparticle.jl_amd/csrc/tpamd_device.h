// tpamd_device.h -- device-side math of the ensemble Boris pusher (gfx950).
//
// Two arithmetic policies share one driver:
//
//  * FP64 ("exact"): state, fields and arithmetic in double with the
//    reference's operation order and no contraction (the library is built
//    with -ffp-contract=off); fma() appears only where the reference's
//    @muladd fuses a scalar a*b+c.  Results are rounding-identical to the CPU
//    restatement of src/boris/boris.jl et al., so trajectories agree bit for
//    bit even in chaotic fields.
//
//  * FP32C ("compensated"): gridded fields stored and gathered in float, the
//    rotation in float with hand-placed fmaf, the position as an unevaluated
//    float pair hi + lo updated with TwoSum, periodic wrap counts as int, and
//    the time in double.  Positions are kept relative to the grid origin so
//    the cell lookup never loses the fraction of a cell.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tpamd_fdlibm.h"

namespace tpamd {

constexpr int kRationalNMax = 12;

// Every buffer the kernels read or write is device (global) memory.  A pointer
// the batch kernel takes from its per-instance parameter table is loaded from
// memory, so the compiler cannot tell its address space and emits flat
// accesses; flat accesses count against the scalar loads' wait counter too, so
// each wait for a parameter would also wait for every gather and record store
// in flight.  Accesses through gbl() compile to global loads and stores (the
// host pass, which only checks these bodies, keeps the plain pointer).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gbl(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}
#else
template <class T>
__device__ __forceinline__ T* gbl(T* p)
{
    return p;
}
#endif
// adaptive solves without a step cap stop after this many steps per particle
constexpr long long kAdaptiveHardCap = 1LL << 31;

struct MeshGrid;

struct DevField {
    int kind;        // tpamd_field_kind
    int extrap;      // tpamd_extrap (grids)
    int dtype;       // storage dtype of `data` (grids)
    int pad0;
    int n[3];
    int pad1;
    double v[3];     // uniform value / affine f0 / UxB drift velocity
    double jac[9];   // affine Jacobian, row-major
    double ft[3];    // affine time derivative
    double b0;       // dipole
    double radius;
    double origin[3];
    double inv_h[3]; // 1 / spacing
    double period[3];// n * spacing (periodic grids)
    float inv_hf[3];
    float pad2;
    const void* data;
    // optional cell-replicated copy: per cell the 8 corner nodes x 3
    // components, corner (dx, dy, dz) at 3 * (dx + 2 dy + 4 dz), one cell per
    // `cell_stride` bytes (float: 128 B, one cache line; double: 192 B);
    // cells per axis nc[d] = n[d] (periodic) or n[d] - 1
    const void* cells;
    int nc[3];
    unsigned cell_stride;
    // time-dependent grids (kind TimeGrid): nsnap snapshots of the node
    // array, snap_bytes apart from `data`, at sorted times[nsnap]
    const double* times;
    long long snap_bytes;
    int nsnap;
    int pad3;
    // kind Mesh: rectilinear / cubic / spherical grid descriptor (device)
    const MeshGrid* mesh;
};

struct Poly {
    double D[kRationalNMax + 1];
    double P1[kRationalNMax + 1];
    double P2[kRationalNMax];
    double P3[kRationalNMax];
    double P6[kRationalNMax];
    int deg;
    int pad;
};

// float copies of per-solve constants for the FP32C specialised policies
struct F32Consts {
    float D[kRationalNMax + 1];
    float P1[kRationalNMax + 1];
    float P2[kRationalNMax];
    float P3[kRationalNMax];
    float P6[kRationalNMax];
    float u[3];    // UxB drift velocity
    float e0[3];   // uniform E
    // float |x|^2 beyond which the sphere tests are decided without the
    // double position: above rin2_skip a particle is not inside r_in, below
    // rout2_skip it is not outside r_out
    float rin2_skip, rout2_skip;
};

struct KParams {
    DevField E, B;
    Poly poly;
    F32Consts f32;
    double q2m_2n;            // q/m / (2 n), multistep
    int alg, order_N, nsub, form;
    int save_start, save_end, save_every, bnd_flags;
    int e_is_uxb, frame_periodic;
    int static_fields;        // neither E nor B depends on t
    int uniform_fields;       // E and B are the same at every x and t (zero / uniform, or
                              // E = -u x B of a uniform B)
    int rec_w;                // record columns: 6, + 6 (E, B), + 4 (work rates)
    int save_fields, save_work;
    int span;                 // adaptive solves: particles per wave (64 x particles per lane)
    double mass;              // for the work rates
    double q2m, t0, t1, dt, C;
    long long nt, nout, cap, K, max_steps, N;
    double r_in, r_out;
    // |x|^2 bounds around each sphere: outside [lo, hi] the radius test is
    // decided without the square root (same outcome as sqrt(|x|^2) vs r)
    double rin2_lo, rin2_hi, rout2_lo, rout2_hi;
    double box_lo[3], box_hi[3];
    double frame_origin[3];   // FP32C: positions stored relative to this point
    double frame_period[3];   // FP32C: wrap period (0 = no wrap)
    const double* xv0;
    const double* dt_pp;
    double* out_u;
    double* out_t;
    int* out_len;
    int* retcode;
    unsigned long long* steps;
};

enum { kFieldZero = 0, kFieldUniform = 1, kFieldAffine = 2, kFieldDipole = 3, kFieldGrid = 4,
       kFieldUxB = 5, kFieldTimeGrid = 6, kFieldMesh = 7 };
enum { kExtrapFillNaN = 0, kExtrapClamp = 1, kExtrapPeriodic = 2 };
}  // namespace tpamd
#include "tpamd_mesh.h"
namespace tpamd {
enum { kBndNaN = 1, kBndSphereIn = 2, kBndSphereOut = 4, kBndBox = 8 };
enum { kRetSuccess = 0, kRetTerminated = 1, kRetMaxIters = 2 };

// ---------------------------------------------------------------------------
// FP64 helpers: StaticArrays order (cross, left-folded dot)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double a[3], const double b[3])
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// boris_velocity_update, src/boris/boris.jl:19-31 (three true divisions, as
// the SVector `2 * t / (1 + t^2)` divides componentwise)
__device__ __forceinline__ void boris_f64(const double v[3], const double E[3], const double B[3],
                                          double h, double out[3])
{
    double t[3], s[3], vm[3], vmt[3], vp[3], vps[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = h * B[c];
    const double den = 1.0 + dot3(t, t);
#pragma unroll
    for (int c = 0; c < 3; ++c) s[c] = (2.0 * t[c]) / den;
#pragma unroll
    for (int c = 0; c < 3; ++c) vm[c] = v[c] + h * E[c];
    cross3(vm, t, vmt);
#pragma unroll
    for (int c = 0; c < 3; ++c) vp[c] = vm[c] + vmt[c];
    cross3(vp, s, vps);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = (vm[c] + vps[c]) + h * E[c];
}

// x / d for several x over one d, by the instruction sequence of the
// compiler's IEEE double division (scale, reciprocal, two Newton steps on
// it, quotient, residual, fused correction, fix-up) with the denominator's
// half formed once.  The two scale steps return their operand unchanged
// when d is in [1, 2^64] and |x| in [2^-512, 2^64] (no overflow or
// underflow in any step), so there the quotient is the division's own, bit
// for bit; callers divide plainly outside that range (zero x included).
struct SharedDiv {
    double d, r;
    __device__ __forceinline__ void init(double den)
    {
        d = den;
        const double r0 = __builtin_amdgcn_rcp(den);
        const double e0 = fma(-den, r0, 1.0);
        const double r1 = fma(r0, e0, r0);
        const double e1 = fma(-den, r1, 1.0);
        r = fma(r1, e1, r1);
    }
    __device__ __forceinline__ double div(double x) const
    {
        const double m = x * r;
        const double f = fma(-d, m, x);
        return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(f, r, m, false), d, x);
    }
};

// boris_f64 with the three divisions of s = 2t / (1 + t.t) sharing their
// denominator (SharedDiv); the same result
__device__ __forceinline__ void boris_f64_sd(const double v[3], const double E[3], const double B[3],
                                             double h, double out[3])
{
    double t[3], s[3], vm[3], vmt[3], vp[3], vps[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = h * B[c];
    const double den = 1.0 + dot3(t, t);
    const double n[3] = {2.0 * t[0], 2.0 * t[1], 2.0 * t[2]};
    // den = 1 + t.t >= 1 (a NaN fails the test below); |2 t_c| <= 2 sqrt(den),
    // so den <= 2^64 bounds the numerators by 2^33.  & keeps the test branch-free.
    if ((den <= 0x1p64) & (fabs(n[0]) >= 0x1p-512) & (fabs(n[1]) >= 0x1p-512) & (fabs(n[2]) >= 0x1p-512)) {
        SharedDiv q;
        q.init(den);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] = q.div(n[c]);
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] = n[c] / den;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) vm[c] = v[c] + h * E[c];
    cross3(vm, t, vmt);
#pragma unroll
    for (int c = 0; c < 3; ++c) vp[c] = vm[c] + vmt[c];
    cross3(vp, s, vps);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = (vm[c] + vps[c]) + h * E[c];
}

// boris_f64 for fields that are the same at every step: t, s and h E depend
// only on the fields and h, so they are formed once (by the same operations,
// so the update is bit-identical) and each step does the rest.
struct BorisConst64 {
    double t[3], s[3], hE[3];
    __device__ __forceinline__ void init(const double E[3], const double B[3], double h)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = h * B[c];
        const double den = 1.0 + dot3(t, t);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] = (2.0 * t[c]) / den;
#pragma unroll
        for (int c = 0; c < 3; ++c) hE[c] = h * E[c];
    }
    __device__ __forceinline__ void apply(const double v[3], double out[3]) const
    {
        double vm[3], vmt[3], vp[3], vps[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) vm[c] = v[c] + hE[c];
        cross3(vm, t, vmt);
#pragma unroll
        for (int c = 0; c < 3; ++c) vp[c] = vm[c] + vmt[c];
        cross3(vp, s, vps);
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = (vm[c] + vps[c]) + hE[c];
    }
};

// Horner over p[0..d], d <= NMAX: r = p[d], then r = fma(r, u, p[j]) for
// j = d-1 .. 0.  The loop runs NMAX+1 unrolled iterations with static
// indices (a runtime index into the kernel-argument block would make the
// compiler copy the block to scratch memory).
template <int NMAX>
__device__ __forceinline__ double horner_f64(const double* p, int d, double u)
{
    double r = 0.0;
#pragma unroll
    for (int j = NMAX; j >= 0; --j) {
        if (j == d) r = p[j];
        else if (j < d) r = fma(r, u, p[j]);
    }
    return r;
}

// update_velocity_multistep, src/boris/multistep_boris.jl:65-142
__device__ __forceinline__ void multistep_f64(const KParams& p, const double v[3], const double E[3],
                                              const double B[3], double dt, double out[3])
{
    const int n = p.nsub;
    const double factor = p.q2m * dt / (2.0 * n);
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
    if (p.form == 1) {
        const double rD = 1.0 / horner_f64<kRationalNMax>(p.poly.D, p.poly.deg, tau2);
        c1 = horner_f64<kRationalNMax>(p.poly.P1, p.poly.deg, tau2) * rD;
        c2 = horner_f64<kRationalNMax - 1>(p.poly.P2, p.poly.deg - 1, tau2) * rD;
        c3 = horner_f64<kRationalNMax - 1>(p.poly.P3, p.poly.deg - 1, tau2) * rD;
        c6 = horner_f64<kRationalNMax - 1>(p.poly.P6, p.poly.deg - 1, tau2) * rD;
    } else {
        const double tmag = sqrt(tau2);
        const double nn = (double)n;
        if (tmag < 1.0e-4) {
            c1 = fma(-(2.0 * nn * nn), tau2, 1.0);
            const double n1 = 2.0 * nn, n3 = 4.0 * nn * nn * nn;
            c2 = fma(-((n1 + n3) / 3.0), tau2, n1);
            c3 = fma(-((4.0 * nn * nn + 2.0 * nn * nn * nn * nn) / 3.0), tau2, 2.0 * nn * nn);
            c6 = (n1 + n3) / 3.0;
        } else {
            // FDLIBM atan and sin/cos (tpamd_fdlibm.h), rounding as the oracle does
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
// FP64 fields
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lerp_f64(double a, double b, double w)
{
    return fma(w, b, (1.0 - w) * a);
}

__device__ __forceinline__ void node3_f64(const DevField& f, const void* data, long long idx,
                                          double o[3])
{
    if (f.dtype == 1) {
        const float* d = (const float*)data + 3 * idx;
        o[0] = (double)d[0]; o[1] = (double)d[1]; o[2] = (double)d[2];
    } else {
        const double* d = (const double*)data + 3 * idx;
        o[0] = d[0]; o[1] = d[1]; o[2] = d[2];
    }
}
__device__ __forceinline__ void node3_f64(const DevField& f, long long idx, double o[3])
{
    node3_f64(f, f.data, idx, o);
}

// build_interpolator(CartesianGrid, A::4D, ...) + linear_interp,
// src/utility/interpolation.jl:119-148, :203-206, :217-235
__device__ __forceinline__ void grid_f64(const DevField& f, const double x[3], double out[3],
                                         const void* data = nullptr)
{
    const void* base = data ? data : f.data;
    int i0[3], i1[3];
    double w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int nd = f.n[d];
        double s = (x[d] - f.origin[d]) * f.inv_h[d];
        const double smax = (double)(nd - 1);
        if (s != s || (f.extrap == kExtrapPeriodic && !(fabs(s) < 4.0e18))) {
            out[0] = out[1] = out[2] = __builtin_nan("");
            return;
        }
        if (f.extrap == kExtrapPeriodic) {
            const double fl = floor(s);
            w[d] = s - fl;
            long long ii = (long long)fl % (long long)nd;
            if (ii < 0) ii += nd;
            i0[d] = (int)ii;
            i1[d] = (i0[d] + 1 == nd) ? 0 : i0[d] + 1;
        } else {
            if (f.extrap == kExtrapFillNaN) {
                if (!(s >= 0.0 && s <= smax)) {
                    out[0] = out[1] = out[2] = __builtin_nan("");
                    return;
                }
            } else {
                s = s < 0.0 ? 0.0 : s;
                s = s > smax ? smax : s;
            }
            double fl = floor(s);
            if (fl > smax - 1.0) fl = smax - 1.0;
            w[d] = s - fl;
            i0[d] = (int)fl;
            i1[d] = i0[d] + 1;
        }
    }
    const long long nx = f.n[0], nxy = (long long)f.n[0] * f.n[1];
    double a[3], b[3], c00[3], c10[3], c01[3], c11[3];
    node3_f64(f, base, i0[0] + nx * i0[1] + nxy * i0[2], a);
    node3_f64(f, base, i1[0] + nx * i0[1] + nxy * i0[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c00[c] = lerp_f64(a[c], b[c], w[0]);
    node3_f64(f, base, i0[0] + nx * i1[1] + nxy * i0[2], a);
    node3_f64(f, base, i1[0] + nx * i1[1] + nxy * i0[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c10[c] = lerp_f64(a[c], b[c], w[0]);
    node3_f64(f, base, i0[0] + nx * i0[1] + nxy * i1[2], a);
    node3_f64(f, base, i1[0] + nx * i0[1] + nxy * i1[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c01[c] = lerp_f64(a[c], b[c], w[0]);
    node3_f64(f, base, i0[0] + nx * i1[1] + nxy * i1[2], a);
    node3_f64(f, base, i1[0] + nx * i1[1] + nxy * i1[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c11[c] = lerp_f64(a[c], b[c], w[0]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[c] = lerp_f64(lerp_f64(c00[c], c10[c], w[1]), lerp_f64(c01[c], c11[c], w[1]), w[2]);
}

// Time-dependent grids, LazyTimeInterpolator (src/utility/interpolation.jl:
// 337-430): k = number of snapshot times <= t (searchsortedlast); k == 0
// clamps to the first snapshot, k == nsnap to the last, otherwise snapshots
// k-1 and k blend with w = (t - t_{k-1}) / (t_k - t_{k-1})
__device__ __forceinline__ int time_bracket(const DevField& f, double t, double& w)
{
    int lo = 0, hi = f.nsnap;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (f.times[mid] <= t) lo = mid + 1;
        else hi = mid;
    }
    w = 0.0;
    if (lo > 0 && lo < f.nsnap) w = (t - f.times[lo - 1]) / (f.times[lo] - f.times[lo - 1]);
    return lo;
}
__device__ __forceinline__ const void* snapshot(const DevField& f, int k)
{
    return (const char*)f.data + (size_t)k * (size_t)f.snap_bytes;
}

// Field-kind classes a generic kernel is compiled for: analytic fields only
// (zero, uniform, affine, dipole), + uniform grids and their time sequences,
// + meshes (through the out-of-line mesh evaluator).  A solve takes the
// smallest class that holds its E and B, so kernels never carry the code
// (and the registers) of field kinds they cannot meet.
enum { kFkAnalytic = 0, kFkGrid = 1, kFkMesh = 2 };

__host__ __device__ inline int field_class(int kind)
{
    return kind == kFieldMesh ? kFkMesh : ((kind == kFieldGrid || kind == kFieldTimeGrid) ? kFkGrid : kFkAnalytic);
}

// the dipole field with its moment along -z, B = b0 R^3 / r^5 (-3xz, -3yz,
// r^2 - 3z^2): the Earth-like dipole of docs/examples/analytic/demo_proton_dipole.jl
__device__ __forceinline__ void dipole_f64(const DevField& f, const double x[3], double o[3])
{
    const double r2 = dot3(x, x);
    const double r = sqrt(r2);
    const double r5 = (r2 * r2) * r;
    const double k = f.b0 * ((f.radius * f.radius) * f.radius);
    const double fac = k / r5;
    o[0] = fac * ((-3.0 * x[0]) * x[2]);
    o[1] = fac * ((-3.0 * x[1]) * x[2]);
    o[2] = fac * (r2 - (3.0 * x[2]) * x[2]);
}

template <int FK = kFkMesh>
__device__ __forceinline__ void field_f64(const DevField& f, const double x[3], double t, double o[3])
{
    if constexpr (FK == kFkMesh) {
        if (f.kind == kFieldMesh) {
            const MeshVal r = mesh_value(f.mesh, x[0], x[1], x[2]);
            o[0] = r.v[0]; o[1] = r.v[1]; o[2] = r.v[2];
            return;
        }
    }
    switch (f.kind) {
    case kFieldUniform:
        o[0] = f.v[0]; o[1] = f.v[1]; o[2] = f.v[2];
        return;
    case kFieldAffine:
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double r = f.v[c] + f.jac[3 * c] * x[0];
            r = r + f.jac[3 * c + 1] * x[1];
            r = r + f.jac[3 * c + 2] * x[2];
            o[c] = r + f.ft[c] * t;
        }
        return;
    case kFieldDipole:
        dipole_f64(f, x, o);
        return;
    case kFieldGrid:
        if constexpr (FK >= kFkGrid) {
            grid_f64(f, x, o);
            return;
        }
        break;
    case kFieldTimeGrid:
        if constexpr (FK >= kFkGrid) {
            double w;
            const int k = time_bracket(f, t, w);
            if (k == 0 || k >= f.nsnap) {
                grid_f64(f, x, o, snapshot(f, k == 0 ? 0 : f.nsnap - 1));
                return;
            }
            double a[3], b[3];
            grid_f64(f, x, a, snapshot(f, k - 1));
            grid_f64(f, x, b, snapshot(f, k));
#pragma unroll
            for (int c = 0; c < 3; ++c) o[c] = (1.0 - w) * a[c] + w * b[c];   // interpolation.jl:377
            return;
        }
        break;
    default:
        break;
    }
    // zero (and kinds outside this instantiation's class, which the host
    // never hands it)
    o[0] = o[1] = o[2] = 0.0;
}

// slope of the trilinear interpolant at x (snapshot `data`): J[i][j] = dB_i/dx_j;
// zero along an axis where a clamped grid is clamped
__device__ __forceinline__ void grid_jac_f64(const DevField& f, const void* data, const double x[3],
                                             double J[3][3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) J[i][0] = J[i][1] = J[i][2] = 0.0;
    // trilinear slope; zero along an axis where a clamped grid is clamped
    int i0[3], i1[3];
    double w[3];
    bool flat[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int nd = f.n[d];
        double s = (x[d] - f.origin[d]) * f.inv_h[d];
        const double smax = (double)(nd - 1);
        flat[d] = false;
        if (s != s || (f.extrap == kExtrapPeriodic && !(fabs(s) < 4.0e18))) {
            J[0][0] = __builtin_nan("");
            return;
        }
        if (f.extrap == kExtrapPeriodic) {
            const double fl = floor(s);
            w[d] = s - fl;
            long long ii = (long long)fl % (long long)nd;
            if (ii < 0) ii += nd;
            i0[d] = (int)ii;
            i1[d] = (i0[d] + 1 == nd) ? 0 : i0[d] + 1;
        } else {
            if (f.extrap == kExtrapFillNaN) {
                if (!(s >= 0.0 && s <= smax)) return;     // B is NaN already
            } else {
                flat[d] = s < 0.0 || s > smax;
                s = s < 0.0 ? 0.0 : s;
                s = s > smax ? smax : s;
            }
            double fl = floor(s);
            if (fl > smax - 1.0) fl = smax - 1.0;
            w[d] = s - fl;
            i0[d] = (int)fl;
            i1[d] = i0[d] + 1;
        }
    }
    const long long nx = f.n[0], nxy = (long long)f.n[0] * f.n[1];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double c[3];
        node3_f64(f, data, ((q & 1) ? i1[0] : i0[0]) + nx * (((q >> 1) & 1) ? i1[1] : i0[1]) +
                         nxy * ((q >> 2) ? i1[2] : i0[2]), c);
        const double ax = (q & 1) ? w[0] : 1.0 - w[0];
        const double ay = ((q >> 1) & 1) ? w[1] : 1.0 - w[1];
        const double az = (q >> 2) ? w[2] : 1.0 - w[2];
        const double sx = (q & 1) ? 1.0 : -1.0;
        const double sy = ((q >> 1) & 1) ? 1.0 : -1.0;
        const double sz = (q >> 2) ? 1.0 : -1.0;
        const double dwx = (sx * ay) * az, dwy = (ax * sy) * az, dwz = (ax * ay) * sz;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            J[i][0] = J[i][0] + c[i] * dwx;
            J[i][1] = J[i][1] + c[i] * dwy;
            J[i][2] = J[i][2] + c[i] * dwz;
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double sd = flat[d] ? 0.0 : f.inv_h[d];
#pragma unroll
        for (int i = 0; i < 3; ++i) J[i][d] = J[i][d] * sd;
    }
    return;
}

// ---------------------------------------------------------------------------
// Saved-record augmentation (src/boris/boris.jl:405-440): B, its Jacobian
// J[i][j] = dB_i/dx_j and dB/dt at (x, t), and the work rates
// (src/equations.jl:394-441, src/utility/utility.jl:379-422).  The reference
// differentiates the field callables with ForwardDiff; here each field kind
// has its derivative in closed form.  B itself is field_f64's value.
// ---------------------------------------------------------------------------
template <int FK = kFkMesh>
__device__ __forceinline__ void field_jac_f64(const DevField& f, const double x[3], double t,
                                              double B[3], double J[3][3], double dBdt[3])
{
    field_f64<FK>(f, x, t, B);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dBdt[i] = 0.0;
        J[i][0] = J[i][1] = J[i][2] = 0.0;
    }
    switch (f.kind) {
    case kFieldAffine:
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            J[i][0] = f.jac[3 * i];
            J[i][1] = f.jac[3 * i + 1];
            J[i][2] = f.jac[3 * i + 2];
            dBdt[i] = f.ft[i];
        }
        return;
    case kFieldDipole: {
        // B_i = k r^-5 g_i, g = (-3xz, -3yz, r^2 - 3z^2):
        // dB_i/dx_j = k (r^-5 dg_i/dx_j - 5 r^-7 x_j g_i)
        const double r2 = dot3(x, x);
        const double r = sqrt(r2);
        const double r5 = (r2 * r2) * r;
        const double k = f.b0 * ((f.radius * f.radius) * f.radius);
        const double a = k / r5;
        const double b = (5.0 * a) / r2;
        const double g[3] = {(-3.0 * x[0]) * x[2], (-3.0 * x[1]) * x[2], r2 - (3.0 * x[2]) * x[2]};
        const double dg[3][3] = {{-3.0 * x[2], 0.0, -3.0 * x[0]},
                                 {0.0, -3.0 * x[2], -3.0 * x[1]},
                                 {2.0 * x[0], 2.0 * x[1], -4.0 * x[2]}};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) J[i][j] = a * dg[i][j] - (b * x[j]) * g[i];
        return;
    }
    case kFieldGrid:
        if constexpr (FK >= kFkGrid) grid_jac_f64(f, f.data, x, J);
        return;
    case kFieldTimeGrid: {
        if constexpr (FK < kFkGrid) return;
        // jacobian and derivative_t of LazyTimeInterpolator (interpolation.jl:396-430)
        double w;
        const int k = time_bracket(f, t, w);
        if (k == 0 || k >= f.nsnap) {
            grid_jac_f64(f, snapshot(f, k == 0 ? 0 : f.nsnap - 1), x, J);
            return;
        }
        double J1[3][3], J2[3][3], a[3], b[3];
        grid_jac_f64(f, snapshot(f, k - 1), x, J1);
        grid_jac_f64(f, snapshot(f, k), x, J2);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) J[i][j] = (1.0 - w) * J1[i][j] + w * J2[i][j];
        grid_f64(f, x, a, snapshot(f, k - 1));
        grid_f64(f, x, b, snapshot(f, k));
        const double dtk = f.times[k] - f.times[k - 1];
#pragma unroll
        for (int i = 0; i < 3; ++i) dBdt[i] = (b[i] - a[i]) / dtk;
        return;
    }
    case kFieldMesh:
        if constexpr (FK == kFkMesh) {
            const MeshValJac r = mesh_value_jac(f.mesh, x[0], x[1], x[2]);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) J[a][b] = r.J[3 * a + b];
        }
        return;
    default:  // zero, uniform
        return;
    }
}

// get_work_rates (src/equations.jl:394-441) with get_magnetic_properties and
// _get_curvature (src/utility/utility.jl:379-422): out = (P_par, P_fermi,
// P_grad, P_betatron)
__device__ __forceinline__ void work_rates_f64(double q2m, double m, const double v[3],
                                               const double E[3], const double B[3],
                                               const double J[3][3], const double dBdt[3],
                                               double out[4])
{
    const double q = q2m * m;
    const double Bmag = sqrt(dot3(B, B));
    if (Bmag == 0.0) {
        out[0] = out[1] = out[2] = out[3] = 0.0;
        return;
    }
    double bh[3], gradB[3], Jb[3], kappa[3], vperp[3], bk[3], bg[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) bh[i] = B[i] / Bmag;
#pragma unroll
    for (int j = 0; j < 3; ++j) gradB[j] = (J[0][j] * bh[0] + J[1][j] * bh[1]) + J[2][j] * bh[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) Jb[i] = (J[i][0] * bh[0] + J[i][1] * bh[1]) + J[i][2] * bh[2];
    const double gb = dot3(gradB, bh);
#pragma unroll
    for (int i = 0; i < 3; ++i) kappa[i] = (Jb[i] + bh[i] * (-gb)) / Bmag;
    const double vpar = dot3(v, bh);
#pragma unroll
    for (int i = 0; i < 3; ++i) vperp[i] = v[i] - vpar * bh[i];
    const double wsq = dot3(vperp, vperp);
    const double mu = (m * wsq) / (2.0 * Bmag);
    cross3(bh, kappa, bk);
    cross3(bh, gradB, bg);
    out[0] = (q * vpar) * dot3(E, bh);
    out[1] = ((m * (vpar * vpar)) / Bmag) * dot3(bk, E);
    out[2] = (mu / Bmag) * dot3(bg, E);
    out[3] = mu * dot3(dBdt, bh);
}

// ---------------------------------------------------------------------------
// FP32 helpers (compensated path).  Contractions are written out with fmaf.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dot3f(const float a[3], const float b[3])
{
    return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}
__device__ __forceinline__ void cross3f(const float a[3], const float b[3], float o[3])
{
    o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

__device__ __forceinline__ void boris_f32(const float v[3], const float E[3], const float B[3],
                                          float h, float out[3])
{
    float t[3], s[3], vm[3], vmt[3], vp[3], vps[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = h * B[c];
    const float r = 2.0f / (1.0f + dot3f(t, t));
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

// boris_f32 with t and s formed once (see BorisConst64); h E stays inside
// the fused multiply-adds, as boris_f32 forms it
struct BorisConst32 {
    float t[3], s[3], E[3], h;
    __device__ __forceinline__ void init(const float Ev[3], const float B[3], float hv)
    {
        h = hv;
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = h * B[c];
        const float r = 2.0f / (1.0f + dot3f(t, t));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s[c] = t[c] * r;
            E[c] = Ev[c];
        }
    }
    __device__ __forceinline__ void apply(const float v[3], float out[3]) const
    {
        float vm[3], vmt[3], vp[3], vps[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) vm[c] = fmaf(h, E[c], v[c]);
        cross3f(vm, t, vmt);
#pragma unroll
        for (int c = 0; c < 3; ++c) vp[c] = vm[c] + vmt[c];
        cross3f(vp, s, vps);
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = fmaf(h, E[c], vm[c] + vps[c]);
    }
};

template <int NMAX>
__device__ __forceinline__ float horner_f32(const double* p, int d, float u)
{
    float r = 0.0f;
#pragma unroll
    for (int j = NMAX; j >= 0; --j) {
        if (j == d) r = (float)p[j];
        else if (j < d) r = fmaf(r, u, (float)p[j]);
    }
    return r;
}

__device__ __forceinline__ void multistep_f32(const KParams& p, const float v[3], const float E[3],
                                              const float B[3], float dt, float out[3])
{
    const int n = p.nsub;
    const float factor = (float)p.q2m * dt / (2.0f * n);
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
    if (p.form == 1) {
        const float rD = 1.0f / horner_f32<kRationalNMax>(p.poly.D, p.poly.deg, tau2);
        c1 = horner_f32<kRationalNMax>(p.poly.P1, p.poly.deg, tau2) * rD;
        c2 = horner_f32<kRationalNMax - 1>(p.poly.P2, p.poly.deg - 1, tau2) * rD;
        c3 = horner_f32<kRationalNMax - 1>(p.poly.P3, p.poly.deg - 1, tau2) * rD;
        c6 = horner_f32<kRationalNMax - 1>(p.poly.P6, p.poly.deg - 1, tau2) * rD;
    } else {
        const float tmag = sqrtf(tau2);
        const float nn = (float)n;
        if (tmag < 1.0e-4f) {
            c1 = fmaf(-(2.0f * nn * nn), tau2, 1.0f);
            const float n1 = 2.0f * nn, n3 = 4.0f * nn * nn * nn;
            c2 = fmaf(-((n1 + n3) / 3.0f), tau2, n1);
            c3 = fmaf(-((4.0f * nn * nn + 2.0f * nn * nn * nn * nn) / 3.0f), tau2, 2.0f * nn * nn);
            c6 = (n1 + n3) / 3.0f;
        } else {
            const float a = atanf(tmag);
            float s, co;
            sincosf(nn * a, &s, &co);
            c1 = fmaf(co, co, -(s * s));
            c2 = 2.0f * s * co / tmag;
            c3 = 2.0f * s * s / tau2;
            c6 = (2.0f * nn - c2) / tau2;
        }
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

// Compensated position: x = frame_origin + wrap * period + hi + lo
struct Pos32 {
    float hi[3];
    float lo[3];
    int wrap[3];
};

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

__device__ __forceinline__ float lerp_f32(float a, float b, float w)
{
    return fmaf(w, b, (1.0f - w) * a);
}

template <typename T>
__device__ __forceinline__ void node3_f32(const T* __restrict__ d, long long idx, float o[3])
{
    const T* q = d + 3 * idx;
    o[0] = (float)q[0]; o[1] = (float)q[1]; o[2] = (float)q[2];
}

// Trilinear gather in float from frame-relative coordinates (frame origin =
// the grid origin, so s = (hi + lo) / h without cancellation).
template <typename T>
__device__ __forceinline__ void grid_f32_t(const DevField& f, const Pos32& x, float out[3],
                                           const void* data = nullptr)
{
    int i0[3], i1[3];
    float w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int nd = f.n[d];
        float s = fmaf(x.hi[d], f.inv_hf[d], x.lo[d] * f.inv_hf[d]);
        const float smax = (float)(nd - 1);
        if (f.extrap == kExtrapPeriodic) {
            if (!(fabsf(s) < 1.0e9f)) {
                out[0] = out[1] = out[2] = __builtin_nanf("");
                return;
            }
            const float fl = floorf(s);
            w[d] = s - fl;
            int ii = (int)fl % nd;
            if (ii < 0) ii += nd;
            i0[d] = ii;
            i1[d] = (ii + 1 == nd) ? 0 : ii + 1;
        } else {
            if (f.extrap == kExtrapFillNaN) {
                if (!(s >= 0.0f && s <= smax)) {
                    out[0] = out[1] = out[2] = __builtin_nanf("");
                    return;
                }
            } else {
                s = s != s ? 0.0f : s;
                s = fminf(fmaxf(s, 0.0f), smax);
            }
            const float fl = fminf(floorf(s), smax - 1.0f);
            w[d] = s - fl;
            i0[d] = (int)fl;
            i1[d] = i0[d] + 1;
        }
    }
    const T* __restrict__ d = (const T*)(data ? data : f.data);
    const long long nx = f.n[0], nxy = (long long)f.n[0] * f.n[1];
    float a[3], b[3], c00[3], c10[3], c01[3], c11[3];
    node3_f32(d, i0[0] + nx * i0[1] + nxy * i0[2], a);
    node3_f32(d, i1[0] + nx * i0[1] + nxy * i0[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c00[c] = lerp_f32(a[c], b[c], w[0]);
    node3_f32(d, i0[0] + nx * i1[1] + nxy * i0[2], a);
    node3_f32(d, i1[0] + nx * i1[1] + nxy * i0[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c10[c] = lerp_f32(a[c], b[c], w[0]);
    node3_f32(d, i0[0] + nx * i0[1] + nxy * i1[2], a);
    node3_f32(d, i1[0] + nx * i0[1] + nxy * i1[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c01[c] = lerp_f32(a[c], b[c], w[0]);
    node3_f32(d, i0[0] + nx * i1[1] + nxy * i1[2], a);
    node3_f32(d, i1[0] + nx * i1[1] + nxy * i1[2], b);
#pragma unroll
    for (int c = 0; c < 3; ++c) c11[c] = lerp_f32(a[c], b[c], w[0]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[c] = lerp_f32(lerp_f32(c00[c], c10[c], w[1]), lerp_f32(c01[c], c11[c], w[1]), w[2]);
}

// absolute position in double (for analytic fields, boundaries and output)
__device__ __forceinline__ void pos_abs(const KParams& p, const Pos32& x, double o[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d)
        o[d] = (p.frame_origin[d] + (double)x.wrap[d] * p.frame_period[d]) +
               ((double)x.hi[d] + (double)x.lo[d]);
}

// The same per field-kind class.  Solves of the analytic class have no
// frame (origin 0, no wraps: only grids set one), so the absolute position
// is the double sum hi + lo; the leading +0.0 keeps the general form's sign
// of an exact zero.
template <int FK>
__device__ __forceinline__ void pos_abs_t(const KParams& p, const Pos32& x, double o[3])
{
    if constexpr (FK == kFkAnalytic) {
#pragma unroll
        for (int d = 0; d < 3; ++d) o[d] = 0.0 + ((double)x.hi[d] + (double)x.lo[d]);
    } else {
        pos_abs(p, x, o);
    }
}

// the absolute position rounded to float, for fields evaluated in float.
// Without a frame it is the float sum hi + lo: rounding a sum of two floats
// first to double and then to float gives the correctly rounded float sum
// (double rounding is harmless for + when 53 >= 2 x 24 + 1; S. A. Figueroa,
// "When is double rounding innocuous?", SIGNUM Newsletter 30(3), 1995).
template <int FK>
__device__ __forceinline__ void pos_f32(const KParams& p, const Pos32& x, float o[3])
{
    if constexpr (FK == kFkAnalytic) {
#pragma unroll
        for (int d = 0; d < 3; ++d) o[d] = x.hi[d] + x.lo[d];
    } else {
        double xa[3];
        pos_abs(p, x, xa);
#pragma unroll
        for (int d = 0; d < 3; ++d) o[d] = (float)xa[d];
    }
}

// dipole_f64 in float at the position of a compensated state
template <int FK>
__device__ __forceinline__ void dipole_f32(const KParams& p, const DevField& f, const Pos32& x, float o[3])
{
    float xf[3];
    pos_f32<FK>(p, x, xf);
    // b0 (R/r)^3 / r^2 keeps every intermediate inside float's normal range
    const float r2 = dot3f(xf, xf);
    const float ir = rsqrtf(r2);
    const float q = (float)f.radius * ir;
    const float fac = (float)f.b0 * (q * q * q) * (ir * ir);
    o[0] = fac * (-3.0f * xf[0] * xf[2]);
    o[1] = fac * (-3.0f * xf[1] * xf[2]);
    o[2] = fac * fmaf(-3.0f * xf[2], xf[2], r2);
}

template <int FK = kFkMesh>
__device__ __forceinline__ void field_f32(const KParams& p, const DevField& f, const Pos32& x,
                                          double t, float o[3])
{
    if constexpr (FK == kFkMesh) {
        if (f.kind == kFieldMesh) {
            // meshes are evaluated in FP64 at the absolute position
            double xa[3];
            pos_abs(p, x, xa);
            const MeshVal r = mesh_value(f.mesh, xa[0], xa[1], xa[2]);
            o[0] = (float)r.v[0]; o[1] = (float)r.v[1]; o[2] = (float)r.v[2];
            return;
        }
    }
    switch (f.kind) {
    case kFieldUniform:
        o[0] = (float)f.v[0]; o[1] = (float)f.v[1]; o[2] = (float)f.v[2];
        return;
    case kFieldGrid:
        if constexpr (FK >= kFkGrid) {
            if (f.dtype == 1) grid_f32_t<float>(f, x, o);
            else grid_f32_t<double>(f, x, o);
        } else {
            o[0] = o[1] = o[2] = 0.0f;
        }
        return;
    case kFieldTimeGrid: {
        if constexpr (FK < kFkGrid) {
            o[0] = o[1] = o[2] = 0.0f;
            return;
        }
        double w;
        const int k = time_bracket(f, t, w);
        const int k0 = k == 0 ? 0 : k - 1, k1 = k >= f.nsnap ? f.nsnap - 1 : k;
        float a[3], b[3];
        if (f.dtype == 1) {
            grid_f32_t<float>(f, x, a, snapshot(f, k0));
            if (k1 != k0) grid_f32_t<float>(f, x, b, snapshot(f, k1));
        } else {
            grid_f32_t<double>(f, x, a, snapshot(f, k0));
            if (k1 != k0) grid_f32_t<double>(f, x, b, snapshot(f, k1));
        }
        if (k == 0 || k >= f.nsnap) {
            o[0] = a[0]; o[1] = a[1]; o[2] = a[2];
            return;
        }
        const float wf = (float)w;
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = fmaf(wf, b[c], (1.0f - wf) * a[c]);
        return;
    }
    case kFieldAffine: {
        float xf[3];
        pos_f32<FK>(p, x, xf);
        const float tf = (float)t;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float r = fmaf((float)f.jac[3 * c], xf[0], (float)f.v[c]);
            r = fmaf((float)f.jac[3 * c + 1], xf[1], r);
            r = fmaf((float)f.jac[3 * c + 2], xf[2], r);
            o[c] = fmaf((float)f.ft[c], tf, r);
        }
        return;
    }
    case kFieldDipole:
        dipole_f32<FK>(p, f, x, o);
        return;
    default:
        o[0] = o[1] = o[2] = 0.0f;
        return;
    }
}

}  // namespace tpamd

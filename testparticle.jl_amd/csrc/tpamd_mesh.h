// Gridded fields beyond the uniform trilinear grid: rectilinear (non-uniform)
// axes, cubic order 3, 2D/1D Cartesian fields and spherical StructuredGrid
// (src/utility/interpolation.jl:15-117,140-178,237-319).
//
// A mesh is described by a MeshGrid in device memory (the field descriptor
// holds a pointer to it); every thread of a solve reads the same descriptor,
// so its fields load through the scalar cache.
//
// Per axis, nodes c[0] < ... < c[n-1]:
//   * uniform axes locate like the uniform grid: s = (x - c[0]) / h,
//     i = floor(s), t = s - i;
//   * non-uniform axes find i = the largest index in [0, n-2] with c[i] <= x
//     (a bucket table gives the start, a short walk the exact cell) and
//     t = (x - c[i]) / (c[i+1] - c[i]);
//   * an axis with one node is ignored (2D and 1D fields);
//   * boundary: FillExtrap(NaN) gives NaN outside [c[0], c[n-1]],
//     ClampExtrap clamps the coordinate, periodic exclusive (uniform axes)
//     wraps with period n h, periodic inclusive wraps with period
//     c[n-1] - c[0] (node n-1 is node 0 one period on).
// Order 1 blends as the uniform grid does: lerp along x, then y, then z.
// Order 3 is the cardinal cubic (Catmull-Rom) on uniform axes: weights of
// nodes i-1 .. i+2 at t
//     ((-t + 2) t - 1) t / 2, ((3 t - 5) t t + 2) / 2,
//     ((-3 t + 4) t + 1) t / 2, (t - 1) t t / 2,
// with missing neighbours past a non-periodic end replaced by the linear
// extrapolation 2 f_0 - f_1 (resp. 2 f_{n-1} - f_{n-2}).
// Spherical meshes take (r, theta, phi) = cart2sph(x) (src/utility/utility.jl:
// 63-71), interpolate (B_r, B_theta, B_phi) with r and theta FillExtrap(NaN)
// and phi periodic inclusive, and rotate the result with sph2cartvec
// (utility.jl:82-89).
#pragma once

namespace tpamd {

enum { kExtrapPeriodicIncl = 3 };
enum { kMeshCartesian = 0, kMeshSpherical = 1 };

struct MeshAxis {
    const double* c;      // node coordinates (device)
    const int* bucket;    // non-uniform: first candidate cell per bucket (device)
    double lo, hi;        // c[0], c[n-1]
    double inv_h;         // uniform: 1 / spacing
    double inv_bw;        // non-uniform: buckets per unit length
    double period;        // periodic: wrap length
    int n;                // nodes (1: the axis is ignored)
    int nb;               // buckets
    int uniform;
    int bc;               // kExtrapFillNaN, kExtrapClamp, kExtrapPeriodic, kExtrapPeriodicIncl
};

struct MeshGrid {
    MeshAxis ax[3];
    const void* data;     // nodes [k][j][i][3], `dtype`
    long long sy, sz;     // node strides of j and k, in nodes
    int coord;            // kMeshCartesian or kMeshSpherical
    int order;            // 1 or 3
    int dtype;            // 0 double, 1 float
    int pad;
};

#ifdef __HIPCC__

__device__ __forceinline__ void mesh_node(const MeshGrid& g, long long idx, double o[3])
{
    if (g.dtype == 1) {
        const float* d = (const float*)g.data + 3 * idx;
        o[0] = (double)d[0]; o[1] = (double)d[1]; o[2] = (double)d[2];
    } else {
        const double* d = (const double*)g.data + 3 * idx;
        o[0] = d[0]; o[1] = d[1]; o[2] = d[2];
    }
}

// cell i (largest index in [0, n-2] with c[i] <= x) of a non-uniform axis,
// x inside [c[0], c[n-1]]
__device__ __forceinline__ int mesh_search(const MeshAxis& a, double x)
{
    int b = (int)((x - a.lo) * a.inv_bw);
    b = b < 0 ? 0 : (b >= a.nb ? a.nb - 1 : b);
    int i = a.bucket[b];
    while (i > 0 && a.c[i] > x) --i;
    while (i < a.n - 2 && a.c[i + 1] <= x) ++i;
    return i;
}

// cell index i in [0, n-2], fraction t and d t / d x of coordinate x on axis
// a; false when the field is NaN there (FillExtrap outside, NaN x).  `flat`
// is set where a clamp holds the coordinate (zero slope).
__device__ __forceinline__ bool mesh_locate(const MeshAxis& a, double x, int& i, double& t,
                                            double& dtdx, bool& flat)
{
    flat = false;
    if (a.n == 1) {
        i = 0; t = 0.0; dtdx = 0.0;
        return true;
    }
    if (x != x) return false;
    if (a.bc == kExtrapPeriodicIncl) {
        double u = x - a.lo;
        u = u - a.period * floor(u / a.period);
        x = a.lo + u;
        x = x > a.hi ? a.hi : x;
    }
    if (a.uniform) {
        double s = (x - a.lo) * a.inv_h;
        const double smax = (double)(a.n - 1);
        if (a.bc == kExtrapPeriodic) {
            if (!(fabs(s) < 4.0e18)) return false;
            const double fl = floor(s);
            t = s - fl;
            long long ii = (long long)fl % (long long)a.n;
            if (ii < 0) ii += a.n;
            i = (int)ii;          // the caller wraps i + 1 past n - 1
            dtdx = a.inv_h;
            return true;
        }
        if (a.bc == kExtrapFillNaN) {
            if (!(s >= 0.0 && s <= smax)) return false;
        } else {
            flat = a.bc == kExtrapClamp && (s < 0.0 || s > smax);
            s = s < 0.0 ? 0.0 : s;
            s = s > smax ? smax : s;
        }
        double fl = floor(s);
        if (fl > smax - 1.0) fl = smax - 1.0;
        t = s - fl;
        i = (int)fl;
        dtdx = flat ? 0.0 : a.inv_h;
        return true;
    }
    if (a.bc == kExtrapFillNaN) {
        if (!(x >= a.lo && x <= a.hi)) return false;
    } else if (a.bc == kExtrapClamp) {
        flat = x < a.lo || x > a.hi;
        x = x < a.lo ? a.lo : x;
        x = x > a.hi ? a.hi : x;
    }
    i = mesh_search(a, x);
    const double h = a.c[i + 1] - a.c[i];
    t = (x - a.c[i]) / h;
    dtdx = flat ? 0.0 : 1.0 / h;
    return true;
}

// node index of stencil entry i + k on axis a (k may step past either end;
// only periodic axes do so)
__device__ __forceinline__ int mesh_wrap(const MeshAxis& a, int j)
{
    if (a.n == 1) return 0;
    if (a.bc == kExtrapPeriodic) {
        j = j % a.n;
        return j < 0 ? j + a.n : j;
    }
    if (a.bc == kExtrapPeriodicIncl && (j < 0 || j > a.n - 1)) {
        const int m = a.n - 1;
        j = j % m;
        return j < 0 ? j + m : j;
    }
    return j;
}

// Catmull-Rom weights and their t-derivatives of nodes i-1 .. i+2, with the
// linear-extrapolation ghosts folded in at a non-periodic end
__device__ __forceinline__ void mesh_cubic_weights(const MeshAxis& a, int i, double t, int idx[4],
                                                   double w[4], double dw[4])
{
    if (a.n == 1) {
        idx[0] = idx[1] = idx[2] = idx[3] = 0;
        w[0] = 0.0; w[1] = 1.0; w[2] = 0.0; w[3] = 0.0;
        dw[0] = dw[1] = dw[2] = dw[3] = 0.0;
        return;
    }
    w[0] = ((-t + 2.0) * t - 1.0) * t * 0.5;
    w[1] = ((3.0 * t - 5.0) * t * t + 2.0) * 0.5;
    w[2] = ((-3.0 * t + 4.0) * t + 1.0) * t * 0.5;
    w[3] = (t - 1.0) * t * t * 0.5;
    dw[0] = ((-3.0 * t + 4.0) * t - 1.0) * 0.5;
    dw[1] = (9.0 * t - 10.0) * t * 0.5;
    dw[2] = ((-9.0 * t + 8.0) * t + 1.0) * 0.5;
    dw[3] = (3.0 * t - 2.0) * t * 0.5;
    const bool per = a.bc == kExtrapPeriodic || a.bc == kExtrapPeriodicIncl;
    if (!per && i == 0) {              // f_{-1} = 2 f_0 - f_1
        w[1] = w[1] + 2.0 * w[0];
        w[2] = w[2] - w[0];
        dw[1] = dw[1] + 2.0 * dw[0];
        dw[2] = dw[2] - dw[0];
        w[0] = 0.0;
        dw[0] = 0.0;
    }
    if (!per && i + 2 > a.n - 1) {     // f_n = 2 f_{n-1} - f_{n-2}
        w[2] = w[2] + 2.0 * w[3];
        w[1] = w[1] - w[3];
        dw[2] = dw[2] + 2.0 * dw[3];
        dw[1] = dw[1] - dw[3];
        w[3] = 0.0;
        dw[3] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = i - 1 + k;
        idx[k] = per ? mesh_wrap(a, j) : (j < 0 ? 0 : (j > a.n - 1 ? a.n - 1 : j));
    }
}

// value (and, when J is given, J[i][j] = dF_i / dq_j in the mesh coordinates
// q) of the interpolant at q; false where it is NaN
template <bool WANT_J>
__device__ __forceinline__ bool mesh_interp(const MeshGrid& g, const double q[3], double out[3],
                                            double J[3][3])
{
    int i[3];
    double t[3], dt[3];
    bool flat[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (!mesh_locate(g.ax[d], q[d], i[d], t[d], dt[d], flat[d])) return false;
    if (g.order == 1) {
        int i1[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) i1[d] = g.ax[d].n == 1 ? 0 : mesh_wrap(g.ax[d], i[d] + 1);
        const long long sy = g.sy, sz = g.sz;
        double c[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            mesh_node(g, (long long)((k & 1) ? i1[0] : i[0]) + sy * (((k >> 1) & 1) ? i1[1] : i[1]) +
                             sz * ((k >> 2) ? i1[2] : i[2]), c[k]);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const double c00 = fma(t[0], c[1][cc], (1.0 - t[0]) * c[0][cc]);
            const double c10 = fma(t[0], c[3][cc], (1.0 - t[0]) * c[2][cc]);
            const double c01 = fma(t[0], c[5][cc], (1.0 - t[0]) * c[4][cc]);
            const double c11 = fma(t[0], c[7][cc], (1.0 - t[0]) * c[6][cc]);
            const double c0 = fma(t[1], c10, (1.0 - t[1]) * c00);
            const double c1 = fma(t[1], c11, (1.0 - t[1]) * c01);
            out[cc] = fma(t[2], c1, (1.0 - t[2]) * c0);
        }
        if constexpr (WANT_J) {
#pragma unroll
            for (int a = 0; a < 3; ++a) J[a][0] = J[a][1] = J[a][2] = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double ax = (k & 1) ? t[0] : 1.0 - t[0];
                const double ay = ((k >> 1) & 1) ? t[1] : 1.0 - t[1];
                const double az = (k >> 2) ? t[2] : 1.0 - t[2];
                const double sx = (k & 1) ? 1.0 : -1.0;
                const double sy_ = ((k >> 1) & 1) ? 1.0 : -1.0;
                const double sz_ = (k >> 2) ? 1.0 : -1.0;
                const double dwx = (sx * ay) * az, dwy = (ax * sy_) * az, dwz = (ax * ay) * sz_;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    J[a][0] = J[a][0] + c[k][a] * dwx;
                    J[a][1] = J[a][1] + c[k][a] * dwy;
                    J[a][2] = J[a][2] + c[k][a] * dwz;
                }
            }
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int a = 0; a < 3; ++a) J[a][d] = J[a][d] * dt[d];
        }
        return true;
    }
    // order 3: tensor-product cardinal cubic, summed over x, then y, then z
    int ix[4], iy[4], iz[4];
    double wx[4], wy[4], wz[4], dwx[4], dwy[4], dwz[4];
    mesh_cubic_weights(g.ax[0], i[0], t[0], ix, wx, dwx);
    mesh_cubic_weights(g.ax[1], i[1], t[1], iy, wy, dwy);
    mesh_cubic_weights(g.ax[2], i[2], t[2], iz, wz, dwz);
    double v[3] = {0.0, 0.0, 0.0};
    double gx[3] = {0.0, 0.0, 0.0}, gy[3] = {0.0, 0.0, 0.0}, gz[3] = {0.0, 0.0, 0.0};
    for (int kz = 0; kz < 4; ++kz) {
        double vy[3] = {0.0, 0.0, 0.0}, gxy[3] = {0.0, 0.0, 0.0}, gyy[3] = {0.0, 0.0, 0.0};
        for (int ky = 0; ky < 4; ++ky) {
            double vx[3] = {0.0, 0.0, 0.0}, gxx[3] = {0.0, 0.0, 0.0};
            const long long row = g.sy * iy[ky] + g.sz * iz[kz];
#pragma unroll
            for (int kx = 0; kx < 4; ++kx) {
                double f[3];
                mesh_node(g, row + ix[kx], f);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    vx[a] = vx[a] + wx[kx] * f[a];
                    if constexpr (WANT_J) gxx[a] = gxx[a] + dwx[kx] * f[a];
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vy[a] = vy[a] + wy[ky] * vx[a];
                if constexpr (WANT_J) {
                    gxy[a] = gxy[a] + wy[ky] * gxx[a];
                    gyy[a] = gyy[a] + dwy[ky] * vx[a];
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = v[a] + wz[kz] * vy[a];
            if constexpr (WANT_J) {
                gx[a] = gx[a] + wz[kz] * gxy[a];
                gy[a] = gy[a] + wz[kz] * gyy[a];
                gz[a] = gz[a] + dwz[kz] * vy[a];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = v[a];
    if constexpr (WANT_J) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            J[a][0] = gx[a] * dt[0];
            J[a][1] = gy[a] * dt[1];
            J[a][2] = gz[a] * dt[2];
        }
    }
    return true;
}

// cart2sph (utility.jl:63-71): r = hypot(x, y, z) as Julia's n-argument
// hypot computes it, theta = acos(z / r), phi = mod2pi(atan(y, x))
__device__ __forceinline__ void mesh_cart2sph(const double x[3], double q[3])
{
    const double m = fmax(fmax(fabs(x[0]), fabs(x[1])), fabs(x[2]));
    if (m == 0.0) {
        q[0] = q[1] = q[2] = 0.0;
        return;
    }
    double r;
    if (m != m || isinf(m)) {
        r = (isinf(x[0]) || isinf(x[1]) || isinf(x[2])) ? __builtin_inf() : m;
    } else {
        const double a = x[0] / m, b = x[1] / m, c = x[2] / m;
        r = m * sqrt((a * a + b * b) + c * c);
    }
    q[0] = r;
    q[1] = acos(x[2] / r);
    double ph = atan2(x[1], x[0]);
    if (ph < 0.0) {
        // ph + 2 pi with 2 pi as a double-double (two-sum, then the low part)
        const double hi = 6.283185307179586, lo = 2.4492935982947064e-16;
        const double s = hi + ph;
        const double bb = s - hi;
        const double err = (hi - (s - bb)) + (ph - bb);
        ph = s + (err + lo);
    }
    q[2] = ph;
}

// F(x) of a mesh field at Cartesian x; J (optional) = dF_i / dx_j
template <bool WANT_J>
__device__ __forceinline__ void mesh_eval(const MeshGrid& g, const double x[3], double out[3],
                                          double J[3][3])
{
    const double nan = __builtin_nan("");
    if (g.coord == kMeshCartesian) {
        if (!mesh_interp<WANT_J>(g, x, out, J)) {
            out[0] = out[1] = out[2] = nan;
            if constexpr (WANT_J) {
#pragma unroll
                for (int a = 0; a < 3; ++a) J[a][0] = J[a][1] = J[a][2] = 0.0;
            }
        }
        return;
    }
    double q[3], s[3], Js[3][3];
    mesh_cart2sph(x, q);
    if (!mesh_interp<WANT_J>(g, q, s, Js)) {
        out[0] = out[1] = out[2] = nan;
        if constexpr (WANT_J) {
#pragma unroll
            for (int a = 0; a < 3; ++a) J[a][0] = J[a][1] = J[a][2] = 0.0;
        }
        return;
    }
    double st, ct, sp, cp;
    sincos(q[1], &st, &ct);
    sincos(q[2], &sp, &cp);
    // sph2cartvec (utility.jl:82-89)
    out[0] = (st * cp * s[0] + ct * cp * s[1]) - sp * s[2];
    out[1] = (st * sp * s[0] + ct * sp * s[1]) + cp * s[2];
    out[2] = ct * s[0] - st * s[1];
    if constexpr (WANT_J) {
        // F = R(theta, phi) S(r, theta, phi):
        // dF/dx_j = sum_q (dR/dq S + R dS/dq) dq/dx_j
        const double r = q[0];
        const double rho = sqrt(x[0] * x[0] + x[1] * x[1]);
        double dq[3][3];     // dq[k][j] = d q_k / d x_j
        for (int j = 0; j < 3; ++j) dq[0][j] = x[j] / r;
        if (rho > 0.0) {
            const double r2 = r * r;
            dq[1][0] = (x[2] * x[0]) / (r2 * rho);
            dq[1][1] = (x[2] * x[1]) / (r2 * rho);
            dq[1][2] = -rho / r2;
            dq[2][0] = -x[1] / (rho * rho);
            dq[2][1] = x[0] / (rho * rho);
            dq[2][2] = 0.0;
        } else {
            // on the polar axis theta and phi are not differentiable
            for (int j = 0; j < 3; ++j) dq[1][j] = dq[2][j] = nan;
        }
        // columns of R: e_r, e_theta, e_phi
        const double R[3][3] = {{st * cp, ct * cp, -sp}, {st * sp, ct * sp, cp}, {ct, -st, 0.0}};
        const double dRt[3][3] = {{ct * cp, -st * cp, 0.0}, {ct * sp, -st * sp, 0.0}, {-st, -ct, 0.0}};
        const double dRp[3][3] = {{-st * sp, -ct * sp, -cp}, {st * cp, ct * cp, -sp}, {0.0, 0.0, 0.0}};
        double dF[3][3];     // dF[a][k] = d F_a / d q_k
        for (int a = 0; a < 3; ++a) {
            double rs = 0.0, ts = 0.0, ps = 0.0;
            for (int b = 0; b < 3; ++b) {
                rs = rs + R[a][b] * Js[b][0];
                ts = ts + (dRt[a][b] * s[b] + R[a][b] * Js[b][1]);
                ps = ps + (dRp[a][b] * s[b] + R[a][b] * Js[b][2]);
            }
            dF[a][0] = rs; dF[a][1] = ts; dF[a][2] = ps;
        }
        for (int a = 0; a < 3; ++a)
            for (int j = 0; j < 3; ++j)
                J[a][j] = (dF[a][0] * dq[0][j] + dF[a][1] * dq[1][j]) + dF[a][2] * dq[2][j];
    }
}

// Out-of-line entry points.  Every generic kernel evaluates fields through a
// switch over the field kinds; keeping the mesh evaluator out of line keeps
// those kernels (and their compile time) the size they are without meshes.
// Arguments and results travel by value, in registers.
struct MeshVal {
    double v[3];
};
struct MeshValJac {
    double v[3];
    double J[9];
};

__device__ __noinline__ MeshVal mesh_value(const MeshGrid* g, double x0, double x1, double x2)
{
    const double x[3] = {x0, x1, x2};
    MeshVal r;
    mesh_eval<false>(*g, x, r.v, nullptr);
    return r;
}

__device__ __noinline__ MeshValJac mesh_value_jac(const MeshGrid* g, double x0, double x1, double x2)
{
    const double x[3] = {x0, x1, x2};
    double J[3][3];
    MeshValJac r;
    mesh_eval<true>(*g, x, r.v, J);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) r.J[3 * a + b] = J[a][b];
    return r;
}

#endif  // __HIPCC__

}  // namespace tpamd

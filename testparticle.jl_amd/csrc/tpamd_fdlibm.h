/*
 * tpamd_fdlibm.h -- double-precision atan and sin/cos, written once and
 * compiled both into the device kernels (hipcc) and into the CPU oracle (gcc,
 * plain C), so that the multistep pusher's reference-form coefficients
 * (src/boris/multistep_boris.jl:109-120: atan, then sincos of n atan) round
 * the same way on both sides.
 *
 * Method: FDLIBM (Sun Microsystems, 1993) as kept in FreeBSD msun and
 * openlibm, which Julia's Base.Math follows for Float64:
 *   - atan: s_atan.c -- reduction of |x| to [0, 7/16) by the identities
 *     atan(x) = atan(c) + atan((x - c) / (1 + c x)) with c in {1/2, 1, 3/2},
 *     or pi/2 + atan(-1/x); then an odd polynomial of degree 23 split into
 *     even and odd halves in w = x^4;
 *   - sin, cos on [-pi/4, pi/4]: k_sin.c / k_cos.c minimax polynomials,
 *     with the tail y of a reduced argument x + y;
 *   - argument reduction: e_rem_pio2.c's medium-size path, x - n pi/2 with
 *     pi/2 split into three 33-bit pieces plus tails (up to 151 bits),
 *     valid for |x| < ~2^20 pi/2 (the multistep pusher asks for n atan(tau)
 *     < n pi/2 with n <= 2^18 subcycles).
 * The polynomial (Horner) steps are fused multiply-adds, as Julia evaluates
 * them (@horner expands to muladd, which becomes an FMA on FMA hardware);
 * every other operation is rounded on its own.  Compile with
 * -ffp-contract=off so that nothing else is fused.
 *
 * Accuracy (checked against the host libm on many arguments by the oracle's
 * tests): atan within 1 ulp, sin and cos within 1 ulp on |x| < 2^19 pi/2.
 */
#ifndef TPAMD_FDLIBM_H
#define TPAMD_FDLIBM_H

#ifdef __HIPCC__
#define TPF_FN __host__ __device__ static inline
#else
#include <math.h>
#include <stdint.h>
#define TPF_FN static inline
#endif

/* biased exponent field of x */
TPF_FN int tpf_expbits(double x)
{
    unsigned long long u;
    __builtin_memcpy(&u, &x, sizeof u);
    return (int)((u >> 52) & 0x7ffull);
}

/* ---- atan (s_atan.c) ---------------------------------------------------- */
TPF_FN double tpf_atan(double x)
{
    /* atan(1/2), atan(1), atan(3/2), atan(inf) as head + tail */
    const double hi0 = 4.63647609000806093515e-01, lo0 = 2.26987774529616870924e-17;
    const double hi1 = 7.85398163397448278999e-01, lo1 = 3.06161699786838301793e-17;
    const double hi2 = 9.82793723247329054082e-01, lo2 = 1.39033110312309984516e-17;
    const double hi3 = 1.57079632679489655800e+00, lo3 = 6.12323399573676603587e-17;
    const double a0 = 3.33333333333329318027e-01, a1 = -1.99999999998764832476e-01,
                 a2 = 1.42857142725034663711e-01, a3 = -1.11111104054623557880e-01,
                 a4 = 9.09088713343650656196e-02, a5 = -7.69187620504482999495e-02,
                 a6 = 6.66107313738753120669e-02, a7 = -5.83357013379057348645e-02,
                 a8 = 4.97687799461593236017e-02, a9 = -3.65315727442169155270e-02,
                 a10 = 1.62858201153657823623e-02;
    if (x != x) return x + x;
    const double ax = fabs(x);
    const int neg = x < 0.0;
    if (ax >= 7.3786976294838206464e19) /* 2^66: atan = +-pi/2 */
        return neg ? -hi3 - lo3 : hi3 + lo3;
    int id;
    double r;
    if (ax < 0.4375) {
        if (ax < 7.450580596923828125e-9) return x; /* 2^-27: atan(x) = x */
        id = -1;
        r = x;
    } else if (ax < 1.1875) {
        if (ax < 0.6875) {
            id = 0;
            r = (2.0 * ax - 1.0) / (2.0 + ax);
        } else {
            id = 1;
            r = (ax - 1.0) / (ax + 1.0);
        }
    } else if (ax < 2.4375) {
        id = 2;
        r = (ax - 1.5) / (1.0 + 1.5 * ax);
    } else {
        id = 3;
        r = -1.0 / ax;
    }
    const double z = r * r;
    const double w = z * z;
    const double s1 = z * fma(w, fma(w, fma(w, fma(w, fma(w, a10, a8), a6), a4), a2), a0);
    const double s2 = w * fma(w, fma(w, fma(w, fma(w, a9, a7), a5), a3), a1);
    if (id < 0) return r - r * (s1 + s2);
    double hi, lo;
    if (id == 0) { hi = hi0; lo = lo0; }
    else if (id == 1) { hi = hi1; lo = lo1; }
    else if (id == 2) { hi = hi2; lo = lo2; }
    else { hi = hi3; lo = lo3; }
    const double q = hi - ((r * (s1 + s2) - lo) - r);
    return neg ? -q : q;
}

/* ---- sin and cos kernels on [-pi/4, pi/4] (k_sin.c, k_cos.c) ------------ *
 * x + y is the argument, |y| < ulp(x) / 2; iy = 0 for an unreduced one
 * (y = 0), 1 after a reduction.                                             */
TPF_FN double tpf_ksin(double x, double y, int iy)
{
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    const double z = x * x;
    const double w = z * z;
    const double r = fma(z, fma(z, S4, S3), S2) + (z * w) * fma(z, S6, S5);
    const double v = z * x;
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

TPF_FN double tpf_kcos(double x, double y)
{
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    const double z = x * x;
    const double w = z * z;
    const double r = z * fma(z, fma(z, C3, C2), C1) + (w * w) * fma(z, fma(z, C6, C5), C4);
    const double hz = 0.5 * z;
    const double u = 1.0 - hz;
    return u + (((1.0 - u) - hz) + (z * r - x * y));
}

/* ---- x - n pi/2 (e_rem_pio2.c, medium path) ----------------------------- *
 * Returns n; *y0 + *y1 is the reduced argument.                            */
TPF_FN int tpf_rem_pio2(double x, double* y0, double* y1)
{
    const double invpio2 = 6.36619772367581382433e-01;
    const double p1 = 1.57079632673412561417e+00, p1t = 6.07710050650619224932e-11;
    const double p2 = 6.07710050630396597660e-11, p2t = 2.02226624879595063154e-21;
    const double p3 = 2.02226624871116645580e-21, p3t = 8.47842766036889956997e-32;
    const double fn = rint(x * invpio2);
    const int n = (int)fn;
    double r = x - fn * p1;
    double w = fn * p1t;                       /* first round: good to 85 bits */
    double a = r - w;
    const int j = tpf_expbits(x);
    if (j - tpf_expbits(a) > 16) {             /* second round: good to 118 bits */
        double t = r;
        w = fn * p2;
        r = t - w;
        w = fn * p2t - ((t - r) - w);
        a = r - w;
        if (j - tpf_expbits(a) > 49) {         /* third round: 151 bits */
            t = r;
            w = fn * p3;
            r = t - w;
            w = fn * p3t - ((t - r) - w);
            a = r - w;
        }
    }
    *y0 = a;
    *y1 = (r - a) - w;
    return n;
}

/* sin(x) and cos(x) as s_sin.c / s_cos.c compute them; NaN from |x| about
 * 2^20 pi/2 on, where the medium reduction no longer holds (never asked for
 * here), and for NaN or infinite x.  The thresholds are FDLIBM's high-word
 * tests written as values. */
TPF_FN void tpf_sincos(double x, double* s, double* c)
{
    const double ax = fabs(x);
    if (!(ax < 1647100.0)) {                   /* high word > 0x413921fb, or NaN */
        *s = *c = __builtin_nan("");
        return;
    }
    if (ax < 0.7853984832763672) {             /* high word <= 0x3fe921fb: ~pi/4 */
        *s = ax < 1.4901161193847656e-08 ? x : tpf_ksin(x, 0.0, 0);     /* 2^-26 */
        *c = ax < 1.0536709282860102e-08 ? 1.0 : tpf_kcos(x, 0.0);      /* 2^-27 sqrt 2 */
        return;
    }
    double y0, y1;
    const int n = tpf_rem_pio2(x, &y0, &y1);
    const double sn = tpf_ksin(y0, y1, 1), cs = tpf_kcos(y0, y1);
    switch (n & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

#endif /* TPAMD_FDLIBM_H */

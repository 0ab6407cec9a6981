/*
 * tp_oracle.h -- CPU oracle for the ensemble Boris pusher.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (TestParticle.jl v0.23.2, src/boris/{boris,multistep_boris,adaptive_boris}.jl,
 * src/utility/interpolation.jl).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the CPU
 * baseline.  The product (testparticle.jl_amd/) never links, loads or calls it.
 *
 * Every function in tp_oracle.c cites the reference file:line it follows.
 */
#ifndef TP_ORACLE_H
#define TP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Field kinds.  The first block is the menu the device library supports;
 * the ORC_FIELD_KAT_* kinds restate closures that appear only in the
 * reference's own tests (test/test_boris.jl, test/test_boundary.jl). */
enum {
    ORC_FIELD_ZERO = 0,      /* src/utility/zero.jl:35-51 */
    ORC_FIELD_UNIFORM = 1,   /* F = v */
    ORC_FIELD_AFFINE = 2,    /* F_i = f0_i + sum_j J_ij x_j + ft_i t */
    ORC_FIELD_DIPOLE = 3,    /* B = b0 R^3 / r^5 (-3xz, -3yz, r^2 - 3z^2) */
    ORC_FIELD_GRID = 4,      /* trilinear on a uniform Cartesian grid */
    ORC_FIELD_UXB = 5,       /* E = -u x B(x, t), B = *src */
    ORC_FIELD_TIME_GRID = 6, /* LazyTimeInterpolator over gridded snapshots */
    ORC_FIELD_MESH = 7,      /* rectilinear / cubic / spherical / 2D / 1D grids */
    ORC_FIELD_KAT_FRONT = 100,      /* test_boris.jl:12-21 time_varying_B */
    ORC_FIELD_KAT_SWITCH_ON = 101,  /* test_boris.jl:252 B = t > t_on ? v : 0 */
    ORC_FIELD_KAT_NAN_OUTSIDE = 102 /* test_boundary.jl:10-26 */
};

enum { ORC_EXTRAP_FILL_NAN = 0, ORC_EXTRAP_CLAMP = 1, ORC_EXTRAP_PERIODIC = 2,
       ORC_EXTRAP_PERIODIC_INCL = 3 };
enum { ORC_MESH_CARTESIAN = 0, ORC_MESH_SPHERICAL = 1 };
enum { ORC_DTYPE_F64 = 0, ORC_DTYPE_F32 = 1 };

typedef struct orc_field {
    int kind;
    int extrap;
    int dtype;
    int pad_;
    double v[3];
    double jac[9];     /* row-major: F_i += jac[3*i+j] * x_j */
    double ft[3];
    double b0;
    double radius;
    double t_on;
    const void* data;  /* 3 x nx x ny x nz, component fastest (Julia layout) */
    long nx, ny, nz;
    double origin[3];
    double spacing[3];
    const struct orc_field* src;
    /* TIME_GRID: nsnap grids of the geometry above stored back to back in
     * data (3 nx ny nz values each), at strictly increasing times[nsnap] */
    const double* times;
    long nsnap;
    /* MESH: axes[d] holds the nx / ny / nz node coordinates of axis d;
     * coord ORC_MESH_*; order 1 or 3; mesh_extrap[d] the boundary per axis */
    const double* axes[3];
    int coord, order;
    int mesh_extrap[3];
    int pad2_;
} orc_field;

enum { ORC_ALG_BORIS = 0, ORC_ALG_MULTISTEP = 1 };
enum { ORC_FORM_REFERENCE = 0, ORC_FORM_RATIONAL = 1 };
enum {
    ORC_BND_NAN = 1,         /* any(isnan, u) */
    ORC_BND_SPHERE_IN = 2,   /* norm(x) < r_in   (removed inside) */
    ORC_BND_SPHERE_OUT = 4,  /* norm(x) > r_out */
    ORC_BND_BOX = 8          /* x outside [lo, hi] in any component */
};
enum { ORC_RET_SUCCESS = 0, ORC_RET_TERMINATED = 1, ORC_RET_MAXITERS = 2 };
enum { ORC_OK = 0, ORC_ERR_DT_TOO_SMALL = 1, ORC_ERR_MAXITERS = 2, ORC_ERR_ARG = 3 };

typedef struct orc_desc {
    int alg;
    int adaptive;
    int order_N;       /* 2, 4, 6 */
    int nsub;          /* n >= 1 */
    int form;          /* ORC_FORM_* (multistep only) */
    int save_start, save_end, save_everystep;
    int bnd_flags;
    int pad_;
    double safety;
    double q2m;
    double t0, t1, dt;
    long maxiters;
    long savestepinterval;
    long max_steps;    /* adaptive step cap, 0 = none (reference has none) */
    long nout_cap;     /* output records per particle */
    double r_in, r_out;
    double box_lo[3], box_hi[3];
    const double* dt_pp;   /* per-particle dt (fixed step), NULL = use dt */
    long nsteps_pp;        /* steps per particle when dt_pp != NULL */
    int save_fields;       /* records gain E, B (boris.jl:405-440) */
    int save_work;         /* records gain P_par, P_fermi, P_grad, P_betatron */
    double mass;           /* p[2], for the work rates */
    long* steps_out;       /* optional [N]: Boris steps each particle advanced */
} orc_desc;

/* record columns: 6 + 6 (save_fields) + 4 (save_work) */
#define ORC_RECORD_WIDTH(d) (6 + ((d)->save_fields ? 6 : 0) + ((d)->save_work ? 4 : 0))

/* boris.jl:19-31 */
void orc_boris_velocity_update(const double v[3], const double E[3],
                               const double B[3], double qdt_2m, double out[3]);
/* multistep_boris.jl:65-142 (form selects atan/sincos or rational coefficients) */
void orc_multistep_velocity(const double v[3], const double E[3], const double B[3],
                            double q2m, double dt, int n, int N, int form,
                            double out[3]);
/* coefficients c1, c2, c3, c6 of multistep_boris.jl:99-120 for |t_n|^2 = tau2 */
void orc_multistep_coeffs(double tau2, int n, int form, double c[4]);
/* integer polynomial coefficients of the rational form (see tp_oracle.c) */
int orc_rational_polys(int n, double* D, double* P1, double* P2, double* P3,
                       double* P6, int* deg);
void orc_field_eval(const orc_field* f, const double x[3], double t, double out[3]);
/* B, J[3*i+j] = dB_i/dx_j and dB/dt at (x, t): the derivatives the reference
 * takes with ForwardDiff (utility.jl:386-396, prepare.jl:64-74), in closed form */
void orc_field_jac(const orc_field* f, const double x[3], double t, double B[3], double J[9],
                   double dBdt[3]);
/* get_work_rates, equations.jl:394-441, with get_magnetic_properties and
 * _get_curvature (utility.jl:379-422): out = P_par, P_fermi, P_grad, P_betatron */
void orc_work_rates(double q2m, double m, const double v[3], const double E[3], const double B[3],
                    const double J[9], const double dBdt[3], double out[4]);
/* the extra record columns at n states: E, B when want_fields, then the work
 * rates when want_work (get_fields / get_work, boris.jl:599-636) */
void orc_record_eval(const orc_field* E, const orc_field* B, double q2m, double m, const double* xv,
                     const double* t, long n, int want_fields, int want_work, double* out);
/* boris.jl:383-403 */
long orc_calculate_nout(long nt, long K, int save_start, int save_end, int save_everystep);
/* boris.jl:345-362 */
int orc_prepare_checks(const orc_desc* d, long* nt, long* nout);
/* Ensemble solve (EnsembleThreads chunking, boris.jl:189-215 and
 * adaptive_boris.jl:196-213).  Records: out_u[(i*cap + k)*6 + c],
 * out_t[i*cap + k]; out_len[i] = records written (may exceed cap for the
 * adaptive solver, in which case the extra records were dropped). */
/* d x(T) / d (x0, v0) of a fixed-step standard Boris solve with E = 0 and a
 * periodic gridded B, by forward-mode dual numbers: xT[3], J[3][6] row-major */
int orc_boris_tangent(const orc_desc* d, const orc_field* B, const double xv0[6], double xT[3],
                      double J[18]);
/* Forward-mode directional derivatives of L = (1/N) sum_i |x_i(T) - x*_i|^2
 * (same solve) along ndir <= ORC_TANGENT_MAX directions in the grid values:
 * dense[ndir][3 nx ny nz] (node layout, component fastest) when dense is not
 * NULL, else the single values flat[ndir] = 3 node + c (distinct).  Particles
 * are split into nthreads contiguous chunks; *loss = L, dloss[k] = dL . dir_k. */
#define ORC_TANGENT_MAX 32
/* The full-orbit branch of the hybrid solver (src/hybrid.jl:257-452), with
 * the buffers and status codes of tpamd_fo_segment: state[N][8] (x, v, t,
 * dt), iters[N][2] (it, steps), flags[N] (in: 1 enter, 2 checked; out: 0
 * check due, 1 end, 2 outside), sync[N][6], records out_u[N][cap][6],
 * out_t[N][cap], out_len[N].  d: q2m, t1 = t_end, safety = safety_fo,
 * maxiters, save_everystep, savestepinterval, nout_cap, the boundary menu. */
int orc_fo_segment(const orc_desc* d, const orc_field* E, const orc_field* B, long N, double dtmin,
                   double dtmax, long check_interval, double* state, long* iters, int* flags,
                   double* sync, double* out_u, double* out_t, int* out_len);
/* atan and sin/cos of the multistep reference form (FDLIBM, shared with the
 * device library's tpamd_fdlibm.h) */
double orc_fdlibm_atan(double x);
void orc_fdlibm_sincos(double x, double* s, double* c);
int orc_loss_tangent(const orc_desc* d, const orc_field* B, long N, const double* xv0,
                     const double* targets, int ndir, const double* dense, const long* flat,
                     int nthreads, double* loss, double* dloss);
int orc_solve(const orc_desc* d, const orc_field* E, const orc_field* B,
              const double* xv0, long N, double* out_u, double* out_t,
              int* out_len, int* retcode, int nthreads, double* elapsed_s);

#ifdef __cplusplus
}
#endif
#endif

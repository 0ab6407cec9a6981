/*
 * tpamd.h -- C ABI of the MI355X (gfx950) ensemble Boris pusher.
 *
 * Drop-in boundary for TestParticle.jl's ensemble Boris path.  Plain C types
 * only: pointers, sizes, doubles.  No exceptions cross the boundary; every
 * entry point returns a tpamd_status and tpamd_last_error() holds the message.
 *
 * Layouts follow the Julia caller so that no host-side reshuffle is needed:
 *   - particle state  xv[N][6]        = Julia 6 x N column-major (x, y, z, vx, vy, vz)
 *   - saved records   u[N][cap][6]    = Julia 6 x cap x N column-major
 *   - saved times     t[N][cap]
 *   - gridded fields  A[nz][ny][nx][3] = Julia 3 x nx x ny x nz column-major
 *
 * Ownership: host buffers are caller-owned; device memory is owned by the
 * context (fields, staging) or by the caller (tpamd_solve_device).
 * Threading: a context is not thread-safe; calls on one context are
 * serialised by the caller.  Calls block until results are in the caller's
 * buffers (tpamd_solve) or the work is complete on the stream
 * (tpamd_solve_device).
 */
#ifndef TPAMD_H
#define TPAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPAMD_ABI_VERSION 1

typedef struct tpamd_ctx tpamd_ctx;
typedef struct tpamd_field tpamd_field;

typedef enum tpamd_status {
    TPAMD_OK = 0,
    TPAMD_ERR_ARGUMENT = 1,      /* invalid argument (ArgumentError) */
    TPAMD_ERR_DT_TOO_SMALL = 2,  /* |dt| < 10 eps, src/boris/boris.jl:349-351 */
    TPAMD_ERR_MAXITERS = 3,      /* nt > maxiters, src/boris/boris.jl:354-356 */
    TPAMD_ERR_ORDER = 4,         /* MultistepBoris N not in {2,4,6}, src/types.jl:70-75 */
    TPAMD_ERR_DEVICE = 5,        /* HIP runtime error */
    TPAMD_ERR_OOM = 6,           /* device allocation failed */
    TPAMD_ERR_UNSUPPORTED = 7,   /* combination the device path does not provide */
    TPAMD_ERR_CAPACITY = 8,      /* adaptive records exceeded nout_cap (stats->max_len says how many) */
    TPAMD_ERR_NONFINITE = 9      /* tpamd_adjoint: a loss or gradient entry is not finite (outputs are
                                    still written, so the caller can see which) */
} tpamd_status;

/* Field kinds: the device-side menu that replaces the Julia field callables
 * carried in p = (q2m, m, E, B, F) (src/prepare.jl:36-50, src/types.jl:32). */
typedef enum tpamd_field_kind {
    TPAMD_FIELD_ZERO = 0,     /* ZeroField, src/utility/zero.jl:1,35-51 */
    TPAMD_FIELD_UNIFORM = 1,  /* constant vector */
    TPAMD_FIELD_AFFINE = 2,   /* F_i = f0_i + sum_j J_ij x_j + ft_i t */
    TPAMD_FIELD_DIPOLE = 3,   /* B = b0 R^3/r^5 (-3xz, -3yz, r^2-3z^2) */
    TPAMD_FIELD_GRID = 4,     /* FieldInterpolator over a uniform CartesianGrid, order 1 */
    TPAMD_FIELD_UXB = 5,      /* E = -u x B(x), B the solve's magnetic field */
    TPAMD_FIELD_TIME_GRID = 6,/* LazyTimeInterpolator over gridded snapshots */
    TPAMD_FIELD_MESH = 7      /* rectilinear / cubic / spherical / 2D / 1D grids */
} tpamd_field_kind;

/* Extrapolation of gridded fields, src/utility/interpolation.jl:203-206 */
typedef enum tpamd_extrap {
    TPAMD_EXTRAP_FILL_NAN = 0, /* FillExtrap(NaN), the prepare() default (src/prepare.jl:257-265) */
    TPAMD_EXTRAP_CLAMP = 1,    /* ClampExtrap() */
    TPAMD_EXTRAP_PERIODIC = 2, /* WrapExtrap() / PeriodicBC(endpoint=:exclusive): period N*h */
    TPAMD_EXTRAP_PERIODIC_INCLUSIVE = 3 /* PeriodicBC(): period x_N - x_1 (meshes only) */
} tpamd_extrap;

typedef enum tpamd_dtype { TPAMD_F64 = 0, TPAMD_F32 = 1 } tpamd_dtype;

typedef enum tpamd_alg { TPAMD_ALG_BORIS = 0, TPAMD_ALG_MULTISTEP = 1 } tpamd_alg;

/* Multistep coefficient evaluation: RATIONAL evaluates cos/sin(2 n atan tau)
 * as integer-polynomial ratios in tau^2 (no transcendental, no Taylor
 * branch); REFERENCE follows src/boris/multistep_boris.jl:94-120 literally. */
typedef enum tpamd_form { TPAMD_FORM_REFERENCE = 0, TPAMD_FORM_RATIONAL = 1 } tpamd_form;

/* FP64: state, fields and arithmetic in double, rounding-identical to the
 * CPU restatement of the reference.  FP32C: field storage and gather in
 * float, Boris rotation in float, positions as a compensated float pair
 * (hi + lo) with periodic wrap counts, time kept in double. */
typedef enum tpamd_precision { TPAMD_FP64 = 0, TPAMD_FP32C = 1 } tpamd_precision;

/* isoutside(u, p, t) menu (src/boris/boris.jl:95,457-461;
 * src/boris/adaptive_boris.jl:26,227-231).  Flags are OR-ed. */
typedef enum tpamd_boundary {
    TPAMD_BND_NAN = 1,        /* any(isnan, u) */
    TPAMD_BND_SPHERE_IN = 2,  /* norm(x) < r_in  (e.g. r < R_E) */
    TPAMD_BND_SPHERE_OUT = 4, /* norm(x) > r_out */
    TPAMD_BND_BOX = 8,        /* x outside [box_lo, box_hi] in any component */
    TPAMD_BND_GRID = 16       /* x outside the node span of B's grid (open boundary) */
} tpamd_boundary;

/* Per-trajectory return codes (SciMLBase ReturnCode) */
typedef enum tpamd_retcode {
    TPAMD_RET_SUCCESS = 0,
    TPAMD_RET_TERMINATED = 1, /* isoutside fired; src/boris/boris.jl:545-551 */
    TPAMD_RET_MAXITERS = 2    /* adaptive step cap reached (max_steps) */
} tpamd_retcode;

/* Solve descriptor: the keyword arguments of
 *   solve(prob, alg, [backend], [ensemblealg]; dt, trajectories, savestepinterval,
 *         isoutside, save_start, save_end, save_everystep, maxiters)
 * (src/boris/boris.jl:88-106; src/boris/adaptive_boris.jl:21-33;
 *  src/boris/boris_kernel.jl:312-317) plus the algorithm struct fields
 * (src/types.jl:51-107).  Zero-initialise, then set what is needed. */
typedef struct tpamd_solve_desc {
    int32_t alg;              /* tpamd_alg */
    int32_t adaptive;         /* 0: fixed step (Boris{false}); 1: AdaptiveBoris / AdaptiveMultistepBoris */
    int32_t order_N;          /* MultistepBoris{N}: 2, 4 or 6 */
    int32_t nsub;             /* MultistepBoris n >= 1 */
    int32_t form;             /* tpamd_form (multistep only) */
    int32_t precision;        /* tpamd_precision */
    int32_t save_start;
    int32_t save_end;
    int32_t save_everystep;
    int32_t bnd_flags;        /* tpamd_boundary bits */
    double safety;            /* adaptive: dt = safety * 2 pi / |q2m B| */
    double q2m;               /* p[1] = q/m (src/prepare.jl:211) */
    double t0, t1;            /* tspan */
    double dt;                /* fixed step */
    int64_t maxiters;         /* default 1_000_000 */
    int64_t savestepinterval; /* K >= 1 */
    int64_t max_steps;        /* adaptive step cap per particle; 0 = none (reference) */
    int64_t nout_cap;         /* record capacity per particle (>= tpamd_plan's nout) */
    double r_in, r_out;
    double box_lo[3], box_hi[3];
    const double* dt_pp;      /* optional per-particle dt (fixed step): tspan_i = (t0, t0 + nsteps_pp dt_i) */
    int64_t nsteps_pp;
    int32_t block_size;       /* 0 = library default */
    int32_t flags;            /* TPAMD_FLAG_* bits */
    double mass;              /* p[2] = m, for save_work (src/boris/boris.jl:405-440) */
    int32_t save_fields;      /* records gain E(x, t) and B(x, t): 6 + 6 columns */
    int32_t save_work;        /* records gain (P_par, P_fermi, P_grad, P_betatron) */
    int32_t particles_per_lane; /* adaptive solves only: each wave works through 64 x this many
                                   particles, a lane taking the next one when its own ends;
                                   0 = library choice; a batch starts from this (0: 4) until
                                   tpamd_batch_autotune picks its own */
    int32_t reserved;
} tpamd_solve_desc;

/* Columns of one record: 6 (x, v) + 6 when save_fields + 4 when save_work,
 * in that order, as _prepare_saved_data builds them (src/boris/boris.jl:405-440).
 * Record buffers are [N][nout_cap][record width]. */
#define TPAMD_RECORD_WIDTH(d) (6 + ((d)->save_fields ? 6 : 0) + ((d)->save_work ? 4 : 0))

/* descriptor flags */
#define TPAMD_FLAG_GENERIC_KERNELS 1  /* skip the specialised kernels (gridded B, uniform fields, the dipole): for cross-checks */

typedef struct tpamd_stats {
    double kernel_ms;         /* device time of the push kernel, HIP events on the solve stream */
    double total_ms;          /* host wall time of the call */
    int64_t particle_steps;   /* Boris steps advanced, summed over particles */
    int64_t nt;               /* fixed-step count (0 for adaptive) */
    int64_t nout;             /* records per particle the schedule asks for (fixed step) */
    int64_t max_len;          /* largest record count of any particle */
    double aux_ms;            /* second phase where there is one (adjoint: reverse sweep) */
} tpamd_stats;

/* ---- library / context -------------------------------------------------- */
int32_t tpamd_abi_version(void);
/* Opens a context on HIP device `device` (its own default-priority stream). */
int32_t tpamd_init(int32_t device, tpamd_ctx** out);
int32_t tpamd_finalize(tpamd_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed tpamd_init when ctx is NULL). */
const char* tpamd_last_error(const tpamd_ctx* ctx);

/* ---- fields ------------------------------------------------------------- *
 * Replace prepare(E, B; ...) / prepare(x, y, z, E, B; order=1, bc) and
 * build_interpolator(CartesianGrid, A::4D, x, y, z, 1, bc)
 * (src/prepare.jl:93-104,145-153; src/utility/interpolation.jl:217-235).   */
int32_t tpamd_field_zero(tpamd_ctx* ctx, tpamd_field** out);
int32_t tpamd_field_uniform(tpamd_ctx* ctx, const double value[3], tpamd_field** out);
int32_t tpamd_field_affine(tpamd_ctx* ctx, const double f0[3], const double jac[9],
                           const double ft[3], tpamd_field** out);
/* Dipole with B(R, equator) = (0, 0, b0); the reference's tests use
 * Magnetostatics.Dipole(BMoment_Earth) (test/test_gc.jl:180). */
int32_t tpamd_field_dipole(tpamd_ctx* ctx, double b0, double radius, tpamd_field** out);
/* Uniform Cartesian grid: node (i,j,k) at origin + (i,j,k) .* spacing.
 * `data` is a host array of `dtype` in the Julia 3 x nx x ny x nz layout; it
 * is copied to the device and stored as `storage_dtype`. */
int32_t tpamd_field_grid(tpamd_ctx* ctx, const void* data, int32_t dtype, int64_t nx,
                         int64_t ny, int64_t nz, const double origin[3],
                         const double spacing[3], int32_t extrap, int32_t storage_dtype,
                         tpamd_field** out);
/* Time-dependent gridded field: the device counterpart of
 * LazyTimeInterpolator(times, loader) whose loader(k) returns
 * build_interpolator(CartesianGrid, A_k, x, y, z, 1, bc) on one grid
 * (src/utility/interpolation.jl:337-430, test/test_utility.jl:312-404).
 * snapshots[k] is the host array of snapshot k (0-based), laid out as for
 * tpamd_field_grid; times[nsnap] are finite and strictly increasing.  All
 * snapshots are copied to the device and stay resident there (the reference
 * loads them lazily and keeps at most three in memory).  At time t, with k = searchsortedlast(times, t):
 * k == 0 gives snapshot 1, k == nsnap snapshot nsnap, otherwise
 * (1 - w) A_k(x) + w A_{k+1}(x) with w = (t - t_k) / (t_{k+1} - t_k).  Its
 * Jacobian blends the same way and its time derivative is
 * (A_{k+1}(x) - A_k(x)) / (t_{k+1} - t_k) inside the time span, 0 outside. */
int32_t tpamd_field_time_grid(tpamd_ctx* ctx, const void* const* snapshots, int32_t dtype,
                              int64_t nx, int64_t ny, int64_t nz, int32_t nsnap,
                              const double* times, const double origin[3],
                              const double spacing[3], int32_t extrap, int32_t storage_dtype,
                              tpamd_field** out);
/* Gridded field on a general mesh, the device counterpart of
 * build_interpolator(gridtype, A, g1, g2, g3, order, bc)
 * (src/utility/interpolation.jl:15-117,140-178,217-319):
 *   coord = TPAMD_MESH_CARTESIAN: CartesianGrid / RectilinearGrid with any
 *     strictly increasing axes (order 1; order 3 on uniform axes); an axis of
 *     one node is ignored, which gives the 2D and 1D interpolators;
 *   coord = TPAMD_MESH_SPHERICAL: StructuredGrid with axes (r, theta, phi),
 *     values (B_r, B_theta, B_phi) rotated to Cartesian with sph2cartvec.
 * `data` is laid out as for tpamd_field_grid with n[0] x n[1] x n[2] nodes;
 * axes[d] holds the n[d] node coordinates of axis d; extrap[d] is that axis's
 * boundary (TPAMD_EXTRAP_*; the exclusive periodic wrap needs a uniform axis).
 * Order 3 is the cardinal cubic: Catmull-Rom weights on nodes i-1 .. i+2,
 * with the linear extrapolation 2 f_0 - f_1 standing in for a missing
 * neighbour at a non-periodic end. */
enum { TPAMD_MESH_CARTESIAN = 0, TPAMD_MESH_SPHERICAL = 1 };
int32_t tpamd_field_mesh(tpamd_ctx* ctx, const void* data, int32_t dtype, const int64_t n[3],
                         const double* const axes[3], int32_t coord, int32_t order,
                         const int32_t extrap[3], int32_t storage_dtype, tpamd_field** out);
/* Builds (enable = 1) or drops (0) a cell-replicated copy of a gridded field:
 * per cell the 8 corner nodes x 3 components in one block (float: 128 B, one
 * cache line; double: 192 B), so a trilinear gather reads one contiguous
 * block instead of eight scattered nodes.  Costs about 10x (float) / 8x
 * (double) the node array in device memory; the cell array must stay below
 * 4 GiB.  Results are unchanged (same values, same blend order). */
int32_t tpamd_field_grid_cells(tpamd_ctx* ctx, tpamd_field* f, int32_t enable);
/* E = -u x B(x) for a uniform drift u; `B` must be the B field of the solve. */
int32_t tpamd_field_uxb(tpamd_ctx* ctx, const double u[3], const tpamd_field* B,
                        tpamd_field** out);
/* Evaluates a field at n points (host arrays x[n][3] -> out[n][3]) on the device:
 * the device counterpart of calling the Julia field callable, E(r, t). */
int32_t tpamd_field_eval(tpamd_ctx* ctx, const tpamd_field* f, const tpamd_field* B,
                         const double* x, int64_t n, double t, int32_t precision,
                         double* out);
int32_t tpamd_field_free(tpamd_ctx* ctx, tpamd_field* f);
/* Fields and work rates at n states (host arrays xv[n][6], t[n]) on the
 * device, in FP64: out[n][w] with w = 6 (E, B) when want_fields, + 4 (P_par,
 * P_fermi, P_grad, P_betatron) when want_work, fields first.  The device
 * counterpart of get_fields / get_work (src/boris/boris.jl:599-636,
 * src/equations.jl:394-441, src/utility/utility.jl:379-422). */
int32_t tpamd_record_eval(tpamd_ctx* ctx, const tpamd_field* E, const tpamd_field* B, double q2m,
                          double mass, const double* xv, const double* t, int64_t n,
                          int32_t want_fields, int32_t want_work, double* out);

/* ---- solve -------------------------------------------------------------- */
/* _prepare_checks + _calculate_nout (src/boris/boris.jl:345-403): step count
 * and records per particle of a fixed-step solve; raises the same errors. */
int32_t tpamd_plan(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int64_t* nt, int64_t* nout);

/* Ensemble solve with host buffers; replaces
 *   solve(prob, ::Boris{false}, backend::KA.Backend, ensemblealg; ...)
 *     (src/boris/boris_kernel.jl:312-382)
 * and the CPU ensemble paths it mirrors
 *   _solve(::EnsembleSerial/::EnsembleThreads, ...) (src/boris/boris.jl:137-215),
 *   _solve_adaptive(...) (src/boris/adaptive_boris.jl:40-71).
 * xv0[N][6] initial states (prob_func already applied by the caller);
 * out_u[N][cap][6], out_t[N][cap], out_len[N], retcode[N]; cap = desc->nout_cap. */
int32_t tpamd_solve(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
                    const tpamd_field* B, const double* xv0, int64_t N, double* out_u,
                    double* out_t, int32_t* out_len, int32_t* retcode, tpamd_stats* stats);

/* Same solve with device-resident buffers on `stream` (a hipStream_t; NULL =
 * the context's stream).  desc->dt_pp, if set, is a device pointer.  Returns
 * after the work on the stream has completed. */
int32_t tpamd_solve_device(tpamd_ctx* ctx, const tpamd_solve_desc* desc,
                           const tpamd_field* E, const tpamd_field* B, const double* d_xv0,
                           int64_t N, double* d_out_u, double* d_out_t, int32_t* d_out_len,
                           int32_t* d_retcode, void* stream, tpamd_stats* stats);

/* ---- adjoint ---------------------------------------------------------------- *
 * Gradient of the trajectory loss L = (1/N) sum_i |x_i(T) - x*_i|^2 (final
 * positions of a fixed-step standard Boris solve, E = 0, gridded B) with
 * respect to every B node value and every initial state.  The reference
 * differentiates with ForwardDiff through the CPU solve (src/utility/utility.jl,
 * test/test_derivatives.jl); there is no reverse-mode counterpart.
 *   desc: alg = BORIS, adaptive = 0, precision FP64 or FP32C, no boundary.
 *   ckpt_interval K: FP64 state checkpoints every K steps; the reverse sweep
 *     recomputes each K-step segment.
 *   targets[N][3]; dL_dB in B's node layout [nz][ny][nx][3] (summed in double);
 *   dL_dxv0[N][6].  stats->kernel_ms = forward pass, aux_ms = reverse.
 *   FP32C: forward, recompute and reverse steps in float; the adjoint state
 *   carries a per-particle power-of-two exponent, so gradients far beyond the
 *   float range (a chaotic flow's grow like exp(lambda T)) stay finite.
 *   Returns TPAMD_ERR_NONFINITE when any output entry is not finite. */
int32_t tpamd_adjoint(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* B,
                      const double* xv0, int64_t N, const double* targets,
                      int32_t ckpt_interval, double* loss, double* dL_dB, double* dL_dxv0,
                      tpamd_stats* stats);

/* ---- AdaptiveHybrid full-orbit segments ----------------------------------- *
 * The full-orbit (Boris) branch of the hybrid guiding-centre / full-orbit
 * solver, _hybrid_adaptive! (src/hybrid.jl:257-452): entering full orbit
 * (dt = clamp(2 pi safety_fo / (|q2m| |B(x, t)|), dtmin, dtmax) and the back
 * half-step at t, :262-267 / :298-304), the fixed-dt leapfrog steps with the
 * final step truncated at t_end (:411-438), and the synchronised state the
 * adiabaticity check (:361-364) and the closing record (:441-452) use.  The
 * guiding-centre branch and every mode decision stay with the caller, which
 * calls this for the particles it keeps in full orbit.
 *   desc: alg BORIS, precision FP64, q2m, t1 = t_end, safety = safety_fo,
 *     maxiters (the solver's step cap), save_everystep, savestepinterval,
 *     the isoutside menu (bnd_flags ...), nout_cap = record capacity per
 *     particle for this call; dt and t0 are not used.
 *   Host buffers, per particle i:
 *     state[i][8]  in/out: x (3), v (3; staggered by -dt/2 once entered), t, dt
 *     iters[i][2]  in/out: the solver's counters it and steps
 *     flags[i]     in: TPAMD_FO_ENTER (enter full orbit first), TPAMD_FO_CHECKED
 *                  (the check at this it is done); out: TPAMD_FO_CHECK (stopped
 *                  at the next it % check_interval == 0), TPAMD_FO_END (t >= t_end
 *                  or steps >= maxiters) or TPAMD_FO_OUTSIDE (isoutside fired:
 *                  x and t kept, v the updated one, as :428-430 leaves them)
 *     sync[i][6]   out: (x, U(v, x, dt/2, t)) at the particle's t
 *     out_u[i][cap][6], out_t[i][cap], out_len[i]: the records this call saved
 *     (out_len may exceed cap: records past cap are dropped and stats->max_len
 *     says how many were due). */
enum { TPAMD_FO_ENTER = 1, TPAMD_FO_CHECKED = 2 };
enum { TPAMD_FO_CHECK = 0, TPAMD_FO_END = 1, TPAMD_FO_OUTSIDE = 2 };
int32_t tpamd_fo_segment(tpamd_ctx* ctx, const tpamd_solve_desc* desc, const tpamd_field* E,
                         const tpamd_field* B, int64_t N, double dtmin, double dtmax,
                         int64_t check_interval, double* state, int64_t* iters, int32_t* flags,
                         double* sync, double* out_u, double* out_t, int32_t* out_len,
                         tpamd_stats* stats);

/* ---- batched ensembles ---------------------------------------------------- *
 * Many field instances with particle populations of uneven size and lifetime
 * in one persistent launch.  Instance k owns particles [offsets[k],
 * offsets[k+1]) of the concatenated state array; one descriptor applies to
 * all instances.  The reference runs such work as one solve per field
 * (src/boris/boris.jl:88-215 / adaptive_boris.jl:21-71 per instance); this
 * entry point has no single reference counterpart. */
typedef struct tpamd_batch tpamd_batch;

/* Binds device buffers (as tpamd_solve_device), uploads the per-instance
 * parameter table and the work chunks, and captures the launch in a HIP graph.
 * E[k], B[k] stay owned by the caller and must outlive the batch. */
int32_t tpamd_batch_create(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int32_t n_inst,
                           const tpamd_field* const* E, const tpamd_field* const* B,
                           const int64_t* offsets, const double* d_xv0, double* d_out_u,
                           double* d_out_t, int32_t* d_out_len, int32_t* d_retcode,
                           tpamd_batch** out);
/* Times every (block size, particles per thread) pair on `stream`, keeps the
 * fastest and re-captures the graph.  Returns the chosen pair. */
int32_t tpamd_batch_autotune(tpamd_batch* b, void* stream, int32_t* block_size, int32_t* ppt);
/* One replay of the captured launch on `stream` (NULL = the context's); blocks
 * until it has completed. */
int32_t tpamd_batch_run(tpamd_batch* b, void* stream, tpamd_stats* stats);
int32_t tpamd_batch_free(tpamd_batch* b);
/* Host-buffer convenience: create, run once, copy back, free. */
int32_t tpamd_solve_batch(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int32_t n_inst,
                          const tpamd_field* const* E, const tpamd_field* const* B,
                          const int64_t* offsets, const double* xv0, double* out_u,
                          double* out_t, int32_t* out_len, int32_t* retcode,
                          tpamd_stats* stats);

/* ---- allocation-free batched solve ----------------------------------------- *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the batched solve of tpamd_batch_create + tpamd_batch_run
 * (same kernels, same results, device buffers as there) as one call that
 * calls no hipMalloc / hipFree.  Its scratch (queue claim counters, step
 * counter, the per-instance parameter table, the chunk list) is a fixed
 * 3.4 MiB device array each GPU holds until the process ends, so:
 *   - at most 256 instances and 131,072 chunks of 64 x particles_per_lane
 *     particles (particles_per_lane from desc, default 8: ~67M particles);
 *     larger batches return TPAMD_ERR_CAPACITY;
 *   - calls that overlap on one GPU (two host threads, or two streams in
 *     flight) race on that array; serialise them.
 * No graph is captured and nothing is autotuned (desc->block_size, default
 * 256).  Host memory for the table is ordinary heap memory.  Returns after
 * the work on `stream` (NULL = the context's) has completed. */
int32_t tpamd_solve_batch_noalloc(tpamd_ctx* ctx, const tpamd_solve_desc* desc, int32_t n_inst,
                                  const tpamd_field* const* E, const tpamd_field* const* B,
                                  const int64_t* offsets, const double* d_xv0, double* d_out_u,
                                  double* d_out_t, int32_t* d_out_len, int32_t* d_retcode,
                                  void* stream, tpamd_stats* stats);

#ifdef __cplusplus
}
#endif
#endif

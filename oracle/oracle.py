"""ctypes binding of the CPU oracle (oracle/tp_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  The product
package never imports this module.

Field and descriptor structs mirror oracle/tp_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtporacle.so")

FIELD_ZERO, FIELD_UNIFORM, FIELD_AFFINE, FIELD_DIPOLE, FIELD_GRID, FIELD_UXB = 0, 1, 2, 3, 4, 5
FIELD_TIME_GRID, FIELD_MESH = 6, 7
EXTRAP_PERIODIC_INCL = 3
MESH_CARTESIAN, MESH_SPHERICAL = 0, 1
FIELD_KAT_FRONT, FIELD_KAT_SWITCH_ON, FIELD_KAT_NAN_OUTSIDE = 100, 101, 102
EXTRAP_FILL_NAN, EXTRAP_CLAMP, EXTRAP_PERIODIC = 0, 1, 2
ALG_BORIS, ALG_MULTISTEP = 0, 1
FORM_REFERENCE, FORM_RATIONAL = 0, 1
BND_NAN, BND_SPHERE_IN, BND_SPHERE_OUT, BND_BOX = 1, 2, 4, 8
RET_SUCCESS, RET_TERMINATED, RET_MAXITERS = 0, 1, 2
ERR_NAMES = {1: "time step dt is too small", 2: "nt exceeds maxiters", 3: "invalid argument"}


class OrcField(C.Structure):
    pass


OrcField._fields_ = [
    ("kind", C.c_int), ("extrap", C.c_int), ("dtype", C.c_int), ("pad_", C.c_int),
    ("v", C.c_double * 3), ("jac", C.c_double * 9), ("ft", C.c_double * 3),
    ("b0", C.c_double), ("radius", C.c_double), ("t_on", C.c_double),
    ("data", C.c_void_p), ("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long),
    ("origin", C.c_double * 3), ("spacing", C.c_double * 3),
    ("src", C.POINTER(OrcField)),
    ("times", C.POINTER(C.c_double)), ("nsnap", C.c_long),
    ("axes", C.POINTER(C.c_double) * 3), ("coord", C.c_int), ("order", C.c_int),
    ("mesh_extrap", C.c_int * 3), ("pad2_", C.c_int),
]


class OrcDesc(C.Structure):
    _fields_ = [
        ("alg", C.c_int), ("adaptive", C.c_int), ("order_N", C.c_int), ("nsub", C.c_int),
        ("form", C.c_int), ("save_start", C.c_int), ("save_end", C.c_int),
        ("save_everystep", C.c_int), ("bnd_flags", C.c_int), ("pad_", C.c_int),
        ("safety", C.c_double), ("q2m", C.c_double),
        ("t0", C.c_double), ("t1", C.c_double), ("dt", C.c_double),
        ("maxiters", C.c_long), ("savestepinterval", C.c_long), ("max_steps", C.c_long),
        ("nout_cap", C.c_long), ("r_in", C.c_double), ("r_out", C.c_double),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
        ("dt_pp", C.POINTER(C.c_double)), ("nsteps_pp", C.c_long),
        ("save_fields", C.c_int), ("save_work", C.c_int), ("mass", C.c_double),
        ("steps_out", C.POINTER(C.c_long)),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d3 = C.POINTER(C.c_double)
        L.orc_boris_velocity_update.argtypes = [d3, d3, d3, C.c_double, d3]
        L.orc_multistep_velocity.argtypes = [d3, d3, d3, C.c_double, C.c_double, C.c_int,
                                             C.c_int, C.c_int, d3]
        L.orc_multistep_coeffs.argtypes = [C.c_double, C.c_int, C.c_int, d3]
        L.orc_rational_polys.argtypes = [C.c_int, d3, d3, d3, d3, d3, C.POINTER(C.c_int)]
        L.orc_boris_tangent.argtypes = [C.POINTER(OrcDesc), C.POINTER(OrcField), d3, d3, d3]
        L.orc_fo_segment.argtypes = [C.POINTER(OrcDesc), C.POINTER(OrcField), C.POINTER(OrcField),
                                     C.c_long, C.c_double, C.c_double, C.c_long, d3,
                                     C.POINTER(C.c_long), C.POINTER(C.c_int), d3, d3, d3,
                                     C.POINTER(C.c_int)]
        L.orc_fdlibm_atan.argtypes = [C.c_double]
        L.orc_fdlibm_atan.restype = C.c_double
        L.orc_fdlibm_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_loss_tangent.argtypes = [C.POINTER(OrcDesc), C.POINTER(OrcField), C.c_long, d3, d3,
                                       C.c_int, d3, C.POINTER(C.c_long), C.c_int, d3, d3]
        L.orc_field_eval.argtypes = [C.POINTER(OrcField), d3, C.c_double, d3]
        L.orc_field_jac.argtypes = [C.POINTER(OrcField), d3, C.c_double, d3, d3, d3]
        L.orc_work_rates.argtypes = [C.c_double, C.c_double, d3, d3, d3, d3, d3, d3]
        L.orc_record_eval.argtypes = [C.POINTER(OrcField), C.POINTER(OrcField), C.c_double,
                                      C.c_double, d3, d3, C.c_long, C.c_int, C.c_int, d3]
        L.orc_calculate_nout.argtypes = [C.c_long, C.c_long, C.c_int, C.c_int, C.c_int]
        L.orc_calculate_nout.restype = C.c_long
        L.orc_prepare_checks.argtypes = [C.POINTER(OrcDesc), C.POINTER(C.c_long),
                                         C.POINTER(C.c_long)]
        L.orc_solve.argtypes = [C.POINTER(OrcDesc), C.POINTER(OrcField), C.POINTER(OrcField),
                                d3, C.c_long, d3, d3, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.c_int, d3]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# Field constructors (keep numpy buffers alive on the returned object)
# ---------------------------------------------------------------------------
class Field:
    def __init__(self, st: OrcField, keep=()):
        self.st = st
        self.keep = keep


def zero() -> Field:
    return Field(OrcField(kind=FIELD_ZERO))


def uniform(v) -> Field:
    st = OrcField(kind=FIELD_UNIFORM)
    st.v[:] = [float(a) for a in v]
    return Field(st)


def affine(f0, jac=None, ft=None) -> Field:
    st = OrcField(kind=FIELD_AFFINE)
    st.v[:] = [float(a) for a in f0]
    st.jac[:] = [float(a) for a in (np.zeros(9) if jac is None else np.ravel(jac))]
    st.ft[:] = [float(a) for a in (np.zeros(3) if ft is None else ft)]
    return Field(st)


def dipole(b0: float, radius: float) -> Field:
    return Field(OrcField(kind=FIELD_DIPOLE, b0=b0, radius=radius))


def grid(data: np.ndarray, origin, spacing, extrap: int) -> Field:
    """data: array of shape (nz, ny, nx, 3) (C order), i.e. Julia's
    3 x nx x ny x nz column-major array."""
    assert data.ndim == 4 and data.shape[3] == 3
    assert data.dtype in (np.float64, np.float32)
    data = np.ascontiguousarray(data)
    nz, ny, nx, _ = data.shape
    st = OrcField(kind=FIELD_GRID, extrap=extrap, dtype=0 if data.dtype == np.float64 else 1,
                  data=data.ctypes.data, nx=nx, ny=ny, nz=nz)
    st.origin[:] = [float(a) for a in origin]
    st.spacing[:] = [float(a) for a in spacing]
    return Field(st, keep=(data,))


def time_grid(snaps: np.ndarray, times, origin, spacing, extrap: int) -> Field:
    """LazyTimeInterpolator over gridded snapshots
    (src/utility/interpolation.jl:337-430).  snaps: (nsnap, nz, ny, nx, 3)
    (C order), snapshot k the grid at times[k]."""
    assert snaps.ndim == 5 and snaps.shape[4] == 3
    assert snaps.dtype in (np.float64, np.float32)
    snaps = np.ascontiguousarray(snaps)
    tm = np.ascontiguousarray(times, dtype=np.float64)
    assert tm.shape == (snaps.shape[0],) and np.all(np.diff(tm) > 0)
    _, nz, ny, nx, _ = snaps.shape
    st = OrcField(kind=FIELD_TIME_GRID, extrap=extrap, dtype=0 if snaps.dtype == np.float64 else 1,
                  data=snaps.ctypes.data, nx=nx, ny=ny, nz=nz, times=_dp(tm), nsnap=len(tm))
    st.origin[:] = [float(a) for a in origin]
    st.spacing[:] = [float(a) for a in spacing]
    return Field(st, keep=(snaps, tm))


def mesh(data: np.ndarray, axes, coord: int, order: int, extrap) -> Field:
    """General mesh (rectilinear / cubic / spherical / 2D / 1D).  data:
    (nz, ny, nx, 3) C order; axes: the three node-coordinate vectors (length
    1 for an ignored axis); extrap: boundary per axis."""
    assert data.ndim == 4 and data.shape[3] == 3
    assert data.dtype in (np.float64, np.float32)
    data = np.ascontiguousarray(data)
    nz, ny, nx, _ = data.shape
    ax = [np.ascontiguousarray(a, dtype=np.float64) for a in axes]
    assert [len(a) for a in ax] == [nx, ny, nz]
    st = OrcField(kind=FIELD_MESH, dtype=0 if data.dtype == np.float64 else 1,
                  data=data.ctypes.data, nx=nx, ny=ny, nz=nz, coord=coord, order=order)
    for d in range(3):
        st.axes[d] = _dp(ax[d])
        st.mesh_extrap[d] = int(extrap[d])
    return Field(st, keep=(data, ax))


def uxb(u, B: Field) -> Field:
    st = OrcField(kind=FIELD_UXB)
    st.v[:] = [float(a) for a in u]
    st.src = C.pointer(B.st)
    return Field(st, keep=(B,))


def kat(kind: int, v=(0.0, 0.0, 0.0), radius: float = 0.0, t_on: float = 0.0) -> Field:
    st = OrcField(kind=kind, radius=radius, t_on=t_on)
    st.v[:] = [float(a) for a in v]
    return Field(st)


def field_eval(f: Field, x, t: float = 0.0) -> np.ndarray:
    xa = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_field_eval(C.byref(f.st), _dp(xa), float(t), _dp(out))
    return out


# ---------------------------------------------------------------------------
# Math kernels
# ---------------------------------------------------------------------------
def boris_velocity_update(v, E, B, h) -> np.ndarray:
    va, Ea, Ba = (np.ascontiguousarray(a, dtype=np.float64) for a in (v, E, B))
    out = np.zeros(3)
    lib().orc_boris_velocity_update(_dp(va), _dp(Ea), _dp(Ba), float(h), _dp(out))
    return out


def multistep_velocity(v, E, B, q2m, dt, n, N, form=FORM_REFERENCE) -> np.ndarray:
    va, Ea, Ba = (np.ascontiguousarray(a, dtype=np.float64) for a in (v, E, B))
    out = np.zeros(3)
    lib().orc_multistep_velocity(_dp(va), _dp(Ea), _dp(Ba), float(q2m), float(dt), int(n),
                                 int(N), int(form), _dp(out))
    return out


def multistep_coeffs(tau2: float, n: int, form: int) -> np.ndarray:
    c = np.zeros(4)
    lib().orc_multistep_coeffs(float(tau2), int(n), int(form), _dp(c))
    return c


def rational_polys(n: int):
    arrs = [np.zeros(16) for _ in range(5)]
    deg = C.c_int(0)
    rc = lib().orc_rational_polys(int(n), *[_dp(a) for a in arrs], C.byref(deg))
    if rc:
        raise ValueError(f"rational form supports 1 <= n <= 12, got {n}")
    d = deg.value
    D, P1, P2, P3, P6 = arrs
    return D[:d + 1], P1[:d + 1], P2[:d], P3[:d], P6[:d]


def calculate_nout(nt, K, save_start=True, save_end=True, save_everystep=True) -> int:
    return int(lib().orc_calculate_nout(int(nt), int(K), int(save_start), int(save_end),
                                        int(save_everystep)))


# ---------------------------------------------------------------------------
# Solve
# ---------------------------------------------------------------------------
class OracleResult:
    def __init__(self, u, t, length, retcode, elapsed, steps=None):
        self.u = u            # (N, cap, 6)
        self.t = t            # (N, cap)
        self.len = length     # (N,)
        self.retcode = retcode
        self.elapsed = elapsed
        self.steps = steps    # (N,) Boris steps advanced per particle

    def traj(self, i):
        n = min(int(self.len[i]), self.u.shape[1])
        return self.t[i, :n], self.u[i, :n]


def solve(xv0, E: Field, B: Field, *, q2m, tspan, dt=0.0, alg=ALG_BORIS, adaptive=False,
          order_N=2, nsub=1, form=FORM_REFERENCE, safety=0.1, savestepinterval=1,
          save_start=True, save_end=True, save_everystep=True, maxiters=1_000_000,
          max_steps=0, nout_cap=None, bnd_flags=0, r_in=0.0, r_out=0.0,
          box_lo=(-np.inf,) * 3, box_hi=(np.inf,) * 3, dt_pp=None, nsteps_pp=0,
          nthreads=1, save_fields=False, save_work=False, mass=0.0) -> OracleResult:
    xv0 = np.ascontiguousarray(np.atleast_2d(xv0), dtype=np.float64)
    N = xv0.shape[0]
    d = OrcDesc(alg=alg, adaptive=int(adaptive), order_N=order_N, nsub=nsub, form=form,
                save_start=int(save_start), save_end=int(save_end),
                save_everystep=int(save_everystep), bnd_flags=bnd_flags, safety=safety,
                q2m=q2m, t0=tspan[0], t1=tspan[1], dt=dt, maxiters=maxiters,
                savestepinterval=savestepinterval, max_steps=max_steps, r_in=r_in,
                r_out=r_out, nsteps_pp=nsteps_pp, save_fields=int(save_fields),
                save_work=int(save_work), mass=mass)
    width = 6 + (6 if save_fields else 0) + (4 if save_work else 0)
    d.box_lo[:] = [float(a) for a in box_lo]
    d.box_hi[:] = [float(a) for a in box_hi]
    dtpp = None
    if dt_pp is not None:
        dtpp = np.ascontiguousarray(dt_pp, dtype=np.float64)
        d.dt_pp = _dp(dtpp)
    L = lib()
    if nout_cap is None:
        if adaptive:
            raise ValueError("adaptive solve needs nout_cap")
        nt, nout = C.c_long(0), C.c_long(0)
        rc = L.orc_prepare_checks(C.byref(d), C.byref(nt), C.byref(nout))
        if rc:
            raise ValueError(ERR_NAMES.get(rc, str(rc)))
        nout_cap = max(int(nout.value), 1)
    d.nout_cap = nout_cap
    out_u = np.full((N, nout_cap, width), np.nan)
    out_t = np.full((N, nout_cap), np.nan)
    ln = np.zeros(N, dtype=np.int32)
    rc_arr = np.zeros(N, dtype=np.int32)
    el = np.zeros(1)
    steps = np.zeros(N, dtype=np.int64)
    d.steps_out = steps.ctypes.data_as(C.POINTER(C.c_long))
    rc = L.orc_solve(C.byref(d), C.byref(E.st), C.byref(B.st), _dp(xv0), N, _dp(out_u),
                     _dp(out_t), ln.ctypes.data_as(C.POINTER(C.c_int)),
                     rc_arr.ctypes.data_as(C.POINTER(C.c_int)), int(nthreads), _dp(el))
    if rc:
        raise ValueError(ERR_NAMES.get(rc, str(rc)))
    return OracleResult(out_u, out_t, ln, rc_arr, float(el[0]), steps)


def field_jac(B: Field, x, t=0.0):
    """B, its Jacobian J[i, j] = dB_i/dx_j and dB/dt at (x, t)."""
    xa = np.ascontiguousarray(x, dtype=np.float64)
    Bv, J, dBdt = np.zeros(3), np.zeros(9), np.zeros(3)
    lib().orc_field_jac(C.byref(B.st), _dp(xa), float(t), _dp(Bv), _dp(J), _dp(dBdt))
    return Bv, J.reshape(3, 3), dBdt


def work_rates(q2m, m, v, E, B, J, dBdt):
    """(P_par, P_fermi, P_grad, P_betatron) from the fields at one state."""
    arr = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (v, E, B, J, dBdt)]
    out = np.zeros(4)
    lib().orc_work_rates(float(q2m), float(m), *[_dp(a) for a in arr], _dp(out))
    return out


def record_eval(E: Field, B: Field, *, q2m, m, xv, t, want_fields=True, want_work=True):
    """The extra record columns at states xv (n, 6) and times t (n,)."""
    xv = np.ascontiguousarray(np.atleast_2d(xv), dtype=np.float64)
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, np.float64), (xv.shape[0],)))
    w = (6 if want_fields else 0) + (4 if want_work else 0)
    out = np.zeros((xv.shape[0], w))
    lib().orc_record_eval(C.byref(E.st), C.byref(B.st), float(q2m), float(m), _dp(xv), _dp(t),
                          xv.shape[0], int(want_fields), int(want_work), _dp(out))
    return out


def boris_tangent(xv0, B: Field, *, q2m, tspan, dt):
    """Final position and its Jacobian d x(T) / d (x0, v0) (3 x 6) of a
    fixed-step standard Boris solve with E = 0, by forward-mode duals."""
    d = OrcDesc(alg=ALG_BORIS, q2m=q2m, t0=tspan[0], t1=tspan[1], dt=dt, maxiters=1 << 40,
                savestepinterval=1, save_start=0, save_end=1, save_everystep=0)
    xa = np.ascontiguousarray(xv0, dtype=np.float64)
    xT = np.zeros(3)
    J = np.zeros(18)
    rc = lib().orc_boris_tangent(C.byref(d), C.byref(B.st), _dp(xa), _dp(xT), _dp(J))
    if rc:
        raise ValueError(ERR_NAMES.get(rc, str(rc)))
    return xT, J.reshape(3, 6)


TANGENT_MAX = 32


def loss_tangent(xv0, targets, B: Field, *, q2m, tspan, dt, dense=None, flat=None, nthreads=None):
    """Loss L = mean_i |x_i(T) - x*_i|^2 of a fixed-step standard Boris solve
    (E = 0, periodic gridded B) and its forward-mode directional derivatives
    along grid-value directions: dense (ndir, 3 nx ny nz) arrays in the node
    layout (component fastest), or flat (ndir,) single values 3 node + c.
    Returns (L, dL (ndir,))."""
    d = OrcDesc(alg=ALG_BORIS, q2m=q2m, t0=tspan[0], t1=tspan[1], dt=dt, maxiters=1 << 40,
                savestepinterval=1, save_start=0, save_end=1, save_everystep=0)
    xa = np.ascontiguousarray(xv0, dtype=np.float64)
    tg = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, np.float64), (len(xa), 3)))
    if dense is not None:
        da = np.ascontiguousarray(dense, dtype=np.float64)
        nd = da.shape[0]
        dp, fp = _dp(da), None
    else:
        fa = np.ascontiguousarray(flat, dtype=np.int64)
        nd = fa.shape[0]
        dp, fp = None, fa.ctypes.data_as(C.POINTER(C.c_long))
    if not 1 <= nd <= TANGENT_MAX:
        raise ValueError(f"1 to {TANGENT_MAX} directions")
    nth = nthreads or max(1, min(16, len(os.sched_getaffinity(0))))
    loss = np.zeros(1)
    dl = np.zeros(nd)
    rc = lib().orc_loss_tangent(C.byref(d), C.byref(B.st), len(xa), _dp(xa), _dp(tg), nd, dp, fp,
                                nth, _dp(loss), _dp(dl))
    if rc:
        raise ValueError(ERR_NAMES.get(rc, str(rc)))
    return float(loss[0]), dl


def fdlibm_atan(x: float) -> float:
    """atan of the FDLIBM restatement shared with the device library."""
    return float(lib().orc_fdlibm_atan(float(x)))


def fdlibm_sincos(x: float):
    """(sin x, cos x) of the FDLIBM restatement shared with the device library."""
    s, c = C.c_double(), C.c_double()
    lib().orc_fdlibm_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def fo_segment(E: Field, B: Field, state, iters, flags, *, q2m, t_end, safety_fo, dtmin, dtmax,
               check_interval, maxiters=10000, savestepinterval=1, save_everystep=True, cap=16,
               bnd_flags=0, r_in=0.0, r_out=0.0, box_lo=(-np.inf,) * 3, box_hi=(np.inf,) * 3):
    """The hybrid solver's full-orbit branch from each particle's state to its
    next check point, t_end or isoutside (src/hybrid.jl:257-452); state (N, 8),
    iters (N, 2) int64 and flags (N,) int32 are updated in place.  Returns
    (sync (N, 6), records (N, cap, 6), times (N, cap), lengths (N,))."""
    d = OrcDesc(alg=ALG_BORIS, q2m=q2m, t1=t_end, safety=safety_fo, maxiters=maxiters,
                savestepinterval=savestepinterval, save_everystep=int(save_everystep),
                nout_cap=cap, bnd_flags=bnd_flags, r_in=r_in, r_out=r_out)
    d.box_lo[:] = [float(a) for a in box_lo]
    d.box_hi[:] = [float(a) for a in box_hi]
    N = state.shape[0]
    sync = np.zeros((N, 6))
    ou = np.zeros((N, cap, 6))
    ot = np.zeros((N, cap))
    ln = np.zeros(N, dtype=np.int32)
    rc = lib().orc_fo_segment(C.byref(d), C.byref(E.st), C.byref(B.st), N, dtmin, dtmax,
                              check_interval, _dp(state), iters.ctypes.data_as(C.POINTER(C.c_long)),
                              flags.ctypes.data_as(C.POINTER(C.c_int)), _dp(sync), _dp(ou), _dp(ot),
                              ln.ctypes.data_as(C.POINTER(C.c_int)))
    if rc:
        raise ValueError(ERR_NAMES.get(rc, str(rc)))
    return sync, ou, ot, ln

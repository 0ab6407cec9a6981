"""TestParticle.solve for the Boris family on the MI355X.

Mirrors
  solve(prob, alg::Union{Boris{false}, MultistepBoris{N,false}}, [ensemblealg];
        trajectories, savestepinterval, dt, isoutside, save_start, save_end,
        save_everystep, save_fields, save_work, maxiters, batch_size, seed)
        (src/boris/boris.jl:88-114),
  solve(prob, alg::Union{Boris{true}, MultistepBoris{N,true}}, [ensemblealg]; ...)
        (src/boris/adaptive_boris.jl:21-38), and
  solve(prob, ::Boris{false}, backend, [ensemblealg]; ..., workgroup_size)
        (src/boris/boris_kernel.jl:312-382).
Every ensemble algorithm runs the whole ensemble as one launch of the HIP
kernels through the C ABI (include/tpamd.h).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .algorithms import AbstractBoris, Boris, EnsembleAlgorithm, MI355X, MultistepBoris
from .fields import AbstractField, UxBField
from .problem import Boundary, EnsembleContext, TraceProblem, DEFAULT_PROB_FUNC
from .solution import EnsembleSolution

_PRECISIONS = {"fp64": _lib.FP64, "fp32c": _lib.FP32C}
_FORMS = {"reference": _lib.FORM_REFERENCE, "rational": _lib.FORM_RATIONAL}


def _split_positional(args):
    backend, ens = None, None
    for a in args:
        if isinstance(a, MI355X):
            backend = a
        elif isinstance(a, EnsembleAlgorithm) or (isinstance(a, type) and issubclass(a, EnsembleAlgorithm)):
            ens = a
        else:
            raise TypeError(f"unexpected positional argument {a!r}")
    return backend, ens


def initial_states(prob: TraceProblem, trajectories: int, seed=None) -> np.ndarray:
    """(N, 6) initial states: prob_func(prob, EnsembleContext(i)) for i = 1..N,
    evaluated on the host as the reference does (boris.jl:501-503)."""
    u0 = np.asarray(prob.u0, dtype=np.float64)
    if u0.ndim == 2:
        if u0.shape[1] != 6:
            raise _lib.ArgumentError("u0 rows must be (x, y, z, vx, vy, vz)")
        if trajectories not in (1, u0.shape[0]):
            raise _lib.ArgumentError("trajectories must match the rows of u0")
        return np.ascontiguousarray(u0)
    if u0.shape != (6,):
        raise _lib.ArgumentError("u0 must be the 6-vector (x, y, z, vx, vy, vz)")
    if prob.prob_func is DEFAULT_PROB_FUNC or prob.prob_func is None:
        return np.ascontiguousarray(np.tile(u0, (trajectories, 1)))
    out = np.empty((trajectories, 6))
    for i in range(1, trajectories + 1):
        rng = None if seed is None else np.random.default_rng([int(seed), i])
        newp = prob.prob_func(prob, EnsembleContext(sim_id=i, rng=rng, seed=seed))
        out[i - 1] = np.asarray(newp.u0 if isinstance(newp, TraceProblem) else newp, np.float64)
    return out


def make_desc(prob: TraceProblem, alg: AbstractBoris, *, dt=None, savestepinterval=1,
              isoutside=None, save_start=True, save_end=True, save_everystep=True,
              maxiters=1_000_000, precision="fp64", form="reference", max_steps=0,
              nout_cap=0, dts=None, nsteps=0, workgroup_size=0,
              kernels="auto", save_fields=False, save_work=False,
              particles_per_lane=0) -> _lib.SolveDesc:
    q2m = float(prob.p[0])
    d = _lib.SolveDesc()
    # records gain E, B and/or the work rates (src/boris/boris.jl:405-440);
    # the work rates need the mass m = p[2]
    d.save_fields, d.save_work = int(bool(save_fields)), int(bool(save_work))
    d.mass = float(prob.p[1]) if len(prob.p) > 1 else 0.0
    if isinstance(alg, MultistepBoris):
        d.alg, d.order_N, d.nsub = _lib.ALG_MULTISTEP, alg.N, alg.n
    elif isinstance(alg, Boris):
        d.alg, d.order_N, d.nsub = _lib.ALG_BORIS, 2, 1
    else:
        raise TypeError(f"unsupported algorithm {alg!r}")
    d.adaptive = int(bool(alg.adaptive))
    d.safety = float(alg.safety)
    if form not in _FORMS:
        raise _lib.ArgumentError(f"form must be one of {sorted(_FORMS)}")
    if precision not in _PRECISIONS:
        raise _lib.ArgumentError(f"precision must be one of {sorted(_PRECISIONS)}")
    d.form = _FORMS[form]
    d.precision = _PRECISIONS[precision]
    d.save_start, d.save_end, d.save_everystep = int(save_start), int(save_end), int(save_everystep)
    d.q2m = q2m
    d.t0, d.t1 = float(prob.tspan[0]), float(prob.tspan[1])
    if not alg.adaptive:
        if dts is None:
            if dt is None:
                raise TypeError("solve() missing required keyword argument: dt")
            d.dt = float(dt)
        else:
            d.nsteps_pp = int(nsteps)
    d.maxiters = int(maxiters)
    d.savestepinterval = int(savestepinterval)
    d.max_steps = int(max_steps)
    d.nout_cap = int(nout_cap)
    d.block_size = int(workgroup_size or 0)
    d.particles_per_lane = int(particles_per_lane or 0)
    if kernels not in ("auto", "generic"):
        raise _lib.ArgumentError("kernels must be 'auto' or 'generic'")
    d.flags = _lib.FLAG_GENERIC_KERNELS if kernels == "generic" else 0
    b = isoutside.condition if hasattr(isoutside, "condition") else isoutside
    if b is None:
        b = Boundary()
    if not isinstance(b, Boundary):
        raise TypeError("isoutside must come from the device menu (OutsideNaN, InsideSphere, "
                        "OutsideSphere, OutsideBox, combined with |): a Python callable cannot "
                        "run on the GPU")
    d.bnd_flags = b.flags
    d.r_in, d.r_out = b.r_in, b.r_out
    d.box_lo[:] = list(b.box_lo)
    d.box_hi[:] = list(b.box_hi)
    return d


def plan(desc: _lib.SolveDesc):
    """(nt, nout) of a fixed-step solve: _prepare_checks + _calculate_nout.
    Pure host logic in the library: needs no device."""
    L = _lib.load_library()
    nt, nout = C.c_int64(), C.c_int64()
    rc = L.tpamd_plan(None, C.byref(desc), C.byref(nt), C.byref(nout))
    if rc:
        msg = L.tpamd_last_error(None).decode()
        if rc in (_lib.ERR_ARGUMENT, _lib.ERR_DT_TOO_SMALL, _lib.ERR_MAXITERS, _lib.ERR_ORDER):
            raise _lib.ArgumentError(msg)
        if rc == _lib.ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        raise _lib.DeviceError(msg)
    return int(nt.value), int(nout.value)


def _fields(prob):
    p = prob.p
    if len(p) < 4:
        raise _lib.ArgumentError("p must be (q2m, m, E, B, F)")
    E, B = p[2], p[3]
    for f in (E, B):
        if not isinstance(f, AbstractField):
            raise TypeError("p[3] and p[4] must be device fields (see testparticle_amd.fields); "
                            "a Python callable cannot run on the GPU")
    if isinstance(B, UxBField):
        raise _lib.ArgumentError("B cannot be a UxB field")
    return E, B


def solve(prob: TraceProblem, alg: AbstractBoris, *args, trajectories: int = 1,
          savestepinterval: int = 1, dt=None, isoutside=None, save_start: bool = True,
          save_end: bool = True, save_everystep: bool = True, save_fields: bool = False,
          save_work: bool = False, maxiters: int = 1_000_000, batch_size=None, seed=None,
          workgroup_size: int = 0, precision: str = "fp64", form: str = "reference",
          max_steps: int = 0, nout_cap=None, dts=None, nsteps: int = 0,
          kernels: str = "auto", particles_per_lane: int = 0, device=None) -> EnsembleSolution:
    """Trace an ensemble with the Boris family on the GPU.

    Extensions beyond the reference's keywords: `precision` ("fp64" |
    "fp32c"), `form` (multistep coefficients: "reference", the default, as
    multistep_boris.jl:94-120 computes them, with atan and sin/cos from
    FDLIBM, bit-identical to the oracle; "rational", the same coefficients as
    ratios of integer polynomials in tau^2, no transcendental),
    `max_steps` (adaptive step cap), `dts` + `nsteps` (per-particle step,
    fixed-step only), `nout_cap` (adaptive record capacity), `kernels`
    ("auto" | "generic": skip the specialised kernels for gridded B, uniform
    fields and the dipole),
    `particles_per_lane` (adaptive only: each wave of 64 lanes works through
    64 x this many particles, a lane taking the next when its own ends; 0 =
    the library's choice from N)."""
    backend, _ens = _split_positional(args)
    if alg.adaptive and dt is not None:
        raise TypeError("the adaptive solver takes no dt (src/boris/adaptive_boris.jl:21-33)")
    dev = backend.device if backend is not None else (0 if device is None else int(device))
    E, B = _fields(prob)
    xv0 = initial_states(prob, trajectories, seed)
    N = xv0.shape[0]
    dts_arr = None
    if dts is not None:
        dts_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, np.float64), (N,)))
    desc = make_desc(prob, alg, dt=dt, savestepinterval=savestepinterval, isoutside=isoutside,
                     save_start=save_start, save_end=save_end, save_everystep=save_everystep,
                     maxiters=maxiters, precision=precision, form=form, max_steps=max_steps,
                     dts=dts_arr, nsteps=nsteps, workgroup_size=workgroup_size,
                     kernels=kernels, save_fields=save_fields, save_work=save_work,
                     particles_per_lane=particles_per_lane)
    if dts_arr is not None:
        desc.dt_pp = dts_arr.ctypes.data
    if not alg.adaptive:
        _, nout = plan(desc)
        cap = nout
    elif nout_cap is not None:
        cap = int(nout_cap)
    elif not save_everystep:
        cap = int(save_start) + 1
    elif max_steps > 0:
        cap = int(save_start) + max_steps // savestepinterval + 2
    else:
        cap = 1024
    ctx = _lib.context(dev)
    hE, hB = E.handle(ctx), B.handle(ctx)
    while True:
        desc.nout_cap = cap
        u = np.empty((N, max(cap, 0), desc.record_width))
        t = np.empty((N, max(cap, 0)))
        ln = np.zeros(N, dtype=np.int32)
        rc = np.zeros(N, dtype=np.int32)
        st = _lib.Stats()
        with ctx.lock:
            code = ctx.L.tpamd_solve(ctx.h, C.byref(desc), hE, hB, xv0.ctypes.data, N,
                                     u.ctypes.data, t.ctypes.data, ln.ctypes.data,
                                     rc.ctypes.data, C.byref(st))
        if code == _lib.ERR_CAPACITY and nout_cap is None:
            cap = int(st.max_len)        # deterministic: run again with room for every record
            continue
        ctx.check(code, st)
        break
    return EnsembleSolution(t, u, ln, rc, st.kernel_ms * 1e-3, stats=st, prob=prob, alg=alg)


def _adaptive_cap(save_start, save_everystep, max_steps, savestepinterval):
    if not save_everystep:
        return int(save_start) + 1
    if max_steps > 0:
        return int(save_start) + max_steps // savestepinterval + 2
    return 1024


def solve_batch(problems, alg: AbstractBoris, *, dt=None, savestepinterval: int = 1,
                isoutside=None, save_start: bool = True, save_end: bool = True,
                save_everystep: bool = True, maxiters: int = 1_000_000,
                precision: str = "fp64", form: str = "reference", max_steps: int = 0,
                nout_cap=None, kernels: str = "auto", device: int = 0,
                save_fields: bool = False, save_work: bool = False):
    """Solve many problems, each with its own fields and (N_k, 6) initial
    states, in one persistent launch (tpamd_solve_batch).  All problems share
    q/m, tspan and the solver keywords.  Returns one EnsembleSolution per
    problem.  The reference has no batched entry point: it would call solve
    once per field (src/boris/boris.jl:88-114)."""
    if not problems:
        return []
    q2m, tspan = float(problems[0].p[0]), tuple(problems[0].tspan)
    for pr in problems:
        if float(pr.p[0]) != q2m or tuple(pr.tspan) != tspan:
            raise _lib.ArgumentError("batched problems must share q/m and tspan")
    if alg.adaptive and dt is not None:
        raise TypeError("the adaptive solver takes no dt (src/boris/adaptive_boris.jl:21-33)")
    fields = [_fields(pr) for pr in problems]
    states = [initial_states(pr, 1) for pr in problems]
    counts = [s.shape[0] for s in states]
    offsets = np.zeros(len(problems) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    xv0 = np.ascontiguousarray(np.concatenate(states) if offsets[-1] else np.zeros((0, 6)))
    N = int(offsets[-1])
    desc = make_desc(problems[0], alg, dt=dt, savestepinterval=savestepinterval,
                     isoutside=isoutside, save_start=save_start, save_end=save_end,
                     save_everystep=save_everystep, maxiters=maxiters, precision=precision,
                     form=form, max_steps=max_steps, kernels=kernels, save_fields=save_fields,
                     save_work=save_work)
    if not alg.adaptive:
        cap = plan(desc)[1]
    elif nout_cap is not None:
        cap = int(nout_cap)
    else:
        cap = _adaptive_cap(save_start, save_everystep, max_steps, savestepinterval)
    ctx = _lib.context(device)
    hE = (C.c_void_p * len(problems))(*[f[0].handle(ctx).value for f in fields])
    hB = (C.c_void_p * len(problems))(*[f[1].handle(ctx).value for f in fields])
    while True:
        desc.nout_cap = cap
        u = np.empty((N, max(cap, 0), desc.record_width))
        t = np.empty((N, max(cap, 0)))
        ln = np.zeros(N, dtype=np.int32)
        rc = np.zeros(N, dtype=np.int32)
        st = _lib.Stats()
        with ctx.lock:
            code = ctx.L.tpamd_solve_batch(ctx.h, C.byref(desc), len(problems), hE, hB,
                                           offsets.ctypes.data_as(C.POINTER(C.c_int64)),
                                           xv0.ctypes.data, u.ctypes.data, t.ctypes.data,
                                           ln.ctypes.data, rc.ctypes.data, C.byref(st))
        if code == _lib.ERR_CAPACITY and nout_cap is None:
            cap = int(st.max_len)
            continue
        ctx.check(code, st)
        break
    out = []
    for k, pr in enumerate(problems):
        a, b = offsets[k], offsets[k + 1]
        out.append(EnsembleSolution(t[a:b], u[a:b], ln[a:b], rc[a:b], st.kernel_ms * 1e-3,
                                    stats=st, prob=pr, alg=alg))
    return out

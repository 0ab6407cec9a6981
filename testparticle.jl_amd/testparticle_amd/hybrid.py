"""AdaptiveHybrid's full-orbit segments on the device (tpamd_fo_segment).

The reference's hybrid solver, _hybrid_adaptive! (src/hybrid.jl:257-452),
alternates between a guiding-centre RK45 branch and a full-orbit Boris branch,
deciding at every `check_interval`-th iteration from the adiabaticity of the
field.  The full-orbit branch is the same leapfrog as the Boris pusher: here it
runs on the GPU, particle by particle, from one check point to the next, and
the host makes every decision:

    res = full_orbit(prob, AdaptiveHybrid(dtmax=..., check_interval=100),
                     stay_full_orbit=lambda i, t, xv: ...)

`stay_full_orbit(i, t, xv_sync)` is asked at the initial decision (:289-293)
and at every check point (:359-379) with the synchronised state; returning
False hands the particle to the caller's guiding-centre branch (its state at
that point is in `res.handoff`).  Without it every particle stays in full
orbit, as AdaptiveHybrid with threshold = 0 does (test/test_hybrid.jl:124-152).
The guiding-centre branch itself (RK45 on the guiding-centre equations) is not
part of this package.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .algorithms import Boris
from .problem import DEFAULT_PROB_FUNC, TraceProblem
from .solution import EnsembleSolution
from .solve import _fields, initial_states, make_desc


@dataclass(frozen=True)
class AdaptiveHybrid:
    """AdaptiveHybrid(; threshold, threshold_gc_to_fo, threshold_fo_to_gc,
    dtmax, dtmin, safety_fo, abstol, reltol, maxiters, check_interval,
    adiabaticity, save_adiabaticity) with the reference's defaults and
    checks (src/hybrid.jl:39-87)."""
    threshold_gc_to_fo: float
    threshold_fo_to_gc: float
    dtmin: float
    dtmax: float
    safety_fo: float = 0.1
    abstol: float = 1.0e-6
    reltol: float = 1.0e-6
    maxiters: int = 10000
    check_interval: int = 10
    adiabaticity: str = "curvature"
    save_adiabaticity: bool = True

    @classmethod
    def make(cls, *, dtmax, threshold=0.1, threshold_gc_to_fo=None, threshold_fo_to_gc=None,
             dtmin=None, safety_fo=0.1, abstol=1.0e-6, reltol=1.0e-6, maxiters=10000,
             check_interval=10, adiabaticity="curvature", save_adiabaticity=True):
        a = threshold if threshold_gc_to_fo is None else threshold_gc_to_fo
        b = threshold if threshold_fo_to_gc is None else threshold_fo_to_gc
        if check_interval <= 0:
            raise _lib.ArgumentError("check_interval must be positive.")
        if not a >= b:
            raise _lib.ArgumentError("threshold_gc_to_fo must be >= threshold_fo_to_gc for hysteresis.")
        if adiabaticity not in ("curvature", "gradB", "both", "jacobian"):
            raise _lib.ArgumentError("adiabaticity must be one of :curvature, :gradB, :both, :jacobian.")
        return cls(float(a), float(b), float(1.0e-2 * dtmax if dtmin is None else dtmin), float(dtmax),
                   float(safety_fo), float(abstol), float(reltol), int(maxiters), int(check_interval),
                   adiabaticity, bool(save_adiabaticity))


def hybrid(**kw) -> AdaptiveHybrid:
    """AdaptiveHybrid(; ...) keyword constructor (src/hybrid.jl:56-87)."""
    return AdaptiveHybrid.make(**kw)


@dataclass
class TraceHybridProblem:
    """TraceHybridProblem(u0, tspan, p; prob_func) (src/hybrid.jl:89-120):
    u0 is the full-orbit state, p = (q2m, m, E, B, F)."""
    u0: object
    tspan: tuple
    p: tuple
    prob_func: object = DEFAULT_PROB_FUNC


@dataclass
class FullOrbitResult:
    """The full-orbit trajectories (an EnsembleSolution with the reference's
    record layout) and, per particle, where it left full orbit:
    `handoff[i]` = (t, xv_sync) for a particle the caller took over for its
    guiding-centre branch, None otherwise."""
    sol: EnsembleSolution
    handoff: list = field(default_factory=list)
    segments: int = 0


def fo_segment(prob, alg: AdaptiveHybrid, state, iters, flags, *, savestepinterval=1,
               isoutside=None, save_everystep=True, cap=None, device: int = 0):
    """One call of tpamd_fo_segment on host arrays, updated in place:
    state (N, 8) x, v, t, dt; iters (N, 2) it, steps; flags (N,) in:
    _lib.FO_ENTER | _lib.FO_CHECKED, out: _lib.FO_CHECK / FO_END / FO_OUTSIDE.
    Returns (sync (N, 6), records (N, cap, 6), times (N, cap), lengths (N,))."""
    E, B = _fields(prob)
    N = state.shape[0]
    K = int(savestepinterval)
    if cap is None:
        # a call runs at most check_interval steps: one record every K
        cap = alg.check_interval // K + 2
    tp = TraceProblem(np.zeros(6), (0.0, float(prob.tspan[1])), prob.p)
    d = make_desc(tp, Boris(), dt=1.0, savestepinterval=K, isoutside=isoutside,
                  save_everystep=save_everystep, maxiters=alg.maxiters, nout_cap=cap)
    d.safety = alg.safety_fo
    for a, dt_ in ((state, np.float64), (iters, np.int64), (flags, np.int32)):
        if a.dtype != dt_ or not a.flags["C_CONTIGUOUS"]:
            raise TypeError("state, iters and flags must be C-contiguous float64, int64, int32 arrays")
    sync = np.zeros((N, 6))
    out_u = np.zeros((N, cap, 6))
    out_t = np.zeros((N, cap))
    ln = np.zeros(N, dtype=np.int32)
    ctx = _lib.context(device)
    st = _lib.Stats()
    with ctx.lock:
        hE, hB = E.handle(ctx), B.handle(ctx)
        rc = ctx.L.tpamd_fo_segment(ctx.h, C.byref(d), hE, hB, N, alg.dtmin, alg.dtmax,
                                    alg.check_interval, _lib.dptr(state), _lib.iptr64(iters),
                                    _lib.iptr(flags), _lib.dptr(sync), _lib.dptr(out_u),
                                    _lib.dptr(out_t), _lib.iptr(ln), C.byref(st))
    ctx.check(rc, st)
    return sync, out_u, out_t, ln


def full_orbit(prob: TraceHybridProblem, alg: AdaptiveHybrid, *, trajectories: int = 1,
               savestepinterval: int = 1, isoutside=None, save_start: bool = True,
               save_end: bool = True, save_everystep: bool = True, seed=None,
               stay_full_orbit=None, device: int = 0) -> FullOrbitResult:
    """The full-orbit branch of solve(prob::TraceHybridProblem, alg::AdaptiveHybrid;
    trajectories, savestepinterval, isoutside, save_start, save_end,
    save_everystep) (src/hybrid.jl:123-138,257-470) on the device, the mode
    decisions on the host (see the module docstring)."""
    xv0 = initial_states(TraceProblem(prob.u0, prob.tspan, prob.p, prob.prob_func), trajectories, seed)
    N = xv0.shape[0]
    t0, t_end = float(prob.tspan[0]), float(prob.tspan[1])
    recs = [[] for _ in range(N)]
    times = [[] for _ in range(N)]
    handoff = [None] * N
    if save_start:
        for i in range(N):
            recs[i].append(xv0[i].copy())
            times[i].append(t0)
    state = np.zeros((N, 8))
    state[:, :6] = xv0
    state[:, 6] = t0
    iters = np.zeros((N, 2), dtype=np.int64)
    iters[:, 0] = 1                                  # it = 1, steps = 0 (:334-335)
    flags = np.full(N, _lib.FO_ENTER, dtype=np.int32)
    # initial mode (:289-293): full orbit unless the caller starts it in GC
    active = np.ones(N, dtype=bool)
    if stay_full_orbit is not None:
        for i in range(N):
            if not stay_full_orbit(i, t0, xv0[i].copy()):
                active[i] = False
                handoff[i] = (t0, xv0[i].copy())
    done = ~active
    last = np.zeros((N, 6))
    segments = 0
    while not done.all():
        idx = np.flatnonzero(~done)
        st, it, fl = (np.ascontiguousarray(state[idx]), np.ascontiguousarray(iters[idx]),
                      np.ascontiguousarray(flags[idx]))
        sync, ou, ot, ln = fo_segment(prob, alg, st, it, fl, savestepinterval=savestepinterval,
                                      isoutside=isoutside, save_everystep=save_everystep,
                                      device=device)
        segments += 1
        state[idx], iters[idx], flags[idx] = st, it, fl
        for k, i in enumerate(idx):
            for r in range(int(ln[k])):
                recs[i].append(ou[k, r].copy())
                times[i].append(float(ot[k, r]))
            last[i] = sync[k]
            if fl[k] == _lib.FO_CHECK:
                # adiabaticity check at this it (:359-379): the caller decides
                if stay_full_orbit is None or stay_full_orbit(int(i), float(st[k, 6]), sync[k].copy()):
                    flags[i] = _lib.FO_CHECKED
                else:
                    handoff[i] = (float(st[k, 6]), sync[k].copy())
                    done[i] = True
            else:
                done[i] = True
    # closing record (:441-452): the synchronised state at the final t
    for i in range(N):
        if handoff[i] is not None:
            continue
        t = float(state[i, 6])
        if save_end and (not times[i] or times[i][-1] != t):
            recs[i].append(last[i].copy())
            times[i].append(t)
    cap = max(1, max(len(r) for r in recs))
    u_all = np.full((N, cap, 6), np.nan)
    t_all = np.full((N, cap), np.nan)
    lengths = np.zeros(N, dtype=np.int32)
    retc = np.zeros(N, dtype=np.int32)
    for i in range(N):
        n = len(recs[i])
        lengths[i] = n
        if n:
            u_all[i, :n] = np.array(recs[i])
            t_all[i, :n] = np.array(times[i])
        if iters[i, 1] >= alg.maxiters:
            retc[i] = _lib.RET_MAXITERS                  # :464
    sol = EnsembleSolution(t_all, u_all, lengths, retc, 0.0, prob=prob, alg=alg)
    return FullOrbitResult(sol, handoff, segments)


def fo_dt(alg: AdaptiveHybrid, q2m: float, Bmag: float) -> float:
    """_fo_dt (src/hybrid.jl:244-249): clamp(2 pi safety_fo / (|q2m| |B|), dtmin, dtmax)."""
    dt = (2.0 * math.pi) * alg.safety_fo / (abs(q2m) * Bmag)
    return alg.dtmax if dt > alg.dtmax else (alg.dtmin if dt < alg.dtmin else dt)

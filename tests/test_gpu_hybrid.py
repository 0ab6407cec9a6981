"""AdaptiveHybrid's full-orbit segments on the device (tpamd_fo_segment)
against the oracle's restatement of src/hybrid.jl:257-452, bit for bit, over
sequences of segments with host-made mode decisions; and the reference's
forced pure-full-orbit test (test/test_hybrid.jl:124-152) through the
full_orbit driver."""
import math

import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import oracle_field

import oracle as O

pytestmark = pytest.mark.gpu


def _lockstep(prob, alg, xv0, *, K, bnd=None, seed=0, enter_at=None):
    """Device and oracle full-orbit segments side by side on the same inputs;
    after every call all outputs must agree bit for bit, then the same
    scripted decisions (stay in full orbit with probability 0.9 at a check)
    apply to both."""
    N = len(xv0)
    Bo = oracle_field(prob.p[3])
    Eo = oracle_field(prob.p[2], Bo)
    state = np.zeros((N, 8))
    state[:, :6] = xv0
    state[:, 6] = prob.tspan[0] if enter_at is None else enter_at
    iters = np.zeros((N, 2), dtype=np.int64)
    iters[:, 0] = 1
    flags = np.full(N, TP._lib.FO_ENTER, dtype=np.int32)
    so, io, fo = state.copy(), iters.copy(), flags.copy()
    rng = np.random.default_rng(seed)
    b = bnd.condition if hasattr(bnd, "condition") else bnd
    kw = {}
    if b is not None:
        kw = dict(bnd_flags=b.flags, r_in=b.r_in, r_out=b.r_out, box_lo=b.box_lo, box_hi=b.box_hi)
    live = np.ones(N, dtype=bool)
    calls, statuses = 0, []
    cap = alg.check_interval // K + 2
    while live.any():
        idx = np.flatnonzero(live)
        st, it, fl = (np.ascontiguousarray(a[idx]) for a in (state, iters, flags))
        sd, ud, td, ld = TP.fo_segment(prob, alg, st, it, fl, savestepinterval=K, isoutside=bnd, cap=cap)
        st2, it2, fl2 = (np.ascontiguousarray(a[idx]) for a in (so, io, fo))
        sy, uo, to, lo = O.fo_segment(Eo, Bo, st2, it2, fl2, q2m=prob.p[0], t_end=prob.tspan[1],
                                      safety_fo=alg.safety_fo, dtmin=alg.dtmin, dtmax=alg.dtmax,
                                      check_interval=alg.check_interval, maxiters=alg.maxiters,
                                      savestepinterval=K, save_everystep=True, cap=cap, **kw)
        calls += 1
        for a, c in ((st, st2), (it, it2), (fl, fl2), (sd, sy), (ld, lo)):
            np.testing.assert_array_equal(a, c)
        for k in range(len(idx)):
            n = int(ld[k])
            np.testing.assert_array_equal(ud[k, :n], uo[k, :n])
            np.testing.assert_array_equal(td[k, :n], to[k, :n])
        state[idx], iters[idx] = st, it
        so[idx], io[idx] = st2, it2
        statuses += list(fl)
        for k, i in enumerate(idx):
            if fl[k] == TP._lib.FO_CHECK and rng.random() < 0.9:
                flags[i] = fo[i] = TP._lib.FO_CHECKED
            elif fl[k] == TP._lib.FO_CHECK and rng.random() < 0.5:
                # back from a guiding-centre stretch: enter full orbit again
                # from the synchronised state, as the GC -> FO switch does
                state[i, :6] = so[i, :6] = sd[k]
                flags[i] = fo[i] = TP._lib.FO_ENTER
            else:
                live[i] = False
    return calls, np.array(statuses), state, iters


def test_segments_turbulent_box_bitwise():
    # 64^3 periodic turbulent box, E = -u x B, FP64 storage; check every 10
    # iterations, a record every 3, re-entries after scripted GC stretches
    w, A = W.turbulent_box(N=512, n=64, nsteps=1, dtype=np.float64, seed=21)
    ax = W.grid_axes(64, w.extra["L"])
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC())
    E = TP.UxBField(w.extra["u"], B)
    Tg = 2 * math.pi / (abs(w.q2m) * 5e-9)
    prob = TP.TraceHybridProblem(w.xv0, (0.0, 300 * Tg / 40), (w.q2m, TP.mi, E, B, TP.ZeroField()))
    alg = TP.hybrid(dtmax=Tg / 40, dtmin=Tg / 4000, check_interval=10)
    calls, st, state, iters = _lockstep(prob, alg, w.xv0, K=3, seed=1)
    assert calls > 5 and np.sum(st == TP._lib.FO_CHECK) > 500
    assert np.sum(state[:, 6] == prob.tspan[1]) > 0      # some reach t_end exactly


def test_segments_dipole_with_boundaries_bitwise():
    # dipole belt, removal inside 1 R_E and outside 8 R_E, step cap
    w = W.dipole_belt(N=256, nsteps=1, seed=22)
    prob = TP.TraceHybridProblem(w.xv0, (0.0, 200.0),
                                 (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]),
                                  TP.ZeroField()))
    alg = TP.hybrid(dtmax=0.05, dtmin=1e-5, check_interval=7, maxiters=3000)
    bnd = TP.InsideSphere(TP.Re) | TP.OutsideSphere(8 * TP.Re)
    calls, st, state, iters = _lockstep(prob, alg, w.xv0, K=4, bnd=bnd, seed=2)
    assert calls > 5
    assert np.any(iters[:, 1] == 3000) or np.any(st == TP._lib.FO_OUTSIDE)


def test_forced_full_orbit_reproduces_boris():
    # test/test_hybrid.jl:124-152: threshold 0 keeps the hybrid in full orbit
    # for the whole trajectory; it must reproduce the standalone Boris solve
    # (the reference holds positions to 1e-8, velocities to 1e-5)
    q2m = TP.qi / TP.mi
    B0 = 0.01
    u0 = np.array([0.0, 0.0, 0.0, 5.0e4, 0.0, 1.0e5])
    Om = abs(q2m) * B0
    Tg = 2 * math.pi / Om
    tspan = (0.0, 30 * Tg)
    p = (q2m, TP.mi, TP.ZeroField(), TP.UniformField([0.0, 0.0, B0]), TP.ZeroField())
    alg = TP.hybrid(threshold=0.0, dtmax=10 * Tg, dtmin=1.0e-6 * Tg, check_interval=100)
    res = TP.full_orbit(TP.TraceHybridProblem(u0, tspan, p), alg)
    sol_h = res.sol.u[0]
    dt_fo = 2 * math.pi * alg.safety_fo / (abs(q2m) * B0)
    sol_b = TP.solve(TP.TraceProblem(u0, tspan, p), TP.Boris(), dt=dt_fo).u[0]
    assert len(sol_h.t) == len(sol_b.t)
    for i in range(len(sol_b.t)):
        uh = sol_h(sol_b.t[i])
        ub = sol_b.u[i]
        assert np.linalg.norm(uh[:3] - ub[:3]) <= 1.0e-8
        assert np.linalg.norm(uh[3:] - ub[3:]) <= 1.0e-5
    assert res.sol.retcodes[0] == 0 and all(h is None for h in res.handoff)


def test_full_orbit_driver_hands_particles_over():
    # the caller's decision hands a particle to its guiding-centre branch at
    # a check point: the driver stops it there and reports its state
    q2m = TP.qi / TP.mi
    p = (q2m, TP.mi, TP.ZeroField(), TP.UniformField([0.0, 0.0, 1e-8]), TP.ZeroField())
    u0 = np.array([[0.0, 0.0, 0.0, 1e5, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0, 1e5, 0.0]])
    Tg = 2 * math.pi / (q2m * 1e-8)
    alg = TP.hybrid(dtmax=Tg, check_interval=20)
    res = TP.full_orbit(TP.TraceHybridProblem(u0, (0.0, 50 * Tg), p), alg, trajectories=2,
                        stay_full_orbit=lambda i, t, xv: i == 0 or t < 5 * Tg)
    assert res.handoff[0] is None and res.handoff[1] is not None
    t1, xv1 = res.handoff[1]
    assert 5 * Tg <= t1 < 7 * Tg
    assert res.sol.t_all[0, res.sol.lengths[0] - 1] == 50 * Tg
    # |v| is conserved through the leapfrog to rounding in a uniform B
    assert abs(np.linalg.norm(xv1[3:]) / 1e5 - 1) < 1e-12

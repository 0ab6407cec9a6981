"""The oracle's restatement of the hybrid solver's full-orbit branch
(src/hybrid.jl:257-452, orc_fo_segment), pinned without the reference:
- in full orbit the hybrid is the Boris leapfrog: with no check points its
  saved states equal the oracle's Boris solve with the same dt bit for bit
  (the reference's own test/test_hybrid.jl:124-152 holds them to 1e-8);
- a check that keeps the particle in full orbit changes nothing: any
  check_interval gives the same trajectory bit for bit;
- the last step is cut at t_end, the step cap stops the particle, entering
  clamps dt to [dtmin, dtmax], and a particle stopped by isoutside keeps its
  position with the updated velocity (:418-424)."""
import math

import numpy as np
import pytest

QI, MI = 1.602176634e-19, 1.67262192369e-27
Q2M = QI / MI


def _run(orc, E, B, xv0, *, t0, t_end, ci, K=1, save_every=True, safety=0.1, dtmin=1e-30,
         dtmax=1e30, maxiters=10 ** 9, cap=4096, **bnd):
    N = len(xv0)
    state = np.zeros((N, 8))
    state[:, :6] = xv0
    state[:, 6] = t0
    iters = np.zeros((N, 2), dtype=np.int64)
    iters[:, 0] = 1
    flags = np.full(N, 1, dtype=np.int32)
    recs = [[] for _ in range(N)]
    times = [[] for _ in range(N)]
    calls = 0
    while True:
        act = np.flatnonzero(flags != 99)
        if act.size == 0:
            break
        st, it, fl = state[act].copy(), iters[act].copy(), flags[act].copy()
        sync, ou, ot, ln = orc.fo_segment(E, B, st, it, fl, q2m=Q2M, t_end=t_end, safety_fo=safety,
                                          dtmin=dtmin, dtmax=dtmax, check_interval=ci,
                                          maxiters=maxiters, savestepinterval=K,
                                          save_everystep=save_every, cap=cap, **bnd)
        calls += 1
        state[act], iters[act] = st, it
        for k, i in enumerate(act):
            recs[i] += [ou[k, r].copy() for r in range(ln[k])]
            times[i] += [ot[k, r] for r in range(ln[k])]
            flags[i] = 2 if fl[k] == 0 else 99          # stay in full orbit at every check
            if fl[k] != 0:
                recs[i].append(sync[k].copy())
                times[i].append(state[i, 6])
    return state, iters, recs, times, calls


def test_full_orbit_equals_boris_solve(orc):
    # uniform B along z plus a periodic grid perturbation; no check points
    rng = np.random.default_rng(5)
    data = np.zeros((6, 5, 4, 3))
    data[..., 2] = 1.0e-8
    data += rng.normal(scale=2e-9, size=data.shape)
    g = orc.grid(data, (0.0, 0.0, 0.0), (4e5, 4e5, 4e5), orc.EXTRAP_PERIODIC)
    xv0 = np.hstack([rng.uniform(0, 1.6e6, (6, 3)), rng.normal(scale=3e4, size=(6, 3))])
    state, iters, recs, times, _ = _run(orc, orc.zero(), g, xv0, t0=0.0, t_end=1e9, ci=10 ** 12,
                                        maxiters=300, K=7)
    for i in range(len(xv0)):
        Bv = orc.field_eval(g, xv0[i, :3], 0.0)
        dt = (2.0 * math.pi) * 0.1 / (abs(Q2M) * math.sqrt((Bv[0] * Bv[0] + Bv[1] * Bv[1]) + Bv[2] * Bv[2]))
        assert state[i, 7] == dt
        r = orc.solve(xv0[i], orc.zero(), g, q2m=Q2M, tspan=(0.0, 300 * dt), dt=dt, savestepinterval=7,
                      save_start=False, save_end=False)
        tb, ub = r.traj(0)
        hyb = np.array(recs[i][:-1])          # in-loop records; the last is the closing one
        assert len(hyb) == len(ub)
        np.testing.assert_array_equal(hyb, ub)
        np.testing.assert_allclose(times[i][:-1], tb, rtol=1e-13)
        assert iters[i, 1] == 300 and iters[i, 0] == 301


def test_checks_truncation_cap_clamp_and_outside(orc):
    rng = np.random.default_rng(6)
    dip = orc.dipole(3.12e-5, 6371.0e3)
    L = rng.uniform(4, 6, 8)
    ph = rng.uniform(0, 2 * np.pi, 8)
    x = np.stack([L * 6371e3 * np.cos(ph), L * 6371e3 * np.sin(ph), np.zeros(8)], axis=1)
    v = rng.normal(scale=3e6, size=(8, 3))
    xv0 = np.hstack([x, v])
    t_end = 7.3
    ref = _run(orc, orc.zero(), dip, xv0, t0=0.25, t_end=t_end, ci=10 ** 12, K=2, dtmax=0.05)
    for ci in (1, 3, 10):
        got = _run(orc, orc.zero(), dip, xv0, t0=0.25, t_end=t_end, ci=ci, K=2, dtmax=0.05)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
        for a, b in zip(got[2], ref[2]):
            np.testing.assert_array_equal(np.array(a), np.array(b))
        assert got[4] > ref[4]                 # more calls, same trajectory
    state, iters = ref[0], ref[1]
    assert np.all(state[:, 6] == t_end)        # the last step is cut at t_end
    assert np.all(state[:, 7] <= 0.05)         # dt clamped to dtmax
    # the step cap stops every particle after maxiters steps
    capd = _run(orc, orc.zero(), dip, xv0, t0=0.25, t_end=t_end, ci=4, maxiters=37, dtmax=0.05)
    assert np.all(capd[1][:, 1] == 37) and np.all(capd[0][:, 6] < t_end)
    # dtmin wins over a larger computed step
    big = _run(orc, orc.zero(), dip, xv0[:1], t0=0.0, t_end=1e-3, ci=5, dtmin=2e-4, dtmax=3e-4,
               safety=1e-9)
    assert big[0][0, 7] == 2e-4 or big[0][0, 6] == 1e-3
    # isoutside: a box the particles leave; the stopping state keeps x and t,
    # and x + v dt with the kept velocity lies outside the box
    lo, hi = x.min(axis=0) - 2e6, x.max(axis=0) + 2e6
    st, it, recs, times, _ = _run(orc, orc.zero(), dip, xv0, t0=0.0, t_end=1e6, ci=6, dtmax=0.5,
                                  bnd_flags=8, box_lo=lo, box_hi=hi)
    for i in range(len(xv0)):
        xn = st[i, :3] + st[i, 3:6] * st[i, 7]
        assert np.all((st[i, :3] >= lo) & (st[i, :3] <= hi))
        assert np.any((xn < lo) | (xn > hi))

"""tpamd_solve_batch_noalloc, the allocation-free batched solve, through the
C ABI against the batched solve whose kernels it shares (tpamd_solve_batch):
bit-identical records, lengths and return codes, on two calls in a row, the
second with a smaller input (the scratch array keeps the first call's
contents, which the second must not see)."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from testparticle_amd.device import solve_batch_noalloc

pytestmark = pytest.mark.gpu


def _problems(n_inst, total, seed, layout):
    inst, tspan = W.batched_turbulence(n_inst=n_inst, total=total, seed=seed, sizes=(16, 24, 32),
                                       max_steps=400)
    return [TP.TraceProblem(xv, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(),
                                        TP.GridField(A, *axes, layout=layout), TP.ZeroField()))
            for A, axes, xv in inst]


def _check(probs, alg, kw, ppl=0):
    ref = TP.solve_batch(probs, alg, **kw)
    u, t, ln, ret, off, st = solve_batch_noalloc(probs, alg, particles_per_lane=ppl, **kw)
    u, t, ln, ret = (a.cpu().numpy() for a in (u, t, ln, ret))
    steps = 0
    for k, s in enumerate(ref):
        a, b = int(off[k]), int(off[k + 1])
        np.testing.assert_array_equal(ln[a:b], s.lengths)
        np.testing.assert_array_equal(ret[a:b], s.retcodes)
        for j in range(b - a):
            n = int(s.lengths[j])
            np.testing.assert_array_equal(u[a + j, :n], s.u_all[j, :n])
            np.testing.assert_array_equal(t[a + j, :n], s.t_all[j, :n])
    assert st.particle_steps > 0
    return st


@pytest.mark.parametrize("case", ["adaptive", "fixed"])
def test_noalloc_batch_matches_batch_twice(case):
    if case == "adaptive":
        alg = TP.AdaptiveBoris(safety=0.1)
        kw = dict(isoutside=TP.OutsideGrid(), max_steps=400, save_everystep=False)
    else:
        alg = TP.MultistepBoris(4, n=2)
        kw = dict(dt=2 * np.pi * TP.mi / (TP.qi * 5e-9) / 40, savestepinterval=25)
    big = _problems(12, 300_000, 31, "cells")
    small = _problems(5, 40_000, 32, "nodes")
    if case == "fixed":
        big = [TP.TraceProblem(p.u0, (0.0, 200 * kw["dt"]), p.p) for p in big]
        small = [TP.TraceProblem(p.u0, (0.0, 120 * kw["dt"]), p.p) for p in small]
    s1 = _check(big, alg, kw)
    s2 = _check(small, alg, kw, ppl=2)
    assert s2.particle_steps < s1.particle_steps


def test_noalloc_batch_limits():
    probs = _problems(3, 3000, 33, "nodes")
    alg = TP.AdaptiveBoris(safety=0.1)
    with pytest.raises(TP.CapacityError):
        solve_batch_noalloc(probs * 100, alg, isoutside=TP.OutsideGrid(), max_steps=10,
                            save_everystep=False)

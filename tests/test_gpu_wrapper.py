"""The trace() caller on the device (bindings/python/src/testparticle/
__init__.py:330-439), with the reference wrapper's own tests restated
(bindings/python/tests/test_wrapper.py:25-155) and the result checked bit
for bit against the oracle and against TP.solve of the same problem."""
import numpy as np
import pytest

import testparticle_amd as TP
from parity import assert_bitwise, run_oracle

pytestmark = pytest.mark.gpu

GRID = (np.linspace(-1e7, 1e7, 3),) * 3


def test_gyrofrequency():
    # test_wrapper.py:25-31
    assert abs(float(TP.get_gyrofrequency(1.0, q=1.0, m=1.0)) - 1.0) < 1e-12


def test_trace_gyroradius_circle():
    # test_wrapper.py:34-62, the field given as a callable tabulated on a grid
    B, v0 = 1.0e-8, 1.0e5

    def get_B(x):
        return [0.0, 0.0, B]

    q, m = TP.Proton.q, TP.Proton.m
    omega = abs(q) * B / m
    r_expected = v0 / omega
    t_end = 2.0 * np.pi / omega * 1.05
    dt = t_end / 2000
    res = TP.trace([0.0, 0.0, 0.0], [v0, 0.0, 0.0], (0.0, t_end), dt, get_B, species="proton",
                   grid=(np.linspace(-2 * r_expected, 2 * r_expected, 3),) * 3)
    centre = res.x[:, 0:2].mean(axis=0)
    radius = np.sqrt((res.x[:, 0] - centre[0]) ** 2 + (res.x[:, 1] - centre[1]) ** 2)
    assert np.all(np.abs(radius - r_expected) / r_expected < 0.05)
    # the same trace through a uniform field description: the tabulated grid
    # returns the constant up to the last bit of its blend
    res_u = TP.trace([0.0, 0.0, 0.0], [v0, 0.0, 0.0], (0.0, t_end), dt, [0.0, 0.0, B])
    np.testing.assert_array_equal(res.t, res_u.t)
    np.testing.assert_allclose(res.x, res_u.x, rtol=0, atol=1e-9 * r_expected)


def test_trace_save_fields_and_work():
    # test_wrapper.py:65-84
    res = TP.trace([0.0, 0.0, 0.0], [1.0e5, 0.0, 0.0], (0.0, 1.0e-7), 5.0e-10,
                   lambda x: [0.0, 0.0, 1.0e-8], species="proton", save_fields=True, grid=GRID)
    assert res.state.shape[1] == 12
    assert res.B.shape == (res.t.size, 3)
    assert np.allclose(res.B, [0.0, 0.0, 1.0e-8], atol=1e-12)
    w = TP.trace([0.0, 0.0, 0.0], [1.0e5, 0.0, 0.0], (0.0, 1.0e-7), 5.0e-10, [0.0, 0.0, 1.0e-8],
                 E=[0.0, 1e-3, 0.0], save_fields=True, save_work=True)
    assert w.state.shape[1] == 16 and w.work.shape == (w.t.size, 4)
    # the result equals TP.solve of the same problem, and the oracle
    prob = TP.TraceProblem([0.0, 0.0, 0.0, 1.0e5, 0.0, 0.0], (0.0, 1.0e-7),
                           (TP.qi / TP.mi, TP.mi, TP.UniformField([0.0, 1e-3, 0.0]),
                            TP.UniformField([0.0, 0.0, 1.0e-8]), TP.ZeroField()))
    kw = dict(dt=5.0e-10, save_fields=True, save_work=True)
    s = TP.solve(prob, TP.Boris(), **kw)
    np.testing.assert_array_equal(w.state, s.u[0].u)
    assert_bitwise(s, run_oracle(prob, TP.Boris(), **kw))


def test_trace_ensemble_init_func():
    # test_wrapper.py:103-130
    def init_func(i):
        return [float(i), 0.0, 0.0, 1.0e5, 0.0, 0.0]

    results = TP.trace([0.0, 0.0, 0.0], [1.0e5, 0.0, 0.0], (0.0, 1.0e-7), 5.0e-10,
                       lambda x: [0.0, 0.0, 1.0e-8], species="proton", trajectories=2,
                       init_func=init_func, grid=GRID)
    assert isinstance(results, list) and len(results) == 2
    assert np.isclose(results[0].x[0, 0], 1.0, atol=1e-9)
    assert np.isclose(results[1].x[0, 0], 2.0, atol=1e-9)


def test_field_time_dependence_detected():
    # test_wrapper.py:133-155: a field B(x, t=0.0) is time-dependent; the
    # tabulated snapshots hold its values at the requested non-zero times and
    # the trace follows them
    seen = []

    def get_B(x, t=0.0):
        seen.append(t)
        return [0.0, 0.0, 1.0e-8 * (1.0 + 1e7 * t)]

    times = np.linspace(0.0, 1.0e-7, 5)
    res = TP.trace([0.0, 0.0, 0.0], [1.0e5, 0.0, 0.0], (0.0, 1.0e-7), 5.0e-10, get_B,
                   species="proton", grid=GRID, times=times, save_fields=True)
    assert len(seen) > 0 and any(t != 0.0 for t in seen)
    # B_z grows linearly in t, which the snapshot blend reproduces
    np.testing.assert_allclose(res.B[:, 2], 1.0e-8 * (1.0 + 1e7 * res.t), rtol=1e-12)
    static = TP.trace([0.0, 0.0, 0.0], [1.0e5, 0.0, 0.0], (0.0, 1.0e-7), 5.0e-10,
                      [0.0, 0.0, 1.0e-8])
    assert not np.array_equal(res.x, static.x)


def test_trace_algorithms_and_species():
    kw = dict(tspan=(0.0, 1e-6), dt=1e-9)
    a = TP.trace([0, 0, 0], [1e5, 0, 0], kw["tspan"], kw["dt"], [0, 0, 1e-8], alg="MultistepBoris4")
    prob = TP.TraceProblem([0, 0, 0, 1e5, 0, 0], kw["tspan"],
                           (TP.qi / TP.mi, TP.mi, TP.ZeroField(), TP.UniformField([0, 0, 1e-8]),
                            TP.ZeroField()))
    np.testing.assert_array_equal(a.state, TP.solve(prob, TP.MultistepBoris(4), dt=1e-9).u[0].u)
    e = TP.trace([0, 0, 0], [1e5, 0, 0], kw["tspan"], kw["dt"], [0, 0, 1e-8], species="electron")
    pe = TP.TraceProblem([0, 0, 0, 1e5, 0, 0], kw["tspan"],
                         (TP.qe / TP.me, TP.me, TP.ZeroField(), TP.UniformField([0, 0, 1e-8]),
                          TP.ZeroField()))
    assert_bitwise(TP.solve(pe, TP.Boris(), dt=1e-9), run_oracle(pe, TP.Boris(), dt=1e-9))
    np.testing.assert_array_equal(e.state, TP.solve(pe, TP.Boris(), dt=1e-9).u[0].u)

"""Time-dependent gridded fields on the device (LazyTimeInterpolator,
src/utility/interpolation.jl:337-430): the reference's own assertions
(test/test_utility.jl:312-404, test/test_derivatives.jl:57-84) through the
device evaluator; FP64 solves, records and evaluations bit-identical to the
oracle; FP32C close to FP64."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_bitwise, oracle_field, run_oracle

import oracle as O

pytestmark = pytest.mark.gpu

BCS = {"periodic": TP.PeriodicBC(), "clamp": TP.ClampExtrap(), "fill_nan": TP.FillExtrap()}


def _linear_itp(times):
    g = np.linspace(0.0, 4.0, 5)

    def loader(i):
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        A = np.stack([X, Y, Z]) * times[i - 1]           # E(x, t_i) = x t_i, A[c, i, j, k]
        return TP.build_interpolator(A, g, g, g)

    return TP.LazyTimeInterpolator(times, loader)


def test_reference_kats_on_device():
    # test_utility.jl:312-358
    itp = _linear_itp([0.0, 1.0, 2.0])
    x = [1.0, 2.0, 3.0]
    np.testing.assert_allclose(itp(x, 0.0), [0.0, 0.0, 0.0], atol=1e-14)
    np.testing.assert_allclose(itp(x, 1.0), [1.0, 2.0, 3.0], rtol=1e-14)
    np.testing.assert_allclose(itp(x, 0.5), [0.5, 1.0, 1.5], rtol=1e-14)
    np.testing.assert_allclose(itp(x, 1.5), [1.5, 3.0, 4.5], rtol=1e-14)
    np.testing.assert_array_equal(itp(x, 3.0), itp(x, 2.0))
    np.testing.assert_array_equal(itp(x, -1.0), itp(x, 0.0))
    # test_utility.jl:359-404
    g = np.linspace(0.0, 1.0, 4)
    B0 = np.zeros((3, 4, 4, 4))
    B1 = np.zeros((3, 4, 4, 4))
    B0[0] = 1.0
    B1[1] = 2.0
    num = TP.LazyTimeInterpolator([0.0, 1.0], lambda i: TP.build_interpolator((B0, B1)[i - 1], g, g, g))
    pt = [0.5, 0.5, 0.5]
    for t, want in ((0.0, [1.0, 0.0, 0.0]), (1.0, [0.0, 2.0, 0.0]), (0.5, [0.5, 1.0, 0.0]),
                    (0.25, [0.75, 0.5, 0.0])):
        np.testing.assert_allclose(num(pt, t), want, rtol=1e-15, atol=1e-15)
    # test_derivatives.jl:57-84, through the record evaluator: a particle at
    # rest in B = x t has zero work rates, but get_fields' Jacobian path runs
    q = TP.TraceProblem([1.0, 2.0, 3.0, 0.0, 0.0, 0.0], (0.0, 1.0),
                        (1.0, 1.0, TP.ZeroField(), itp, TP.ZeroField()))
    r = TP.record_eval(q, np.array([[1.0, 2.0, 3.0, 0.0, 0.0, 0.0]]), np.array([0.5]))
    np.testing.assert_allclose(r[0, 3:6], [0.5, 1.0, 1.5], rtol=1e-14)


def _turbulent_itp(n, nsnap, bc, storage=None, dtype=np.float64, seed=50):
    L = 1e7
    ax = W.grid_axes(n, L, periodic=isinstance(bc, TP.PeriodicBC))
    As = [W.turbulent_field(n, L, 5e-9, 0.5, seed + k, dtype=dtype) for k in range(nsnap)]
    times = np.array([0.0, 13.0, 21.5, 40.0, 47.0][:nsnap])
    return TP.LazyTimeInterpolator(times, lambda i: TP.GridField(As[i - 1], *ax, bc=bc,
                                                                 storage=storage)), L


@pytest.mark.parametrize("extrap", ["periodic", "clamp", "fill_nan"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_device_eval_bitwise(extrap, storage):
    itp, L = _turbulent_itp(12, 4, BCS[extrap], storage=storage)
    orc = oracle_field(itp)
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.1 * L, 1.1 * L, size=(300, 3))
    for t in (-3.0, 0.0, 5.2, 13.0, 20.0, 39.99, 40.0, 55.0):
        got = itp(x, t)
        want = np.array([O.field_eval(orc, xi, t) for xi in x])
        np.testing.assert_array_equal(got, want)


def _ensemble(L, N, seed, vth=6e5):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.1 * L, 0.9 * L, size=(N, 3))
    v = W.isotropic_directions(rng, N) * vth
    return np.hstack([x, v])


@pytest.mark.parametrize("extrap", ["periodic", "clamp", "fill_nan"])
def test_solves_bitwise(extrap):
    itp, L = _turbulent_itp(12, 4, BCS[extrap])
    E = TP.UxBField([3e5, -1e5, 0.0], itp)
    prob = TP.TraceProblem(_ensemble(L, 256, 5), (-4.0, 52.0), (TP.qi / TP.mi, TP.mi, E, itp,
                                                                   TP.ZeroField()))
    kw = dict(dt=0.35, savestepinterval=16)
    for alg in (TP.Boris(), TP.MultistepBoris(2, n=2), TP.MultistepBoris(4, n=3)):
        assert_bitwise(TP.solve(prob, alg, **kw), run_oracle(prob, alg, **kw))
    # records: fields, and work rates whose betatron term holds dB/dt
    kr = dict(kw, save_fields=True, save_work=True)
    g = TP.solve(prob, TP.Boris(), **kr)
    assert_bitwise(g, run_oracle(prob, TP.Boris(), **kr))
    filled = np.arange(g.u_all.shape[1])[None, :] < g.lengths[:, None]
    assert np.any(np.abs(g.u_all[filled][:, 15]) > 0)
    # adaptive: time-dependent fields are evaluated afresh at every step
    ka = dict(savestepinterval=4, max_steps=4000, nout_cap=1200)
    pa = TP.TraceProblem(prob.u0[:64], (-4.0, 52.0), prob.p)
    for alg in (TP.AdaptiveBoris(safety=0.1), TP.AdaptiveMultistepBoris[2](n=2)):
        assert_bitwise(TP.solve(pa, alg, **ka), run_oracle(pa, alg, **ka))


def test_outside_grid_and_batch():
    itp, L = _turbulent_itp(10, 3, TP.ClampExtrap())
    prob = TP.TraceProblem(_ensemble(L, 200, 9, vth=3e6), (0.0, 30.0),
                           (TP.qi / TP.mi, TP.mi, TP.ZeroField(), itp, TP.ZeroField()))
    kw = dict(dt=0.25, savestepinterval=10, isoutside=TP.OutsideGrid())
    g = TP.solve(prob, TP.Boris(), **kw)
    assert np.any(g.retcodes == 1)                      # some particles leave the grid
    assert_bitwise(g, run_oracle(prob, TP.Boris(), **kw))
    # batched: one instance with a time grid beside a static one
    B2, _ = _turbulent_itp(10, 1, TP.ClampExtrap(), seed=70)
    p2 = TP.TraceProblem(prob.u0[:120], (0.0, 30.0), (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B2,
                                                      TP.ZeroField()))
    sols = TP.solve_batch([prob, p2], TP.Boris(), **kw)
    for pr, s in zip((prob, p2), sols):
        r = TP.solve(pr, TP.Boris(), **kw)
        np.testing.assert_array_equal(s.lengths, r.lengths)
        np.testing.assert_array_equal(s.retcodes, r.retcodes)
        filled = np.arange(r.u_all.shape[1])[None, :] < r.lengths[:, None]   # records past a
        np.testing.assert_array_equal(s.u_all[filled], r.u_all[filled])       # length are unset


def test_fp32c_close_to_fp64():
    itp, L = _turbulent_itp(16, 3, TP.PeriodicBC(), storage="f32", dtype=np.float32)
    # 10 steps across the first snapshot interval: FP32C rounding only (the
    # same float storage on both sides), before chaos amplifies it
    prob = TP.TraceProblem(_ensemble(L, 512, 12), (10.0, 13.0),
                           (TP.qi / TP.mi, TP.mi, TP.UxBField([4e5, 0, 0], itp), itp, TP.ZeroField()))
    kw = dict(dt=0.3, savestepinterval=1)
    a = TP.solve(prob, TP.Boris(), precision="fp32c", **kw)
    b = TP.solve(prob, TP.Boris(), **kw)
    np.testing.assert_array_equal(a.lengths, b.lengths)
    np.testing.assert_allclose(a.u_all[:, :, :3], b.u_all[:, :, :3], rtol=0, atol=1e-6 * L)
    np.testing.assert_allclose(a.u_all[:, :, 3:], b.u_all[:, :, 3:], rtol=0, atol=1e-4 * 6e5)
    x = np.random.default_rng(1).uniform(0, L, size=(500, 3))
    for t in (0.0, 7.7, 25.0, 31.0):
        np.testing.assert_allclose(itp(x, t, precision="fp32c"), itp(x, t), rtol=0, atol=1e-6 * 1e-8)


def test_single_snapshot_equals_static_grid():
    itp, L = _turbulent_itp(12, 1, TP.PeriodicBC())
    static = itp._first
    mk = lambda B: TP.TraceProblem(_ensemble(L, 128, 2), (0.0, 20.0),
                                   (TP.qi / TP.mi, TP.mi, TP.UxBField([2e5, 0, 0], B), B, TP.ZeroField()))
    kw = dict(dt=0.2, savestepinterval=10)
    np.testing.assert_array_equal(TP.solve(mk(itp), TP.Boris(), **kw).u_all,
                                  TP.solve(mk(static), TP.Boris(), **kw).u_all)
    with pytest.raises(ValueError):
        TP.LazyTimeInterpolator([1.0, 1.0], lambda i: static)

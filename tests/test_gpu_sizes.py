"""Ensemble sizes at the edges of the launch shapes: none, one particle, one
short of a wave, one past it, and a ragged last block of 256, on each kernel
family (uniform fields, the dipole, gridded B in the specialised and the
generic kernels, adaptive).  FP64 against the oracle bit for bit; FP32C
against the generic kernels (kernels="generic") for the same sizes."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_bitwise, run_oracle

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 63, 65, 257]


def check(prob, alg, fp32c=True, **kw):
    g = TP.solve(prob, alg, **kw)
    assert len(g) == len(np.asarray(prob.u0))
    assert_bitwise(g, run_oracle(prob, alg, **kw))
    if fp32c:
        a = TP.solve(prob, alg, precision="fp32c", **kw)
        b = TP.solve(prob, alg, precision="fp32c", kernels="generic", **kw)
        np.testing.assert_array_equal(a.lengths, b.lengths)
        np.testing.assert_array_equal(a.retcodes, b.retcodes)
        np.testing.assert_array_equal(a.u_all, b.u_all)
        np.testing.assert_array_equal(a.t_all, b.t_all)


@pytest.fixture(scope="module")
def turb32():
    w, A = W.turbulent_box(N=257, n=32, nsteps=200, dtype=np.float64, seed=31)
    return w, A


@pytest.mark.parametrize("N", SIZES)
def test_sizes_uniform_fields(N):
    w = W.uniform_gyration(N=257, nsteps=640)
    prob = TP.TraceProblem(w.xv0[:N], w.tspan, (w.q2m, TP.mi, TP.UniformField([1e-6, 0.0, 0.0]),
                                                TP.UniformField(w.extra["B"]), TP.ZeroField()))
    check(prob, TP.Boris(), dt=w.dt, savestepinterval=w.savestepinterval)


@pytest.mark.parametrize("N", SIZES)
def test_sizes_dipole(N):
    w = W.dipole_belt(N=257, nsteps=500, seed=9)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0[:N], (0.0, 0.0), param)
    check(prob, TP.Boris(), dts=w.dts[:N], nsteps=w.nsteps, savestepinterval=50,
          isoutside=TP.InsideSphere(3 * TP.Re))


@pytest.mark.parametrize("N", SIZES)
def test_sizes_grid_fixed_step(turb32, N):
    w, A = turb32
    ax = W.grid_axes(32, w.extra["L"])
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC(), layout="cells")
    prob = TP.TraceProblem(w.xv0[:N], w.tspan, (w.q2m, TP.mi, TP.UxBField(w.extra["u"], B), B,
                                                TP.ZeroField()))
    for alg in (TP.Boris(), TP.MultistepBoris(4, n=2)):
        for kernels in ("auto", "generic"):
            g = TP.solve(prob, alg, dt=w.dt, savestepinterval=20, kernels=kernels)
            assert_bitwise(g, run_oracle(prob, alg, dt=w.dt, savestepinterval=20))


@pytest.mark.parametrize("N", SIZES)
def test_sizes_adaptive(turb32, N):
    w, A = turb32
    L = w.extra["L"]
    ax = W.grid_axes(32, L, periodic=False)
    B = TP.GridField(A, *ax, bc=TP.FillExtrap())
    prob = TP.TraceProblem(w.xv0[:N], (0.0, 200 * w.dt), (w.q2m, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    kw = dict(isoutside=TP.OutsideBox((0, 0, 0), (L, L, L)), savestepinterval=5, max_steps=30, nout_cap=12)
    g = TP.solve(prob, TP.AdaptiveBoris(safety=0.05), **kw)
    assert len(g) == N
    assert_bitwise(g, run_oracle(prob, TP.AdaptiveBoris(safety=0.05), **kw))

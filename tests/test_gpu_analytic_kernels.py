"""The kernels compiled for one pair of analytic fields against the generic
analytic kernels (kernels="generic"), bit for bit: fixed-step Boris in fields
that are the same everywhere (the update's rotation vectors formed once per
particle) and in the dipole with no E (no field-kind switch per step).  FP64
is also held to the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_bitwise, run_oracle

pytestmark = pytest.mark.gpu


def assert_same(a, b):
    np.testing.assert_array_equal(a.lengths, b.lengths)
    np.testing.assert_array_equal(a.retcodes, b.retcodes)
    np.testing.assert_array_equal(a.u_all, b.u_all)
    np.testing.assert_array_equal(a.t_all, b.t_all)


def both(prob, precision, **kw):
    a = TP.solve(prob, TP.Boris(), precision=precision, **kw)
    b = TP.solve(prob, TP.Boris(), precision=precision, kernels="generic", **kw)
    assert_same(a, b)
    return a


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_dipole_kernel_matches_generic(precision):
    w = W.dipole_belt(N=2048, nsteps=3000, seed=5)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    base = dict(dts=w.dts, nsteps=w.nsteps)
    # removal at 1 R_E, no intermediate records
    a = both(prob, precision, save_everystep=False, isoutside=TP.InsideSphere(TP.Re), **base)
    assert a.lengths.min() >= 1
    # records every 250 steps, loss cone and outer sphere both removing
    a = both(prob, precision, savestepinterval=250,
             isoutside=TP.InsideSphere(3 * TP.Re) | TP.OutsideSphere(6 * TP.Re), **base)
    assert (a.retcodes == 1).any() and (a.retcodes == 0).any()
    # no boundary at all
    both(prob, precision, savestepinterval=1000, **base)


def test_dipole_kernel_fp64_matches_oracle():
    w = W.dipole_belt(N=1024, nsteps=2000, seed=6)
    xv0 = w.xv0.copy()
    # every 7th particle starts on the equator (B_x = B_y = 0: the Boris
    # divisions have zero numerators) and every 11th 1e-300 m above it (tiny
    # numerators): their first steps take the plain-division path, lanes of
    # the same waves the shared-denominator one
    xv0[::7, 2] = 0.0
    xv0[::11, 2] = 1e-300
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(xv0, (0.0, 0.0), param)
    kw = dict(dts=w.dts, nsteps=w.nsteps, savestepinterval=100,
              isoutside=TP.InsideSphere(3 * TP.Re) | TP.OutsideSphere(6 * TP.Re))
    assert_bitwise(TP.solve(prob, TP.Boris(), **kw), run_oracle(prob, TP.Boris(), **kw))


@pytest.mark.parametrize("scale", [1e-160, 1e-152, 1e-100, 1.0, 1e9, 1e10, 1e11, 1e30])
def test_dipole_kernel_fp64_across_field_magnitudes(scale):
    # the dipole kernel's three Boris divisions share their denominator
    # 1 + t.t while it is <= 2^64 and every 2 t_c is >= 2^-512 in magnitude,
    # and divide plainly otherwise.  Scaling the field with the same steps
    # moves |t| from ~1e-161 to ~1e29 (about 0.1 unscaled): 1e-160 and
    # 1e11, 1e30 lie outside the shared range, 1e-152 (|2 t| ~ 2e-153) and
    # 1e10 (1 + t.t ~ 1e18) just inside its edges.  Bit for bit either way.
    w = W.dipole_belt(N=256, nsteps=200, seed=13)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"] * scale, w.extra["R"]),
             TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    kw = dict(dts=w.dts, nsteps=w.nsteps, savestepinterval=20)
    assert_bitwise(TP.solve(prob, TP.Boris(), **kw), run_oracle(prob, TP.Boris(), **kw))


def test_shared_division_equals_ieee_division():
    # SharedDiv (the dipole kernel's three divisions over one denominator)
    # against the compiler's IEEE division on 2^29 random pairs across the
    # range it is used in (d in [1, 2^64], |x| in [2^-512, 2^33]), edges
    # included: not one bit may differ
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "testparticle.jl_amd",
                       "_build", "shared_div_check")
    r = subprocess.run([exe, "1024"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout)
    assert out["pairs"] >= 5e8 and out["differing"] == 0


def test_negative_rotation_factor():
    # h = q/m dt / 2 < 0 for electrons and for tracing backward in time: the
    # numerators of the shared divisions and the hoisted t, s change sign.
    # Backward in time the reference keeps no closing record at t1 (its test
    # is tsave[end] < t_final, boris.jl:531), so those particles end
    # Terminated with one record fewer; the device follows it.
    w = W.dipole_belt(N=512, nsteps=300, seed=17)
    dip = TP.DipoleField(w.extra["b0"], w.extra["R"])
    cases = [(TP.TraceProblem(w.xv0, (0.0, 0.0), (-w.q2m * TP.mi / TP.me, TP.mi, TP.ZeroField(), dip,
                                                  TP.ZeroField())), dict(dts=w.dts, nsteps=w.nsteps)),
             (TP.TraceProblem(w.xv0, (0.0, 0.0), (w.q2m, TP.mi, TP.ZeroField(), dip, TP.ZeroField())),
              dict(dts=-w.dts, nsteps=w.nsteps))]
    g = W.uniform_gyration(N=256, nsteps=640)
    cases.append((TP.TraceProblem(g.xv0, (g.tspan[1], g.tspan[0]),
                                  (g.q2m, TP.mi, TP.UniformField([1e-6, 0.0, 0.0]),
                                   TP.UniformField(g.extra["B"]), TP.ZeroField())), dict(dt=-g.dt)))
    for prob, kw in cases:
        kw = dict(kw, savestepinterval=50)
        a = both(prob, "fp64", **kw)
        assert_bitwise(a, run_oracle(prob, TP.Boris(), **kw))
        both(prob, "fp32c", **kw)
    assert (a.retcodes == 1).all()           # backward: no closing record


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_uniform_kernel_matches_generic(precision):
    w = W.uniform_gyration(N=1024)
    B = w.extra["B"]
    for E in (TP.ZeroField(), TP.UniformField([1e-6, -2e-6, 5e-7])):
        prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, E, TP.UniformField(B), TP.ZeroField()))
        both(prob, precision, dt=w.dt, savestepinterval=w.savestepinterval)
        both(prob, precision, dt=w.dt, save_everystep=False)
        # a box on the scale of the gyroradii (~1e5 m) that particles leave
        # along B and across it at uneven times
        a = both(prob, precision, dt=w.dt, savestepinterval=7,
                 isoutside=TP.OutsideBox([-1.5e5, -1.5e5, -1.5e5], [1.5e5, 1.5e5, 1.5e5]))
        assert len(np.unique(a.lengths)) > 10


def test_uniform_kernel_fp64_matches_oracle():
    w = W.uniform_gyration(N=512)
    B = TP.UniformField(w.extra["B"])
    # uniform E, and E = -u x B of the uniform B (both one quad of lanes a particle)
    for E in (TP.UniformField([1e-6, -2e-6, 5e-7]), TP.UxBField([1e2, -3e1, 5e1], B)):
        prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, E, B, TP.ZeroField()))
        kw = dict(dt=w.dt, savestepinterval=w.savestepinterval)
        assert_bitwise(TP.solve(prob, TP.Boris(), **kw), run_oracle(prob, TP.Boris(), **kw))
        both(prob, "fp32c", **kw)
    # a step of its own per particle (dts with nsteps): each particle's t1 differs
    dts = w.dt * np.random.default_rng(3).uniform(0.5, 1.5, len(w.xv0))
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), (w.q2m, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    kw = dict(dts=dts, nsteps=700, savestepinterval=30)
    assert_bitwise(TP.solve(prob, TP.Boris(), **kw), run_oracle(prob, TP.Boris(), **kw))
    both(prob, "fp32c", **kw)

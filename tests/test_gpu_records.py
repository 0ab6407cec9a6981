"""save_fields / save_work records on the device (src/boris/boris.jl:405-440):
FP64 records bit-identical to the oracle for every field kind, solver and
kernel family; the reference's own assertions (test/test_boris.jl:359-507);
get_fields / get_work post-processing equal to the saved columns; FP32C
records close to FP64."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_bitwise, run_oracle

pytestmark = pytest.mark.gpu

ME, QE = 9.1093837015e-31, -1.602176634e-19


def _electron(E, B, u0, tspan):
    return TP.TraceProblem(u0, tspan, (QE / ME, ME, E, B, TP.ZeroField()))


def test_reference_save_fields_and_work():
    # test_boris.jl:359-372
    prob = _electron(TP.ZeroField(), TP.UniformField([0, 0, 0.01]), [0, 0, 0, 0, 1e5, 0], (0.0, 3e-8))
    s = TP.solve(prob, TP.Boris(), dt=3e-11, savestepinterval=10, save_fields=True)
    u = s.u[0].u
    assert u.shape[1] == 12
    for row in (u[0], u[-1]):
        assert list(row[6:9]) == [0.0, 0.0, 0.0] and list(row[9:12]) == [0.0, 0.0, 0.01]
    assert_bitwise(s, run_oracle(prob, TP.Boris(), dt=3e-11, savestepinterval=10, save_fields=True))
    # test_boris.jl:374-459
    E = TP.UniformField([0.0, 1e5, 0.0])
    gradB = TP.AffineField([0, 0, 0.01], np.array([[0, 0, 0], [0, 0, 0], [0.01, 0, 0]]), [0, 0, 0])
    prob = _electron(E, gradB, [0, 0, 0, 1e5, 0, 0], (0.0, 1e-9))
    kw = dict(dt=1e-11, savestepinterval=1)
    w = TP.solve(prob, TP.Boris(), save_work=True, **kw)
    wu = w.u[0].u
    assert wu.shape[1] == 10
    assert abs(wu[0, 6]) < 1e-10 and abs(wu[0, 7]) < 1e-10 and abs(wu[0, 9]) < 1e-10
    assert abs(wu[0, 8]) > 0
    assert_bitwise(w, run_oracle(prob, TP.Boris(), save_work=True, **kw))
    both = TP.solve(prob, TP.Boris(), save_work=True, save_fields=True, **kw)
    assert both.u[0].u.shape[1] == 16
    np.testing.assert_array_equal(both.u[0].u[:, 12:16], wu[:, 6:10])
    beta = TP.AffineField([0, 0, 0.01], np.zeros((3, 3)), [0, 0, 0.01 * 1e7])
    pb = _electron(TP.ZeroField(), beta, [0, 0, 0, 1e5, 0, 0], (0.0, 1e-9))
    sb = TP.solve(pb, TP.Boris(), save_work=True, **kw)
    assert sb.u[0].u[0, 9] > 0
    assert_bitwise(sb, run_oracle(pb, TP.Boris(), save_work=True, **kw))
    for alg in (TP.MultistepBoris(2, n=2), TP.MultistepBoris(4, n=2)):
        sm = TP.solve(prob, alg, save_work=True, save_fields=True, **kw)
        assert_bitwise(sm, run_oracle(prob, alg, save_work=True, save_fields=True, **kw))
    alg = TP.AdaptiveBoris(safety=0.1)
    sa = TP.solve(prob, alg, save_work=True, save_fields=True, nout_cap=3000)
    assert_bitwise(sa, run_oracle(prob, alg, save_work=True, save_fields=True, nout_cap=3000))


def test_postprocessing_equals_saved_columns():
    # test_boris.jl:461-507: get_fields / get_work of a plain solve equal the
    # columns a save_fields + save_work solve stores
    E = TP.UniformField([0.0, 1e5, 0.0])
    gradB = TP.AffineField([0, 0, 0.01], np.array([[0, 0, 0], [0, 0, 0], [0.01, 0, 0]]), [0, 0, 0])
    prob = _electron(E, gradB, [0, 0, 0, 1e5, 0, 0], (0.0, 1e-9))
    ref = TP.solve(prob, TP.Boris(), dt=1e-11, savestepinterval=1, save_fields=True,
                   save_work=True).u[0]
    sol = TP.solve(prob, TP.Boris(), dt=1e-11, savestepinterval=1).u[0]
    Ep, Bp = TP.get_fields(sol)
    wp = TP.get_work(sol)
    np.testing.assert_array_equal(Ep, ref.u[:, 6:9])
    np.testing.assert_array_equal(Bp, ref.u[:, 9:12])
    np.testing.assert_array_equal(wp, ref.u[:, 12:16])


@pytest.mark.parametrize("kernels", ["auto", "generic"])
@pytest.mark.parametrize("extrap", ["periodic", "clamp", "fill_nan"])
def test_gridded_records_bitwise(extrap, kernels):
    w, A = W.turbulent_box(N=256, n=24, nsteps=1, dtype=np.float64, seed=13)
    L = w.extra["L"]
    bc = {"periodic": TP.PeriodicBC(), "clamp": TP.ClampExtrap(), "fill_nan": TP.FillExtrap()}[extrap]
    ax = W.grid_axes(24, L, periodic=extrap == "periodic")
    B = TP.GridField(A, *ax, bc=bc)
    E = TP.UxBField(w.extra["u"], B)
    prob = TP.TraceProblem(w.xv0, (0.0, 300 * w.dt), (w.q2m, TP.mi, E, B, TP.ZeroField()))
    kw = dict(dt=w.dt, savestepinterval=37, save_fields=True, save_work=True)
    for alg in (TP.Boris(), TP.MultistepBoris(4, n=2)):
        g = TP.solve(prob, alg, kernels=kernels, **kw)
        assert_bitwise(g, run_oracle(prob, alg, **kw))


def test_dipole_records_bitwise_with_removal():
    w = W.dipole_belt(N=1024, nsteps=2000)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    kw = dict(dts=w.dts, nsteps=w.nsteps, savestepinterval=250, isoutside=TP.InsideSphere(TP.Re),
              save_fields=True, save_work=True)
    g = TP.solve(prob, TP.Boris(), **kw)
    assert_bitwise(g, run_oracle(prob, TP.Boris(), **kw))
    # trapped particles: the mirror force does no work in a static B, so only
    # the E-free rates enter; every work column is zero with E = 0
    filled = np.arange(g.u_all.shape[1])[None, :] < g.lengths[:, None]
    assert np.all(g.u_all[filled][:, 12:16] == 0.0)


def test_stored_E_grid_and_time_dependent_adaptive_bitwise():
    rng = np.random.default_rng(8)
    A = W.turbulent_field(16, 1e7, 5e-9, 0.4, 3, dtype=np.float64)
    ax = W.grid_axes(16, 1e7)
    Eg = TP.GridField(A * 1e5, *ax, bc=TP.PeriodicBC())
    Bg = TP.GridField(A, *ax, bc=TP.PeriodicBC())
    x = rng.uniform(0, 1e7, size=(200, 3))
    v = W.isotropic_directions(rng, 200) * 5e5
    prob = TP.TraceProblem(np.hstack([x, v]), (0.0, 50.0), (TP.qi / TP.mi, TP.mi, Eg, Bg, TP.ZeroField()))
    kw = dict(dt=0.2, savestepinterval=20, save_fields=True, save_work=True)
    assert_bitwise(TP.solve(prob, TP.Boris(), **kw), run_oracle(prob, TP.Boris(), **kw))
    # adaptive, time-dependent affine fields: betatron term non-zero
    Ea = TP.AffineField([0.1, -0.2, 0.05], rng.normal(size=9) * 0.05, [0.01, 0.0, -0.02])
    Ba = TP.AffineField([0.0, 0.3, 1.0], rng.normal(size=9) * 0.1, [0.0, 0.05, 0.0])
    pa = TP.TraceProblem(np.hstack([rng.normal(size=(128, 3)), rng.normal(size=(128, 3))]), (0.5, 20.5),
                         (1.3, 2.0, Ea, Ba, TP.ZeroField()))
    kwa = dict(savestepinterval=5, max_steps=3000, nout_cap=800, save_fields=True, save_work=True)
    for alg in (TP.AdaptiveBoris(safety=0.05), TP.AdaptiveMultistepBoris[4](n=2)):
        g = TP.solve(pa, alg, **kwa)
        assert_bitwise(g, run_oracle(pa, alg, **kwa))
        filled = np.arange(g.u_all.shape[1])[None, :] < g.lengths[:, None]
        assert np.any(np.abs(g.u_all[filled][:, 15]) > 0)
    # the wide-record kernels with lanes refilling from 64 x 3 particles per
    # wave, lives cut unevenly by a sphere (121 of 128 leave, after 3 to 31
    # records): same records
    kwr = dict(kwa, isoutside=TP.OutsideSphere(8.0))
    g1 = TP.solve(pa, TP.AdaptiveBoris(safety=0.05), **kwr)
    g3 = TP.solve(pa, TP.AdaptiveBoris(safety=0.05), particles_per_lane=3, **kwr)
    assert (g3.retcodes == 1).any() and (g3.retcodes == 0).any()
    assert_bitwise(g3, run_oracle(pa, TP.AdaptiveBoris(safety=0.05), **kwr))
    np.testing.assert_array_equal(g3.lengths, g1.lengths)


def test_fp32c_records_close_to_fp64():
    w, A = W.turbulent_box(N=512, n=32, nsteps=1, dtype=np.float64, seed=17)
    ax = W.grid_axes(32, w.extra["L"])
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC())
    prob = TP.TraceProblem(w.xv0, (0.0, 5 * w.dt), (w.q2m, TP.mi, TP.UxBField(w.extra["u"], B), B,
                                                    TP.ZeroField()))
    kw = dict(dt=w.dt, savestepinterval=1, save_fields=True, save_work=True)
    g64 = TP.solve(prob, TP.Boris(), **kw)
    g32 = TP.solve(prob, TP.Boris(), precision="fp32c", **kw)
    a, b = g32.u_all[:, :, 6:12], g64.u_all[:, :, 6:12]
    # fields at FP32C positions after 5 steps: float rounding of states and field storage
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-5 * np.abs(b).max())
    # the records' fields equal a direct evaluation at the FP32C record states
    x = g32.u_all.reshape(-1, 16)
    t = g32.t_all.reshape(-1)
    r = TP.record_eval(prob, x[:, :6], t)
    np.testing.assert_array_equal(r, x[:, 6:])


def test_batch_records_match_single_solves():
    rng = np.random.default_rng(4)
    probs = []
    for k, n in enumerate((12, 20, 9)):
        A = W.turbulent_field(n, 1e7, 5e-9, 0.5, 40 + k, dtype=np.float64)
        B = TP.GridField(A, *W.grid_axes(n, 1e7), bc=TP.PeriodicBC())
        x = rng.uniform(0, 1e7, size=(100 + 50 * k, 3))
        v = W.isotropic_directions(rng, len(x)) * 1e6
        probs.append(TP.TraceProblem(np.hstack([x, v]), (0.0, 40.0),
                                     (TP.qi / TP.mi, TP.mi, TP.UxBField([4e5, 0, 0], B), B, TP.ZeroField())))
    kw = dict(dt=0.33, savestepinterval=10, save_fields=True, save_work=True)
    sols = TP.solve_batch(probs, TP.Boris(), **kw)
    for pr, g in zip(probs, sols):
        s = TP.solve(pr, TP.Boris(), **kw)
        np.testing.assert_array_equal(g.u_all, s.u_all)

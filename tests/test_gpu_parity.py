"""GPU parity: the HIP path through the C ABI against the oracle and against
the reference's own known answers.

FP64 results must be bit-identical to the oracle (the device code keeps the
reference's operation order); the reference-form multistep coefficients
(atan/sincos from the device math library) are held to 1e-10 relative.
"""
import math

import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_bitwise, assert_close, run_oracle

pytestmark = pytest.mark.gpu


def q2m_of(kats, species):
    c = kats["constants"]
    return c["qe"] / c["me"] if species == "electron" else c["qi"] / c["mi"]


def sp(species):
    return TP.Electron if species == "electron" else TP.Proton


# ---------------------------------------------------------------------------
# Reference known answers, run on the device
# ---------------------------------------------------------------------------
def test_basic_boris_kat_bitexact(kats):
    k = kats["basic_boris"]
    param = TP.prepare(TP.ZeroField(), TP.UniformField(k["B"]), species=TP.Electron)
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    sol = TP.solve(prob, TP.Boris(), dt=k["dt"], savestepinterval=k["savestepinterval"]).u[0]
    assert len(sol.t) == len(sol.u) == 101
    assert np.array_equal(sol.u[-1], np.array(k["u_end"]))
    np.testing.assert_allclose(sol(k["interp_t"]), k["u_interp"], rtol=k["rtol"], atol=1e-20)
    # the KernelAbstractions-style positional backend routes to the same kernels
    sol2 = TP.solve(prob, TP.Boris(), TP.MI355X(0), dt=k["dt"], savestepinterval=10).u[0]
    assert np.array_equal(sol2.u, sol.u)


def test_ensemble_sums(kats):
    k = kats["basic_boris"]
    param = TP.prepare(TP.ZeroField(), TP.UniformField(k["B"]), species=TP.Electron)

    def prob_func(prob, ctx):
        u = list(prob.u0)
        u[4] = ctx.sim_id * 1.0e5
        return TP.remake(prob, u0=u)

    prob = TP.TraceProblem(k["u0"], k["tspan"], param, prob_func=prob_func)
    e = kats["ensemble_threads_sum"]
    sols = TP.solve(prob, TP.Boris(), TP.EnsembleThreads(), dt=k["dt"],
                    savestepinterval=e["savestepinterval"], trajectories=e["trajectories"])
    assert sum(s.u[-1][3] for s in sols.u) == pytest.approx(e["expected"], rel=1.5e-8)
    e = kats["ensemble_serial_sum"]
    sols = TP.solve(prob, TP.Boris(), dt=k["dt"], savestepinterval=e["savestepinterval"],
                    trajectories=e["trajectories"])
    assert sum(np.sum(s.u[-1]) for s in sols.u) == pytest.approx(e["expected"], rel=1.5e-8)


def _algs(spec):
    out = []
    for name, N, n in spec:
        out.append(TP.Boris() if name == "boris" else TP.MultistepBoris(N, n=n))
    return out


@pytest.mark.parametrize("form", ["rational", "reference"])
def test_exb_drift_and_gyration(kats, form):
    k = kats["exb_drift"]
    param = (k["q2m"], 1.0, TP.UniformField(k["E"]), TP.UniformField(k["B"]), TP.ZeroField())
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    for alg in _algs(k["algs"]):
        u = TP.solve(prob, alg, dt=k["dt"], form=form).u[0].u[-1]
        assert abs(u[0] - k["x_end"]) <= k["atol"] and abs(u[3] - k["vx_end"]) <= k["atol"], alg
    g = kats["gyration"]
    param = (g["q2m"], 1.0, TP.ZeroField(), TP.UniformField(g["B"]), TP.ZeroField())
    prob = TP.TraceProblem(g["u0"], g["tspan"], param)
    for alg in _algs(g["algs"]):
        u = TP.solve(prob, alg, dt=g["dt"], form=form).u[0].u[-1]
        assert math.hypot(*u[:3]) < g["max_radius_end"], alg
        assert abs(math.hypot(*u[3:]) - 1.0) <= g["speed_atol"], alg


def test_save_flags(kats):
    k = kats["basic_boris"]
    param = TP.prepare(TP.ZeroField(), TP.UniformField(k["B"]), species=TP.Electron)
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    t0, t1 = k["tspan"]
    dt = k["dt"]
    for cs in kats["save_flags"]["cases"]:
        sol = TP.solve(prob, TP.Boris(), dt=dt, savestepinterval=cs["K"],
                       save_everystep=cs["save_everystep"], save_start=cs["save_start"],
                       save_end=cs["save_end"]).u[0]
        assert len(sol.t) == cs["nout"]
        ref = {"t0": t0, "t1": t1}
        if "t_first" in cs:
            assert sol.t[0] == ref[cs["t_first"]]
        if "t_last" in cs:
            assert sol.t[-1] == ref[cs["t_last"]]
        if "t_first_steps" in cs:
            assert sol.t[0] == pytest.approx(t0 + cs["t_first_steps"] * dt)
            assert sol.t[-1] == pytest.approx(t0 + cs["t_last_steps"] * dt)
    ms = TP.solve(prob, TP.MultistepBoris(2, n=2), dt=dt, savestepinterval=10,
                  save_everystep=False).u[0]
    assert list(ms.t) == [t0, t1]


def test_kernel_path_kats(kats):
    k = kats["kernel_path"]
    c = kats["constants"]
    param = TP.prepare(TP.UniformField(k["E"]), TP.UniformField(k["B"]), species=TP.Proton)
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    sol = TP.solve(prob, TP.Boris(), TP.MI355X(), dt=k["dt"], savestepinterval=10).u[0]
    assert sol.t[0] == k["tspan"][0] and sol.t[-1] == pytest.approx(k["tspan"][1])
    assert abs(sol.u[-1][0]) > 0 and abs(sol.u[-1][1]) > 0
    Tg = 2 * math.pi / (abs(c["qi"]) * k["B"][2] / c["mi"])
    e = k["energy"]
    pg = TP.TraceProblem(k["u0"], (0.0, e["n_periods"] * Tg), param)
    u = TP.solve(pg, TP.Boris(), dt=Tg / e["steps_per_period"], savestepinterval=10).u[0].u[-1]
    assert math.sqrt(float(np.dot(u[3:], u[3:]))) == pytest.approx(1.0e5, rel=e["rtol"])
    with pytest.raises(TP.ArgumentError):
        TP.solve(prob, TP.Boris(), TP.MI355X(), dt=k["dt"], maxiters=k["maxiters_error"])
    with pytest.raises(TP.ArgumentError):
        TP.solve(prob, TP.Boris(), dt=np.finfo(np.float64).eps)


def test_nonzero_tspan_switch(kats):
    # B switches on at t > 5: an AffineField with a time slope reproduces
    # "B present at t in (10, 10.1)" (test_boris.jl:250-272) on the device
    k = kats["nonzero_tspan"]
    param = TP.prepare(TP.ZeroField(), TP.UniformField(k["B_on"]), species=TP.Proton)
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    u = TP.solve(prob, TP.Boris(), dt=k["dt"]).u[0].u[-1]
    assert abs(u[3]) < k["vx_bound"]
    u = TP.solve(prob, TP.AdaptiveBoris(safety=0.1)).u[0].u[-1]
    assert abs(u[3]) < k["vx_bound"]


def _kat_grid_field(k, extrap):
    nx, ny, nz = k["nx"], k["ny"], k["nz"]
    A = np.zeros((3, nx, ny, nz), dtype=np.float32)
    for i in range(nx):
        for j in range(ny):
            for kk in range(nz):
                A[:, i, j, kk] = (i + 1) + (j + 1) + (kk + 1)
    ax = [np.linspace(k["lo"], k["hi"], m) for m in (nx, ny, nz)]
    bc = {"fill_nan": TP.FillExtrap(), "clamp": TP.ClampExtrap(), "periodic": TP.PeriodicBC()}[extrap]
    return TP.build_interpolator(A, *ax, 1, bc)


def test_interpolation_kats_on_device(kats, orc):
    from parity import oracle_field
    k = kats["interpolation"]
    for cs in k["cases"]:
        f = _kat_grid_field(k, cs["extrap"])
        val = f(cs["x"])
        np.testing.assert_allclose(val, [cs["value"]] * 3, rtol=1.5e-8)
        # and bit-identical to the oracle at random points, inside and outside
        rng = np.random.default_rng(7)
        pts = rng.uniform(-30, 30, size=(500, 3))
        got = f(pts)
        of = oracle_field(f)
        ref = np.array([orc.field_eval(of, p) for p in pts])
        assert np.array_equal(got, ref, equal_nan=True)
    f = _kat_grid_field(k, "fill_nan")
    assert np.all(np.isnan(f([20.0, 0.0, 0.0])))


def test_nan_boundary_grid(kats):
    # fields NaN outside their grid; isoutside = any NaN | norm(x) > 1
    k = kats["nan_boundary"]
    ax = np.linspace(-1.0, 1.0, 5)
    Bg = np.zeros((3, 5, 5, 5)); Bg[2] = 1.0
    Eg = np.zeros((3, 5, 5, 5))
    param = (k["q2m"], 1.0, TP.GridField(Eg, ax, ax, ax), TP.GridField(Bg, ax, ax, ax), TP.ZeroField())
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    bnd = TP.TerminateOutside(TP.OutsideNaN() | TP.OutsideSphere(k["radius"]))
    for kw in (dict(dt=k["dt"]), dict()):
        alg = TP.Boris() if kw else TP.AdaptiveBoris(safety=k["adaptive_safety"])
        sol = TP.solve(prob, alg, isoutside=bnd.condition, **kw).u[0]
        assert not np.any(np.isnan(sol.u[-1]))
        assert math.hypot(*sol.u[-1][:3]) <= 1.0
        assert sol.retcode == "Terminated"


def test_spatial_rejection_and_integer_tspan(kats):
    k = kats["spatial_rejection"]
    param = TP.prepare(TP.ZeroField(), TP.UniformField([1.0, 0, 0]), species=TP.Proton)
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    sol = TP.solve(prob, TP.AdaptiveBoris(safety=k["safety"]),
                   isoutside=TP.OutsideBox((-np.inf,) * 3, (k["x_max"], np.inf, np.inf))).u[0]
    assert sol.u[-1][0] <= k["x_max"] and sol.t[-1] < k["tspan"][1]
    assert sol.retcode == "Terminated"
    k = kats["integer_tspan"]
    prob = TP.TraceProblem(k["u0"], tuple(k["tspan"]), param)
    sol = TP.solve(prob, TP.Boris(), dt=k["dt"]).u[0]
    assert sol.t[-1] == 10 and sol.retcode == "Success"


def test_adaptive_multistep_consistency(kats):
    k = kats["adaptive_multistep_consistency"]
    param = TP.prepare(TP.ZeroField(), TP.UniformField(k["B"]))
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    a = TP.solve(prob, TP.AdaptiveBoris(safety=k["safety"])).u[0]
    b = TP.solve(prob, TP.AdaptiveMultistepBoris[2](n=1, safety=k["safety"])).u[0]
    assert len(a.t) == len(b.t)
    np.testing.assert_allclose(b.t, a.t, rtol=1.5e-8)
    np.testing.assert_allclose(b.u, a.u, rtol=1.5e-8, atol=1e-12)


def test_solver_limits(kats):
    k = kats["solver_limits"]
    param = TP.prepare(TP.UniformField([0, 1e5, 0]),
                       TP.AffineField([0, 0, 0.01], [0, 0, 0, 0, 0, 0, 0.01, 0, 0]),
                       species=TP.Electron)
    prob = TP.TraceProblem(k["u0"], k["tspan"], param)
    with pytest.raises(TP.ArgumentError, match="maxiters"):
        TP.solve(prob, TP.Boris(), dt=k["dt"], maxiters=k["maxiters_error"])
    with pytest.raises(TP.ArgumentError, match="too small"):
        TP.solve(prob, TP.Boris(), dt=np.finfo(np.float64).eps)
    for N in k["bad_orders"]:
        with pytest.raises(TP.ArgumentError):
            TP.MultistepBoris(N, n=2)


# ---------------------------------------------------------------------------
# Device vs oracle, bit for bit, on the benchmark workloads (reduced sizes)
# ---------------------------------------------------------------------------
def test_parity_uniform_gyration_full():
    # the CPU-runnable configuration at full size: 1024 protons, 6400 steps, K = 10
    w = W.uniform_gyration()
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.UniformField(w.extra["B"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, w.tspan, param)
    g = TP.solve(prob, TP.Boris(), dt=w.dt, savestepinterval=w.savestepinterval)
    assert g.u_all.shape[1] == 641
    o = run_oracle(prob, TP.Boris(), dt=w.dt, savestepinterval=w.savestepinterval)
    assert_bitwise(g, o)
    assert g.stats.particle_steps == 1024 * 6400
    v0 = np.linalg.norm(w.xv0[:, 3:], axis=1)
    drift = np.abs(np.linalg.norm(g.final_states()[:, 3:], axis=1) ** 2 / v0 ** 2 - 1)
    assert drift.max() < 1e-12


def test_parity_dipole_belt():
    w = W.dipole_belt(N=2048, nsteps=10000)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    bnd = TP.InsideSphere(TP.Re)
    kw = dict(dts=w.dts, nsteps=w.nsteps, save_everystep=False, isoutside=bnd)
    g = TP.solve(prob, TP.Boris(), **kw)
    o = run_oracle(prob, TP.Boris(), **kw)
    assert_bitwise(g, o)
    # with save_everystep on a stride as well
    kw.update(save_everystep=True, savestepinterval=250)
    g = TP.solve(prob, TP.Boris(), **kw)
    o = run_oracle(prob, TP.Boris(), **kw)
    assert_bitwise(g, o)



@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_device_ensemble_matches_host_solve(precision):
    # device-resident buffers (tpamd_solve_device on torch's stream) against
    # the host-buffer entry point, with per-particle dt and a removal sphere
    pytest.importorskip("torch")
    from testparticle_amd.device import DeviceEnsemble

    w = W.dipole_belt(N=4096, nsteps=3000)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    kw = dict(dts=w.dts, nsteps=w.nsteps, save_everystep=False, isoutside=TP.InsideSphere(TP.Re),
              precision=precision)
    g = TP.solve(prob, TP.Boris(), **kw)
    e = DeviceEnsemble(prob, TP.Boris(), w.xv0, **kw)
    st = e.run()
    u, t, ln, ret = e.host()
    np.testing.assert_array_equal(ln, g.lengths)
    np.testing.assert_array_equal(ret, g.retcodes)
    np.testing.assert_array_equal(u[:, :g.u_all.shape[1]], g.u_all)
    np.testing.assert_array_equal(t[:, :g.t_all.shape[1]], g.t_all)
    assert st.particle_steps > 0

@pytest.fixture(scope="module")
def turb64():
    w, A = W.turbulent_box(N=1024, n=64, nsteps=2000, dtype=np.float64)
    ax = W.grid_axes(64, w.extra["L"])
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC())
    E = TP.UxBField(w.extra["u"], B)
    return w, A, ax, B, E


@pytest.mark.parametrize("form", ["reference", "rational"])
@pytest.mark.parametrize("kernels", ["auto", "generic"])
@pytest.mark.parametrize("N,n", [(4, 2), (2, 1), (6, 3), (2, 0)])
def test_parity_turbulent_multistep(turb64, N, n, kernels, form):
    # n = 0 stands for the standard Boris update; "auto" takes the gridded-B
    # specialisations where they apply (n <= 2), "generic" never does.  The
    # reference form (atan, sincos, Taylor branch) is bit-identical too: the
    # device and the oracle share one FDLIBM atan and sin/cos.  2,000 steps of
    # a chaotic flow, so a single differing bit anywhere would show.
    if n == 0 and form == "rational":
        pytest.skip("Boris has no multistep coefficients")
    w, A, ax, B, E = turb64
    prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, E, B, TP.ZeroField()))
    alg = TP.Boris() if n == 0 else TP.MultistepBoris(N, n=n)
    g = TP.solve(prob, alg, dt=w.dt, savestepinterval=100, kernels=kernels, form=form)
    o = run_oracle(prob, alg, dt=w.dt, savestepinterval=100, form=form)
    assert_bitwise(g, o)


@pytest.mark.parametrize("extrap", ["periodic", "clamp", "fill_nan"])
def test_parity_specialised_f32_storage_and_extrap(turb64, extrap):
    # FP64 arithmetic on float storage, every wrap mode, uniform E
    w, A, ax, _, _ = turb64
    bc = {"periodic": TP.PeriodicBC(), "clamp": TP.ClampExtrap(), "fill_nan": TP.FillExtrap()}[extrap]
    B = TP.GridField(A.astype(np.float32), *ax, bc=bc)
    prob = TP.TraceProblem(w.xv0[:512], (0.0, 300 * w.dt),
                           (w.q2m, TP.mi, TP.UniformField([1e-4, -2e-4, 0.0]), B, TP.ZeroField()))
    for alg in (TP.Boris(), TP.MultistepBoris(4, n=2)):
        g = TP.solve(prob, alg, dt=w.dt, savestepinterval=30)
        o = run_oracle(prob, alg, dt=w.dt, savestepinterval=30)
        assert_bitwise(g, o)


@pytest.mark.parametrize("extrap", ["periodic", "clamp", "fill_nan"])
def test_cell_layout_is_bit_identical(turb64, extrap):
    # the cell-replicated copy changes where node values live, not the blend:
    # FP64 stays bit-identical to the oracle, FP32C to its node-layout run
    w, A, ax, _, _ = turb64
    bc = {"periodic": TP.PeriodicBC, "clamp": TP.ClampExtrap, "fill_nan": TP.FillExtrap}[extrap]
    Bc = TP.GridField(A, *ax, bc=bc(), layout="cells")
    prob = TP.TraceProblem(w.xv0[:512], (0.0, 300 * w.dt),
                           (w.q2m, TP.mi, TP.UxBField(w.extra["u"], Bc), Bc, TP.ZeroField()))
    for alg in (TP.Boris(), TP.MultistepBoris(4, n=2)):
        g = TP.solve(prob, alg, dt=w.dt, savestepinterval=30)
        assert_bitwise(g, run_oracle(prob, alg, dt=w.dt, savestepinterval=30))
    A32 = A.astype(np.float32)
    Bn = TP.GridField(A32, *ax, bc=bc(), layout="nodes")
    Bc = TP.GridField(A32, *ax, bc=bc(), layout="cells")
    res = []
    for B in (Bn, Bc):
        pr = TP.TraceProblem(w.xv0[:512], (0.0, 300 * w.dt),
                             (w.q2m, TP.mi, TP.UxBField(w.extra["u"], B), B, TP.ZeroField()))
        res.append(TP.solve(pr, TP.MultistepBoris(4, n=2), dt=w.dt, savestepinterval=30,
                            precision="fp32c"))
        # FP64 over float storage, nodes and cells (one line per gather)
        for alg in (TP.Boris(), TP.MultistepBoris(4, n=2)):
            g = TP.solve(pr, alg, dt=w.dt, savestepinterval=30)
            assert_bitwise(g, run_oracle(pr, alg, dt=w.dt, savestepinterval=30))
    np.testing.assert_array_equal(res[0].u_all, res[1].u_all)


@pytest.mark.parametrize("form", ["reference", "rational"])
def test_fp32c_specialised_matches_generic(turb64, form):
    w, A, ax, _, _ = turb64
    B = TP.GridField(A.astype(np.float32), *ax, bc=TP.PeriodicBC())
    E = TP.UxBField(w.extra["u"], B)
    prob = TP.TraceProblem(w.xv0, (0.0, 10 * w.dt), (w.q2m, TP.mi, E, B, TP.ZeroField()))
    for alg in (TP.Boris(), TP.MultistepBoris(4, n=2), TP.MultistepBoris(2, n=1)):
        a = TP.solve(prob, alg, dt=w.dt, precision="fp32c", form=form).final_states()
        b = TP.solve(prob, alg, dt=w.dt, precision="fp32c", kernels="generic", form=form).final_states()
        d = np.linalg.norm(a[:, :3] - b[:, :3], axis=1) / np.linalg.norm(b[:, :3] - w.xv0[:, :3], axis=1)
        assert np.median(d) < 1e-5, np.median(d)


def test_parity_turbulent_stored_E_f32_storage(turb64):
    # E stored as its own grid (two gathers per evaluation), FP32 storage of
    # both grids, FP64 arithmetic
    w, A, ax, _, _ = turb64
    u = np.array(w.extra["u"])
    Eg = -np.cross(u[:, None, None, None], A, axis=0)
    B = TP.GridField(A.astype(np.float32), *ax, bc=TP.PeriodicBC())
    E = TP.GridField(Eg.astype(np.float32), *ax, bc=TP.PeriodicBC())
    prob = TP.TraceProblem(w.xv0[:256], w.tspan, (w.q2m, TP.mi, E, B, TP.ZeroField()))
    g = TP.solve(prob, TP.Boris(), dt=w.dt, savestepinterval=500)
    o = run_oracle(prob, TP.Boris(), dt=w.dt, savestepinterval=500)
    assert_bitwise(g, o)


def test_reference_form_multistep_bitwise_and_forms_agree():
    # the reference form (atan, sincos, Taylor branch below 1e-4) is
    # bit-identical to the oracle over the full 6,400-step gyration workload,
    # in a smooth field where the two forms must also agree to 1e-10 (the
    # kernel-path precedent, test_boris_kernel.jl:81-93); a large rotation per
    # subcycle (|t_n| up to ~3) takes sincos through its argument reduction
    w = W.uniform_gyration(N=256)
    B = TP.AffineField([0.0, 0.0, 1.0e-8], [0, 0, 0, 0, 0, 0, 1e-15, 0, 0])
    E = TP.UniformField([1e-4, 0.0, 0.0])
    prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, E, B, TP.ZeroField()))
    for alg in (TP.MultistepBoris(4, n=2), TP.MultistepBoris(6, n=3), TP.MultistepBoris(2, n=1)):
        g = TP.solve(prob, alg, dt=w.dt, savestepinterval=64, form="reference")
        o = run_oracle(prob, alg, dt=w.dt, savestepinterval=64, form="reference")
        assert_bitwise(g, o)
        r = TP.solve(prob, alg, dt=w.dt, savestepinterval=64, form="rational")
        np.testing.assert_array_equal(g.lengths, r.lengths)
        np.testing.assert_allclose(g.u_all[..., :3], r.u_all[..., :3], rtol=1e-10, atol=1e-10 * 1e5)
        np.testing.assert_allclose(g.u_all[..., 3:], r.u_all[..., 3:], rtol=1e-10, atol=1e-10 * 1e5)
    for dtf in (20.0, 200.0):
        for alg in (TP.MultistepBoris(4, n=2), TP.MultistepBoris(2, n=7)):
            pr = TP.TraceProblem(w.xv0[:64], (0.0, 300 * dtf * w.dt), prob.p)
            g = TP.solve(pr, alg, dt=dtf * w.dt, savestepinterval=30, form="reference")
            assert_bitwise(g, run_oracle(pr, alg, dt=dtf * w.dt, savestepinterval=30, form="reference"))


@pytest.mark.parametrize("ms", [None, (4, 2)])
def test_parity_adaptive_open_box(ms):
    w, A = W.turbulent_box(N=512, n=32, nsteps=1, dtype=np.float64, seed=11)
    L = w.extra["L"]
    ax = W.grid_axes(32, L, periodic=False)
    B = TP.GridField(A, *ax, bc=TP.FillExtrap())
    prob = TP.TraceProblem(w.xv0, (0.0, 400 * w.dt), (w.q2m, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    alg = TP.AdaptiveBoris(safety=0.1) if ms is None else TP.AdaptiveMultistepBoris[ms[0]](n=ms[1])
    bnd = TP.OutsideBox((0, 0, 0), (L, L, L))
    kw = dict(isoutside=bnd, savestepinterval=7, max_steps=2000, nout_cap=400)
    g = TP.solve(prob, alg, **kw)
    o = run_oracle(prob, alg, **kw)
    assert (g.retcodes == 1).any()          # some particles leave the box
    assert_bitwise(g, o)


@pytest.mark.parametrize("ppl", [1, 3, 8])
def test_adaptive_lane_refill_bitwise(ppl):
    # a wave works through 64 x ppl particles, a lane taking the next one as
    # soon as its own leaves the box or reaches the step cap; N = 1000 leaves
    # a ragged last wave at every ppl.  Results cannot depend on the hand-out.
    w, A = W.turbulent_box(N=1000, n=32, nsteps=1, dtype=np.float64, seed=12)
    L = w.extra["L"]
    ax = W.grid_axes(32, L, periodic=False)
    B = TP.GridField(A, *ax, bc=TP.FillExtrap())
    prob = TP.TraceProblem(w.xv0, (0.0, 400 * w.dt), (w.q2m, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    alg = TP.AdaptiveBoris(safety=0.02)
    kw = dict(isoutside=TP.OutsideBox((0, 0, 0), (L, L, L)), savestepinterval=5, max_steps=40,
              nout_cap=16)
    g = TP.solve(prob, alg, particles_per_lane=ppl, **kw)
    # exits (most after a few steps) and step caps both occur: lives are uneven
    assert (g.retcodes == 1).any() and (g.retcodes == 2).any()
    assert_bitwise(g, run_oracle(prob, alg, **kw))
    f = TP.solve(prob, alg, precision="fp32c", particles_per_lane=ppl, **kw)
    f1 = TP.solve(prob, alg, precision="fp32c", particles_per_lane=1, **kw)
    np.testing.assert_array_equal(f.lengths, f1.lengths)
    np.testing.assert_array_equal(f.u_all, f1.u_all)
    np.testing.assert_array_equal(f.t_all, f1.t_all)


def test_particles_per_lane_checks():
    z = TP.ZeroField()
    prob = TP.TraceProblem(np.zeros((4, 6)) + [0, 0, 0, 1, 0, 0], (0.0, 1.0),
                           (1.0, 1.0, z, TP.UniformField([0, 0, 1.0]), z))
    with pytest.raises(TP.ArgumentError, match="adaptive solves only"):
        TP.solve(prob, TP.Boris(), dt=0.1, particles_per_lane=2)
    with pytest.raises(TP.ArgumentError, match="particles_per_lane"):
        TP.solve(prob, TP.AdaptiveBoris(), particles_per_lane=-1)
    # more lanes than particles: idle lanes, same answer
    a = TP.solve(prob, TP.AdaptiveBoris(), particles_per_lane=16)
    np.testing.assert_array_equal(a.u_all, TP.solve(prob, TP.AdaptiveBoris()).u_all)


def test_parity_affine_time_dependent_all_orders():
    rng = np.random.default_rng(5)
    N = 256
    xv0 = np.hstack([rng.normal(size=(N, 3)), rng.normal(size=(N, 3))])
    E = TP.AffineField([0.1, -0.2, 0.05], rng.normal(size=9) * 0.05, [0.01, 0.0, -0.02])
    B = TP.AffineField([0.0, 0.3, 1.0], rng.normal(size=9) * 0.1, [0.0, 0.05, 0.0])
    prob = TP.TraceProblem(xv0, (0.5, 20.5), (1.3, 1.0, E, B, TP.ZeroField()))
    for alg in (TP.Boris(), TP.MultistepBoris(2, n=3), TP.MultistepBoris(4, n=1),
                TP.MultistepBoris(6, n=2)):
        g = TP.solve(prob, alg, dt=0.05, savestepinterval=13)
        o = run_oracle(prob, alg, dt=0.05, savestepinterval=13)
        assert_bitwise(g, o)



@pytest.mark.parametrize("ft", [0.0, 0.05])
def test_parity_affine_adaptive_static_and_time_dependent(ft):
    # ft = 0: static fields, one evaluation per adaptive step serves every
    # update; ft != 0: the mid-step update evaluates at t + dt/2 on its own
    rng = np.random.default_rng(6)
    N = 256
    xv0 = np.hstack([rng.normal(size=(N, 3)), rng.normal(size=(N, 3))])
    E = TP.AffineField([0.1, -0.2, 0.05], rng.normal(size=9) * 0.05, [0.5 * ft, 0.0, -ft])
    B = TP.AffineField([0.0, 0.3, 1.0], rng.normal(size=9) * 0.1, [0.0, ft, 0.0])
    prob = TP.TraceProblem(xv0, (0.5, 20.5), (1.3, 1.0, E, B, TP.ZeroField()))
    kw = dict(savestepinterval=5, max_steps=3000, nout_cap=800)
    for alg in (TP.AdaptiveBoris(safety=0.05), TP.AdaptiveMultistepBoris[4](n=2)):
        g = TP.solve(prob, alg, **kw)
        o = run_oracle(prob, alg, **kw)
        assert_bitwise(g, o)
        g = TP.solve(prob, alg, save_everystep=False, **kw)
        o = run_oracle(prob, alg, save_everystep=False, **kw)
        assert_bitwise(g, o)

# ---------------------------------------------------------------------------
# FP32C: error against the FP64 path
# ---------------------------------------------------------------------------
def test_fp32c_sphere_tests_decided_in_float_agree_with_exact():
    # analytic fields: FP32C settles sphere tests in float away from the
    # radii; adding OutsideNaN (never true here) sends every step through the
    # exact double test instead.  Loss-cone and outer-L particles cross the
    # spheres (the oracle ends 125 / 766 / 554 of 2048 early in these cases).
    w = W.dipole_belt(N=2048, nsteps=2000)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    for bnd in (TP.InsideSphere(3 * TP.Re), TP.OutsideSphere(6 * TP.Re),
                TP.InsideSphere(3 * TP.Re) | TP.OutsideSphere(6.5 * TP.Re)):
        kw = dict(dts=w.dts, nsteps=w.nsteps, savestepinterval=100, precision="fp32c")
        a = TP.solve(prob, TP.Boris(), isoutside=bnd, **kw)
        b = TP.solve(prob, TP.Boris(), isoutside=bnd | TP.OutsideNaN(), **kw)
        assert (a.retcodes == 1).any() and (a.retcodes == 0).any()
        np.testing.assert_array_equal(a.retcodes, b.retcodes)
        np.testing.assert_array_equal(a.lengths, b.lengths)
        np.testing.assert_array_equal(a.u_all, b.u_all)
        np.testing.assert_array_equal(a.t_all, b.t_all)


def test_fp32c_uniform_gyration_error():
    w = W.uniform_gyration(N=1024)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.UniformField(w.extra["B"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, w.tspan, param)
    g64 = TP.solve(prob, TP.Boris(), dt=w.dt, save_everystep=False)
    g32 = TP.solve(prob, TP.Boris(), dt=w.dt, save_everystep=False, precision="fp32c")
    x64, x32 = g64.final_states()[:, :3], g32.final_states()[:, :3]
    disp = np.linalg.norm(x64 - w.xv0[:, :3], axis=1)
    err = np.linalg.norm(x32 - x64, axis=1) / np.maximum(disp, 1.0)
    print(f"fp32c uniform: median {np.median(err):.2e} p95 {np.percentile(err, 95):.2e} max {err.max():.2e}")
    assert np.median(err) < 1e-4


def test_fp32c_turbulent_error(turb64):
    # The turbulent box is chaotic (errors e-fold every ~10 steps), so FP32C
    # is held to (a) a tight bound over a few steps, and (b) at 200 steps, the
    # divergence of FP64 runs whose initial positions differ by one FP32 ulp
    # (the floor that rounding the input to float alone would set).
    w, A, ax, _, _ = turb64
    B64 = TP.GridField(A, *ax, bc=TP.PeriodicBC())
    B32 = TP.GridField(A.astype(np.float32), *ax, bc=TP.PeriodicBC())
    alg = TP.MultistepBoris(4, n=2)

    def final(B, prec, nsteps, xv0):
        E = TP.UxBField(w.extra["u"], B)
        prob = TP.TraceProblem(xv0, (0.0, nsteps * w.dt), (w.q2m, TP.mi, E, B, TP.ZeroField()))
        return TP.solve(prob, alg, dt=w.dt, save_everystep=False, precision=prec).final_states()

    def err(a, b):
        return np.linalg.norm(a[:, :3] - b[:, :3], axis=1) / np.linalg.norm(b[:, :3] - w.xv0[:, :3], axis=1)

    e10 = err(final(B32, "fp32c", 10, w.xv0), final(B64, "fp64", 10, w.xv0))
    print(f"fp32c turbulent 10 steps: median {np.median(e10):.2e} p95 {np.percentile(e10, 95):.2e}")
    assert np.median(e10) < 1e-5
    ref = final(B64, "fp64", 200, w.xv0)
    e32 = err(final(B32, "fp32c", 200, w.xv0), ref)
    xp = w.xv0.copy()
    xp[:, :3] *= 1.0 + 2.0 ** -24
    floor = err(final(B64, "fp64", 200, xp), ref)
    print(f"fp32c turbulent 200 steps: median {np.median(e32):.2e}; "
          f"FP64 one-float-ulp floor {np.median(floor):.2e}")
    assert np.median(e32) < 10 * np.median(floor)

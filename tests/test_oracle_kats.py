"""Pin the CPU oracle against the reference's own known answers.

Each test restates one testset of TestParticle.jl's test suite (the fixture
entry names the file:line) and runs it through oracle/tp_oracle.c.  Where the
reference only states a tolerance, so does the test; the two Basic-Boris
answers and the time-varying-B answer are checked bit-exactly, which pins the
(x_n, t_{n+1/2}) field-time convention and the save schedule.
"""
import math

import numpy as np
import pytest


def q2m_of(kats, species):
    c = kats["constants"]
    return c["qe"] / c["me"] if species == "electron" else c["qi"] / c["mi"]


def test_basic_boris_bitexact(kats, orc):
    k = kats["basic_boris"]
    r = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m_of(kats, "electron"),
                  tspan=k["tspan"], dt=k["dt"], savestepinterval=k["savestepinterval"])
    t, u = r.traj(0)
    assert len(t) == 101 and r.retcode[0] == orc.RET_SUCCESS
    # the oracle reproduces test_boris.jl:51-54 to the last bit
    assert np.array_equal(u[-1], np.array(k["u_end"]))
    # sol(t): SciMLBase LinearInterpolation between saved records
    tq = k["interp_t"]
    j = int(np.searchsorted(t, tq))
    w = (tq - t[j - 1]) / (t[j] - t[j - 1])
    ui = u[j - 1] + w * (u[j] - u[j - 1])
    np.testing.assert_allclose(ui, k["u_interp"], rtol=k["rtol"], atol=1e-20)


def _prob_func_u0(u0, i):
    u = list(u0)
    u[4] = i * 1.0e5
    return u


def test_ensemble_sums(kats, orc):
    k = kats["basic_boris"]
    q2m = q2m_of(kats, "electron")
    for key, reduce in (("ensemble_threads_sum", lambda u: u[3]), ("ensemble_serial_sum", np.sum)):
        e = kats[key]
        xv0 = np.array([_prob_func_u0(k["u0"], i) for i in range(1, e["trajectories"] + 1)])
        r = orc.solve(xv0, orc.zero(), orc.uniform(k["B"]), q2m=q2m, tspan=k["tspan"],
                      dt=k["dt"], savestepinterval=e["savestepinterval"], nthreads=2)
        total = sum(reduce(r.traj(i)[1][-1]) for i in range(e["trajectories"]))
        assert total == pytest.approx(e["expected"], rel=1.5e-8)


def test_time_varying_B_bitexact(kats, orc):
    k = kats["time_varying_B"]
    r = orc.solve(k["u0"], orc.zero(), orc.kat(orc.FIELD_KAT_FRONT), q2m=q2m_of(kats, "electron"),
                  tspan=k["tspan"], dt=k["dt"], savestepinterval=k["savestepinterval"])
    assert r.traj(0)[1][-1][0] == k["x_end"]


def _alg_kwargs(orc, name, N, n, form):
    if name == "boris":
        return dict(alg=orc.ALG_BORIS)
    return dict(alg=orc.ALG_MULTISTEP, order_N=N, nsub=n, form=form)


@pytest.mark.parametrize("form", ["reference", "rational"])
def test_exb_drift(kats, orc, form):
    k = kats["exb_drift"]
    fm = orc.FORM_REFERENCE if form == "reference" else orc.FORM_RATIONAL
    for name, N, n in k["algs"]:
        r = orc.solve(k["u0"], orc.uniform(k["E"]), orc.uniform(k["B"]), q2m=k["q2m"],
                      tspan=k["tspan"], dt=k["dt"], **_alg_kwargs(orc, name, N, n, fm))
        u = r.traj(0)[1][-1]
        assert abs(u[0] - k["x_end"]) <= k["atol"], (name, N, n)
        assert abs(u[3] - k["vx_end"]) <= k["atol"], (name, N, n)


@pytest.mark.parametrize("form", ["reference", "rational"])
def test_gyration_closure(kats, orc, form):
    k = kats["gyration"]
    fm = orc.FORM_REFERENCE if form == "reference" else orc.FORM_RATIONAL
    for name, N, n in k["algs"]:
        r = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=k["q2m"], tspan=k["tspan"],
                      dt=k["dt"], **_alg_kwargs(orc, name, N, n, fm))
        u = r.traj(0)[1][-1]
        assert math.hypot(*u[:3]) < k["max_radius_end"], (name, N, n)
        assert abs(math.hypot(*u[3:]) - 1.0) <= k["speed_atol"], (name, N, n)


def test_adaptive_gradient_field(kats, orc):
    k = kats["adaptive_gradient"]
    c = kats["constants"]
    q2m = q2m_of(kats, "electron")
    tperiod = abs(2 * math.pi / (c["qe"] * k["B_ref"] / c["me"]))
    E = orc.uniform([0.0, 1.0e5, 0.0])
    B = orc.affine([0.0, 0.0, 0.01], jac=[0, 0, 0, 0, 0, 0, 0.01, 0, 0])
    r = orc.solve(k["u0"], E, B, q2m=q2m, tspan=(0.0, k["n_periods"] * tperiod),
                  adaptive=True, safety=k["safety"], nout_cap=200000)
    t, u = r.traj(0)
    assert r.len[0] <= 200000 and r.retcode[0] == orc.RET_SUCCESS
    dt_end = t[-2] - t[-3]
    dt_start = t[10] - t[9]
    assert dt_end < k["dt_ratio_max"] * dt_start

    def energy(s):
        return 0.5 * c["me"] * float(np.dot(s[3:], s[3:])) + c["qe"] * (-1.0e5 * s[1])

    assert energy(u[-1]) == pytest.approx(energy(u[0]), rel=k["energy_rtol"])


def test_nonzero_tspan(kats, orc):
    k = kats["nonzero_tspan"]
    q2m = q2m_of(kats, "proton")
    B = orc.kat(orc.FIELD_KAT_SWITCH_ON, v=k["B_on"], t_on=k["t_on"])
    r = orc.solve(k["u0"], orc.zero(), B, q2m=q2m, tspan=k["tspan"], dt=k["dt"])
    assert abs(r.traj(0)[1][-1][3]) < k["vx_bound"]
    r = orc.solve(k["u0"], orc.zero(), B, q2m=q2m, tspan=k["tspan"], adaptive=True, safety=0.1,
                  nout_cap=100000)
    assert abs(r.traj(0)[1][-1][3]) < k["vx_bound"]


def test_save_flags(kats, orc):
    k = kats["basic_boris"]
    dt = k["dt"]
    t0, t1 = k["tspan"]
    for cs in kats["save_flags"]["cases"]:
        r = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m_of(kats, "electron"),
                      tspan=k["tspan"], dt=dt, savestepinterval=cs["K"],
                      save_everystep=cs["save_everystep"], save_start=cs["save_start"],
                      save_end=cs["save_end"])
        t, u = r.traj(0)
        assert len(t) == cs["nout"] == orc.calculate_nout(1000, cs["K"], cs["save_start"],
                                                         cs["save_end"], cs["save_everystep"])
        ref = {"t0": t0, "t1": t1}
        if "t_first" in cs:
            assert t[0] == ref[cs["t_first"]]
        if "t_last" in cs:
            assert t[-1] == ref[cs["t_last"]]
        if "t_first_steps" in cs:
            assert t[0] == pytest.approx(t0 + cs["t_first_steps"] * dt)
            assert t[-1] == pytest.approx(t0 + cs["t_last_steps"] * dt)
        if "t_second_steps" in cs:
            assert t[1] == pytest.approx(t0 + cs["t_second_steps"] * dt)
            assert t[-2] == pytest.approx(t0 + cs["t_penultimate_steps"] * dt)


def test_backward_tracing_keeps_no_closing_record(kats, orc):
    # No reference test traces Boris backward; boris.jl:523-551 decides it.
    # With t1 < t0 and dt < 0 the closing record is taken only if
    # tsave[iout] < t_final, which never holds going backward, so the
    # solution has one record fewer than nout and the retcode is Terminated;
    # the same problem forward keeps the record at t1 and ends Success.
    k = kats["basic_boris"]
    q2m = q2m_of(kats, "electron")
    t0, t1 = k["tspan"]
    K = 10
    nout = orc.calculate_nout(1000, K, True, True, True)
    fwd = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m, tspan=(t0, t1), dt=k["dt"],
                    savestepinterval=K)
    back = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m, tspan=(t1, t0), dt=-k["dt"],
                     savestepinterval=K)
    assert fwd.len[0] == nout and fwd.retcode[0] == 0 and fwd.t[0, nout - 1] == t1
    assert back.len[0] == nout - 1 and back.retcode[0] == 1
    tb = back.t[0, :nout - 1]
    assert tb[0] == t1 and np.all(np.diff(tb) < 0) and tb[-1] > t0


def test_kernel_path_energy_and_saving(kats, orc):
    k = kats["kernel_path"]
    c = kats["constants"]
    q2m = q2m_of(kats, "proton")
    Tg = 2 * math.pi / (abs(c["qi"]) * k["B"][2] / c["mi"])
    e = k["energy"]
    r = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m,
                  tspan=(0.0, e["n_periods"] * Tg), dt=Tg / e["steps_per_period"],
                  savestepinterval=10)
    u = r.traj(0)[1][-1]
    assert math.sqrt(float(np.dot(u[3:], u[3:]))) == pytest.approx(1.0e5, rel=e["rtol"])
    r = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m, tspan=k["tspan"],
                  dt=k["dt"], save_everystep=False)
    t, _ = r.traj(0)
    assert list(t) == [k["tspan"][0], k["tspan"][1]]
    r = orc.solve(k["u0"], orc.zero(), orc.uniform(k["B"]), q2m=q2m, tspan=k["tspan"],
                  dt=k["dt"], save_everystep=False, save_start=False)
    assert list(r.traj(0)[0]) == [k["tspan"][1]]


def test_solver_limits(kats, orc):
    k = kats["solver_limits"]
    args = dict(q2m=q2m_of(kats, "electron"), tspan=k["tspan"])
    E = orc.uniform([0.0, 1.0e5, 0.0])
    B = orc.affine([0.0, 0.0, 0.01], jac=[0, 0, 0, 0, 0, 0, 0.01, 0, 0])
    with pytest.raises(ValueError, match="maxiters"):
        orc.solve(k["u0"], E, B, dt=k["dt"], maxiters=k["maxiters_error"], **args)
    with pytest.raises(ValueError, match="too small"):
        orc.solve(k["u0"], E, B, dt=np.finfo(np.float64).eps, **args)
    for N in k["bad_orders"]:
        with pytest.raises(ValueError):
            orc.solve(k["u0"], E, B, dt=k["dt"], alg=orc.ALG_MULTISTEP, order_N=N, nsub=2, **args)


def test_nan_boundary(kats, orc):
    k = kats["nan_boundary"]
    E = orc.kat(orc.FIELD_KAT_NAN_OUTSIDE, v=[0, 0, 0], radius=k["radius"])
    B = orc.kat(orc.FIELD_KAT_NAN_OUTSIDE, v=[0, 0, 1.0], radius=k["radius"])
    bnd = dict(bnd_flags=orc.BND_NAN | orc.BND_SPHERE_OUT, r_out=k["radius"])
    for kw in (dict(dt=k["dt"]), dict(adaptive=True, safety=k["adaptive_safety"], nout_cap=10000)):
        r = orc.solve(k["u0"], E, B, q2m=k["q2m"], tspan=k["tspan"], **bnd, **kw)
        u = r.traj(0)[1][-1]
        assert not np.any(np.isnan(u))
        assert math.hypot(*u[:3]) <= 1.0
        assert r.retcode[0] == orc.RET_TERMINATED


def test_spatial_rejection_and_integer_tspan(kats, orc):
    k = kats["spatial_rejection"]
    q2m = q2m_of(kats, "proton")
    B = orc.uniform([1.0, 0.0, 0.0])
    r = orc.solve(k["u0"], orc.zero(), B, q2m=q2m, tspan=k["tspan"], adaptive=True,
                  safety=k["safety"], nout_cap=100000, bnd_flags=orc.BND_BOX,
                  box_hi=(k["x_max"], np.inf, np.inf))
    t, u = r.traj(0)
    assert u[-1][0] <= k["x_max"] and t[-1] < k["tspan"][1]
    assert r.retcode[0] == orc.RET_TERMINATED
    k = kats["integer_tspan"]
    r = orc.solve(k["u0"], orc.zero(), B, q2m=q2m, tspan=k["tspan"], dt=k["dt"])
    t, _ = r.traj(0)
    assert t[-1] == 10 and r.retcode[0] == orc.RET_SUCCESS


def _kat_grid(k):
    nx, ny, nz = k["nx"], k["ny"], k["nz"]
    i = np.arange(1, nx + 1)[None, None, :]
    j = np.arange(1, ny + 1)[None, :, None]
    kk = np.arange(1, nz + 1)[:, None, None]
    n = (i + j + kk).astype(np.float32)
    data = np.repeat(n[..., None], 3, axis=3)
    spacing = [(k["hi"] - k["lo"]) / (m - 1) for m in (nx, ny, nz)]
    return data, [k["lo"]] * 3, spacing


def test_interpolation_kats(kats, orc):
    k = kats["interpolation"]
    data, origin, spacing = _kat_grid(k)
    modes = {"fill_nan": orc.EXTRAP_FILL_NAN, "clamp": orc.EXTRAP_CLAMP,
             "periodic": orc.EXTRAP_PERIODIC}
    for cs in k["cases"]:
        f = orc.grid(data, origin, spacing, modes[cs["extrap"]])
        val = orc.field_eval(f, cs["x"])
        np.testing.assert_allclose(val, [cs["value"]] * 3, rtol=1.5e-8)
    # outside the domain, FillExtrap(NaN) gives NaN (prepare.jl default bc)
    f = orc.grid(data, origin, spacing, orc.EXTRAP_FILL_NAN)
    assert np.all(np.isnan(orc.field_eval(f, [20.0, 0.0, 0.0])))


def test_adaptive_multistep_consistency(kats, orc):
    k = kats["adaptive_multistep_consistency"]
    q2m = q2m_of(kats, "proton")
    B = orc.uniform(k["B"])
    base = dict(q2m=q2m, tspan=k["tspan"], adaptive=True, safety=k["safety"], nout_cap=100000)
    a = orc.solve(k["u0"], orc.zero(), B, **base)
    for form in (orc.FORM_REFERENCE, orc.FORM_RATIONAL):
        b = orc.solve(k["u0"], orc.zero(), B, alg=orc.ALG_MULTISTEP, order_N=2, nsub=1,
                      form=form, **base)
        ta, ua = a.traj(0)
        tb, ub = b.traj(0)
        assert len(ta) == len(tb)
        np.testing.assert_allclose(tb, ta, rtol=1.5e-8)
        np.testing.assert_allclose(ub, ua, rtol=1.5e-8, atol=1e-12)

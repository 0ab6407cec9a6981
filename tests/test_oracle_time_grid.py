"""The oracle's time-dependent gridded field (LazyTimeInterpolator,
src/utility/interpolation.jl:337-430), pinned by the reference's own
assertions (test/test_utility.jl:312-404, test/test_derivatives.jl:57-84), by
central differences in space and time of the oracle's own evaluation, and by
special cases that reduce to the static grid (one snapshot, t on a snapshot
time)."""
import numpy as np
import pytest

RNG = np.random.default_rng(77)


def _linear_snapshots(times, n=5, L=4.0):
    """Snapshots of E(x, t_k) = x t_k on an n^3 grid over [0, L]^3: trilinear
    interpolation reproduces a linear field, so these stand in for the
    reference's analytic loader x -> x * t_k (test_utility.jl:325-333)."""
    g = np.linspace(0.0, L, n)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([X, Y, Z], axis=-1)                     # (nz, ny, nx, 3)
    return np.stack([base * tk for tk in times]), (0.0, 0.0, 0.0), (L / (n - 1),) * 3


def test_reference_linear_loader_kats(orc):
    # test_utility.jl:312-358
    times = [0.0, 1.0, 2.0]
    snaps, o, h = _linear_snapshots(times)
    itp = orc.time_grid(snaps, times, o, h, orc.EXTRAP_FILL_NAN)
    x = [1.0, 2.0, 3.0]
    ev = lambda t: orc.field_eval(itp, x, t)
    np.testing.assert_allclose(ev(0.0), [0.0, 0.0, 0.0], atol=1e-14)
    np.testing.assert_allclose(ev(1.0), [1.0, 2.0, 3.0], rtol=1e-14)
    np.testing.assert_allclose(ev(0.5), [0.5, 1.0, 1.5], rtol=1e-14)
    np.testing.assert_allclose(ev(1.5), [1.5, 3.0, 4.5], rtol=1e-14)
    np.testing.assert_array_equal(ev(3.0), ev(2.0))        # clamp to the end
    np.testing.assert_array_equal(ev(-1.0), ev(0.0))       # clamp to the start


def test_reference_numerical_field_kats(orc):
    # test_utility.jl:359-404: B0 = (1, 0, 0), B1 = (0, 2, 0) on a 4^3 grid
    B0 = np.zeros((4, 4, 4, 3))
    B1 = np.zeros((4, 4, 4, 3))
    B0[..., 0] = 1.0
    B1[..., 1] = 2.0
    itp = orc.time_grid(np.stack([B0, B1]), [0.0, 1.0], (0.0,) * 3, (1.0 / 3,) * 3,
                        orc.EXTRAP_FILL_NAN)
    pt = [0.5, 0.5, 0.5]
    for t, want in ((0.0, [1.0, 0.0, 0.0]), (1.0, [0.0, 2.0, 0.0]), (0.5, [0.5, 1.0, 0.0]),
                    (0.25, [0.75, 0.5, 0.0])):
        np.testing.assert_allclose(orc.field_eval(itp, pt, t), want, rtol=1e-15, atol=1e-15)


def test_reference_derivative_kats(orc):
    # test_derivatives.jl:57-84: loader x -> x t_k, at x = (1, 2, 3), t = 0.5
    times = [0.0, 1.0, 2.0]
    snaps, o, h = _linear_snapshots(times)
    itp = orc.time_grid(snaps, times, o, h, orc.EXTRAP_FILL_NAN)
    x = [1.0, 2.0, 3.0]
    _, J, dBdt = orc.field_jac(itp, x, 0.5)
    np.testing.assert_allclose(J, 0.5 * np.eye(3), atol=1e-14)      # d(x t)/dx = t I
    np.testing.assert_allclose(dBdt, x, rtol=1e-14)                 # d(x t)/dt = x
    # out of the time span: the Jacobian clamps, the time derivative is zero
    np.testing.assert_array_equal(orc.field_jac(itp, x, -1.0)[1], orc.field_jac(itp, x, 0.0)[1])
    np.testing.assert_array_equal(orc.field_jac(itp, x, 3.0)[1], orc.field_jac(itp, x, 2.0)[1])
    np.testing.assert_array_equal(orc.field_jac(itp, x, -1.0)[2], np.zeros(3))
    np.testing.assert_array_equal(orc.field_jac(itp, x, 3.0)[2], np.zeros(3))


def _random_time_grid(orc, extrap, n=(7, 6, 5), nsnap=4, dtype=np.float64):
    snaps = RNG.normal(size=(nsnap, n[2], n[1], n[0], 3)).astype(dtype)
    times = np.cumsum(RNG.uniform(0.3, 1.5, size=nsnap)) - 0.7
    o = RNG.normal(size=3)
    h = RNG.uniform(0.5, 2.0, size=3)
    return orc.time_grid(snaps, times, o, h, extrap), snaps, times, o, h


@pytest.mark.parametrize("extrap", [0, 1, 2])
def test_snapshot_times_and_single_snapshot_equal_static_grid(orc, extrap):
    itp, snaps, times, o, h = _random_time_grid(orc, extrap)
    grids = [orc.grid(s, o, h, extrap) for s in snaps]
    for _ in range(50):
        x = o + RNG.uniform(0.1, 0.9, size=3) * np.array([6, 5, 4]) * h
        for k, tk in enumerate(times):
            # t on a snapshot time: w = 0 selects that snapshot exactly
            np.testing.assert_array_equal(orc.field_eval(itp, x, tk), orc.field_eval(grids[k], x))
        np.testing.assert_array_equal(orc.field_eval(itp, x, times[0] - 5.0),
                                      orc.field_eval(grids[0], x))
        np.testing.assert_array_equal(orc.field_eval(itp, x, times[-1] + 5.0),
                                      orc.field_eval(grids[-1], x))
    one = orc.time_grid(snaps[:1], times[:1], o, h, extrap)
    for t in (-1e9, times[0], 3.7):
        x = o + RNG.uniform(-2, 8, size=3) * h
        a, b = orc.field_eval(one, x, t), orc.field_eval(grids[0], x)
        np.testing.assert_array_equal(a, b)
        Ja, Jb = orc.field_jac(one, x, t), orc.field_jac(grids[0], x, t)
        for u, v in zip(Ja, Jb):
            np.testing.assert_array_equal(u, v)


def test_linear_in_time_and_derivatives_by_differences(orc):
    itp, snaps, times, o, h = _random_time_grid(orc, 2)
    for _ in range(40):
        x = o + RNG.uniform(0.05, 0.95, size=3) * np.array([6, 5, 4]) * h
        k = RNG.integers(1, len(times))
        t = times[k - 1] + RNG.uniform(0.2, 0.8) * (times[k] - times[k - 1])
        _, J, dBdt = orc.field_jac(itp, x, t)
        # within one interval the field is linear in t: a central difference is exact
        dt = 1e-3 * (times[k] - times[k - 1])
        fd_t = (orc.field_eval(itp, x, t + dt) - orc.field_eval(itp, x, t - dt)) / (2 * dt)
        np.testing.assert_allclose(dBdt, fd_t, rtol=1e-7, atol=1e-9)
        # the spatial Jacobian against central differences inside a cell
        fd = np.zeros((3, 3))
        for j in range(3):
            e = np.zeros(3)
            e[j] = 1e-6 * h[j]
            fd[:, j] = (orc.field_eval(itp, x + e, t) - orc.field_eval(itp, x - e, t)) / (2 * e[j])
        np.testing.assert_allclose(J, fd, rtol=1e-6, atol=1e-8)
        # the midpoint of an interval is the mean of the two snapshots' values
        tm = 0.5 * (times[k - 1] + times[k])
        ga = orc.grid(snaps[k - 1], o, h, 2)
        gb = orc.grid(snaps[k], o, h, 2)
        np.testing.assert_allclose(orc.field_eval(itp, x, tm),
                                   0.5 * (orc.field_eval(ga, x) + orc.field_eval(gb, x)),
                                   rtol=1e-13, atol=1e-14)


def test_time_grid_solve_matches_static_for_one_snapshot(orc):
    snaps = RNG.normal(size=(1, 8, 8, 8, 3)) * 1e-8
    o, h = (0.0, 0.0, 0.0), (1e6 / 8,) * 3
    tg = orc.time_grid(snaps, [3.0], o, h, orc.EXTRAP_PERIODIC)
    g = orc.grid(snaps[0], o, h, orc.EXTRAP_PERIODIC)
    xv0 = np.hstack([RNG.uniform(0, 1e6, size=(20, 3)), RNG.normal(size=(20, 3)) * 1e5])
    kw = dict(q2m=9.58e7, tspan=(0.0, 5.0), dt=0.01, savestepinterval=50)
    a = orc.solve(xv0, orc.zero(), tg, **kw)
    b = orc.solve(xv0, orc.zero(), g, **kw)
    np.testing.assert_array_equal(a.u, b.u)

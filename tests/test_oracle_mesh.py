"""The oracle's general meshes (rectilinear axes, cubic order 3, spherical
StructuredGrid, 2D/1D fields; src/utility/interpolation.jl:15-117,140-178,
237-319), pinned by the reference's own assertions (test/test_utility.jl:
95-165,261-311; test/runtests.jl:150-183), by reduction to the pinned uniform
grid, by exact reproduction of multilinear (order 1) and quadratic (order 3,
interior cells) data, and by central differences for the Jacobians."""
import numpy as np
import pytest

RNG = np.random.default_rng(2024)
FILL, CLAMP, PER, PER_INCL = 0, 1, 2, 3


def _nodes(A):
    """(3, n1, n2, n3) reference layout -> the oracle's (n3, n2, n1, 3)."""
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).transpose(3, 2, 1, 0))


def _mesh(orc, A, axes, coord=0, order=1, bc=(FILL,) * 3):
    return orc.mesh(_nodes(A), axes, coord, order, bc)


def test_reference_cartesian_kats(orc):
    # test_utility.jl:95-121: n = i + j + k on ranges of 4, 6, 8 nodes (component 1)
    x, y, z = np.linspace(-10, 10, 4), np.linspace(-10, 10, 6), np.linspace(-10, 10, 8)
    I, J, K = np.meshgrid(np.arange(1, 5), np.arange(1, 7), np.arange(1, 9), indexing="ij")
    A = np.zeros((3, 4, 6, 8))
    A[0] = I + J + K
    for order, bc, pt, want in [(1, FILL, (9, 0, 0), 11.85), (1, PER, (20, 0, 0), 9.5),
                                (1, CLAMP, (20, 0, 0), 12.0), (3, FILL, (9, 0, 0), 11.85),
                                (3, PER, (20, 0, 0), 9.25), (3, CLAMP, (20, 0, 0), 12.0)]:
        f = _mesh(orc, A, (x, y, z), order=order, bc=(bc,) * 3)
        np.testing.assert_allclose(orc.field_eval(f, np.array(pt, float))[0], want, rtol=1e-14)


def test_reference_nonuniform_kats(orc):
    # test_utility.jl:261-311 (RectilinearGrid)
    g = np.linspace(0.0, 10.0, 11)
    B = np.zeros((3, 11, 11, 11))
    B[0] = 1.0
    np.testing.assert_allclose(orc.field_eval(_mesh(orc, B, (g, g, g)), [5.0, 5.0, 5.0]), [1, 0, 0])
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    A = np.zeros((3, 11, 11, 11))
    A[0] = I + J + K
    f = _mesh(orc, A, (g, g, g))
    np.testing.assert_allclose(orc.field_eval(f, [5.0, 5.0, 5.0])[0], 15.0, rtol=1e-14)
    np.testing.assert_allclose(orc.field_eval(f, [5.5, 5.5, 5.5])[0], 16.5, rtol=1e-14)
    nu = np.array([0.0, 1.0, 4.0, 9.0])
    I, J, K = np.meshgrid(nu, nu, nu, indexing="ij")
    A = np.zeros((3, 4, 4, 4))
    A[0] = np.sqrt(I) + np.sqrt(J) + np.sqrt(K)
    f = _mesh(orc, A, (nu, nu, nu))
    np.testing.assert_allclose(orc.field_eval(f, [4.0, 4.0, 4.0])[0], 6.0, rtol=1e-14)
    np.testing.assert_allclose(orc.field_eval(f, [1.0, 4.0, 9.0])[0], 6.0, rtol=1e-14)


def _radial_B(r, th, ph):
    B = np.zeros((3, len(r), len(th), len(ph)))
    B[0] = 1.0
    return B


@pytest.mark.parametrize("order", [1, 3])
@pytest.mark.parametrize("log_r", [False, True])
def test_reference_spherical_kats(orc, order, log_r):
    # test_utility.jl:124-165: B_r = 1 everywhere -> the unit radial vector
    r = np.geomspace(1.0, 10.0, 11) if log_r else np.linspace(0.1, 10, 11)
    th, ph = np.linspace(0, np.pi, 11), np.linspace(0, 2 * np.pi, 11)
    f = _mesh(orc, _radial_B(r, th, ph), (r, th, ph), coord=1, order=order, bc=(FILL, FILL, PER_INCL))
    np.testing.assert_allclose(orc.field_eval(f, [1.0, 1.0, 1.0]), [0.57735] * 3, atol=1e-5)
    assert np.all(np.isnan(orc.field_eval(f, [0.0, 0.0, 0.0])))        # r = 0 < r_1


def test_reference_spherical_dipole_axis_kat(orc):
    # runtests.jl:150-183: B = B0 z-hat in spherical components on a log r grid
    r = np.geomspace(0.1, 10.0, 4)
    th, ph = np.linspace(0, np.pi, 8), np.linspace(0, 2 * np.pi, 8)
    B = np.zeros((3, 4, 8, 8))
    for it, tv in enumerate(th):
        B[0, :, it, :] = 1e-8 * np.cos(tv)
        B[1, :, it, :] = -1e-8 * np.sin(tv)
    f = _mesh(orc, B, (r, th, ph), coord=1, bc=(FILL, FILL, PER_INCL))
    np.testing.assert_allclose(orc.field_eval(f, [1.0, 1.0, 1.0])[2], 9.888387888463716e-9, rtol=1e-14)


@pytest.mark.parametrize("bc", [FILL, CLAMP, PER])
def test_uniform_mesh_equals_uniform_grid(orc, bc):
    n = (7, 5, 6)
    A = RNG.normal(size=(3,) + n)
    o, h = RNG.normal(size=3), RNG.uniform(0.5, 2.0, size=3)
    axes = [o[d] + h[d] * np.arange(n[d]) for d in range(3)]
    f = _mesh(orc, A, axes, bc=(bc,) * 3)
    # a uniform axis takes its spacing as (x_N - x_1) / (N - 1)
    g = orc.grid(_nodes(A), o, [(a[-1] - a[0]) / (len(a) - 1) for a in axes], bc)
    for _ in range(200):
        x = o + RNG.uniform(-1.5, 1.2, size=3) * np.array(n) * h
        np.testing.assert_array_equal(orc.field_eval(f, x), orc.field_eval(g, x))
        a, b = orc.field_jac(f, x), orc.field_jac(g, x)
        np.testing.assert_array_equal(a[1], b[1])


def _rect_axes(n):
    return [np.cumsum(np.r_[RNG.uniform(-3, 3), RNG.uniform(0.2, 2.0, size=k - 1)]) for k in n]


@pytest.mark.parametrize("bc", [FILL, CLAMP, PER_INCL])
def test_rectilinear_reproduces_multilinear(orc, bc):
    n = (6, 7, 5)
    axes = _rect_axes(n)
    c = RNG.normal(size=(3, 8))
    X, Y, Z = np.meshgrid(*axes, indexing="ij")

    def ml(x, y, z, k):
        return (c[k, 0] + c[k, 1] * x + c[k, 2] * y + c[k, 3] * z + c[k, 4] * x * y + c[k, 5] * y * z
                + c[k, 6] * x * z + c[k, 7] * x * y * z)

    A = np.stack([ml(X, Y, Z, k) for k in range(3)])
    f = _mesh(orc, A, axes, bc=(bc,) * 3)
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    for _ in range(300):
        x = lo + RNG.uniform(0, 1, size=3) * (hi - lo)
        want = np.array([ml(*x, k) for k in range(3)])
        np.testing.assert_allclose(orc.field_eval(f, x), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # on the nodes the interpolant is the node value (an inclusive wrap maps
    # the last node onto the first)
    for _ in range(50):
        i = [RNG.integers(0, k - (bc == PER_INCL)) for k in n]
        x = np.array([axes[d][i[d]] for d in range(3)])
        np.testing.assert_allclose(orc.field_eval(f, x), A[:, i[0], i[1], i[2]], rtol=1e-13, atol=1e-13)


def test_rectilinear_outside_behaviour(orc):
    n = (5, 4, 6)
    axes = _rect_axes(n)
    A = RNG.normal(size=(3,) + n)
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    x = 0.5 * (lo + hi)
    out = x.copy()
    out[1] = hi[1] + 0.3
    assert np.all(np.isnan(orc.field_eval(_mesh(orc, A, axes, bc=(FILL,) * 3), out)))
    fc = _mesh(orc, A, axes, bc=(CLAMP,) * 3)
    edge = x.copy()
    edge[1] = hi[1]
    np.testing.assert_array_equal(orc.field_eval(fc, out), orc.field_eval(fc, edge))
    np.testing.assert_array_equal(orc.field_jac(fc, out)[1][:, 1], np.zeros(3))   # clamped: flat
    fp = _mesh(orc, A, axes, bc=(PER_INCL,) * 3)
    shifted = x + np.array([2, -1, 3]) * (hi - lo)
    np.testing.assert_allclose(orc.field_eval(fp, shifted), orc.field_eval(fp, x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("bc", [FILL, CLAMP, PER])
def test_cubic_reproduces_quadratics_and_nodes(orc, bc):
    n = (9, 8, 10)
    h = np.array([0.7, 1.3, 0.9])
    axes = [h[d] * np.arange(n[d]) for d in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    q = RNG.normal(size=(3, 3, 3))

    def quad(x, y, z, k):
        return (q[k, 0, 0] + q[k, 0, 1] * x + q[k, 0, 2] * x * x) * (q[k, 1, 0] + q[k, 1, 1] * y) + \
            q[k, 2, 0] * z + q[k, 2, 2] * z * z

    A = np.stack([quad(X, Y, Z, k) for k in range(3)])
    f = _mesh(orc, A, axes, order=3, bc=(bc,) * 3)
    for _ in range(200):
        if bc == PER:
            # a periodic wrap joins the ends: only node values are reproduced
            i = [RNG.integers(0, k) for k in n]
            x = np.array([axes[d][i[d]] for d in range(3)])
            np.testing.assert_allclose(orc.field_eval(f, x), A[:, i[0], i[1], i[2]], rtol=1e-12, atol=1e-12)
            continue
        # interior cells (full four-node stencil): Catmull-Rom is exact for quadratics
        x = np.array([h[d] * RNG.uniform(1.0, n[d] - 2.0) for d in range(3)])
        want = np.array([quad(*x, k) for k in range(3)])
        np.testing.assert_allclose(orc.field_eval(f, x), want, rtol=1e-11, atol=1e-11 * np.abs(want).max())
    # boundary cells reproduce linear data (the ghost node is the linear extrapolation)
    L = RNG.normal(size=(3, 4))
    Al = np.stack([L[k, 0] + L[k, 1] * X + L[k, 2] * Y + L[k, 3] * Z for k in range(3)])
    fl = _mesh(orc, Al, axes, order=3, bc=(CLAMP,) * 3)
    for _ in range(100):
        x = np.array([h[d] * RNG.uniform(0.0, n[d] - 1.0) for d in range(3)])
        want = L[:, 0] + L[:, 1:] @ x
        np.testing.assert_allclose(orc.field_eval(fl, x), want, rtol=1e-11, atol=1e-11)


def _fd_jac(orc, f, x, h):
    J = np.zeros((3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        J[:, j] = (orc.field_eval(f, x + e) - orc.field_eval(f, x - e)) / (2 * h)
    return J


@pytest.mark.parametrize("kind", ["rect1", "cubic", "sph1", "sph3", "two_d"])
def test_mesh_jacobians_by_central_differences(orc, kind):
    if kind == "rect1":
        axes = _rect_axes((6, 7, 5))
        f = _mesh(orc, RNG.normal(size=(3, 6, 7, 5)), axes, bc=(CLAMP,) * 3)
        box = [(a[0], a[-1]) for a in axes]
    elif kind == "cubic":
        axes = [0.8 * np.arange(8), 1.1 * np.arange(7), 0.9 * np.arange(9)]
        f = _mesh(orc, RNG.normal(size=(3, 8, 7, 9)), axes, order=3, bc=(PER,) * 3)
        box = [(a[0], a[-1]) for a in axes]
    elif kind == "two_d":
        axes = [_rect_axes((6,))[0], 0.5 * np.arange(7), np.array([0.0])]
        f = _mesh(orc, RNG.normal(size=(3, 6, 7, 1)), axes, bc=(FILL,) * 3)
        box = [(axes[0][0], axes[0][-1]), (0, 3.0), (-5.0, 5.0)]
    else:
        r = np.geomspace(1.0, 6.0, 9) if kind == "sph1" else np.linspace(1.0, 6.0, 9)
        th, ph = np.linspace(0, np.pi, 13), np.linspace(0, 2 * np.pi, 17)
        f = _mesh(orc, RNG.normal(size=(3, 9, 13, 17)), (r, th, ph), coord=1,
                  order=1 if kind == "sph1" else 3, bc=(FILL, FILL, PER_INCL))
        box = None
    checked = 0
    for _ in range(60):
        if box is None:
            rr, tt, pp = RNG.uniform(1.5, 5.5), RNG.uniform(0.3, 2.8), RNG.uniform(0, 2 * np.pi)
            x = rr * np.array([np.sin(tt) * np.cos(pp), np.sin(tt) * np.sin(pp), np.cos(tt)])
        else:
            x = np.array([RNG.uniform(a + 0.05 * (b - a), b - 0.05 * (b - a)) for a, b in box])
        _, J, dBdt = orc.field_jac(f, x)
        assert np.all(dBdt == 0)
        fd = _fd_jac(orc, f, x, 1e-7)
        # a step across a cell face sees the other cell's slope: skip those
        if not np.allclose(fd, _fd_jac(orc, f, x, 2e-7), rtol=1e-4, atol=1e-6):
            continue
        np.testing.assert_allclose(J, fd, rtol=1e-5, atol=1e-6 * max(1.0, np.abs(fd).max()))
        checked += 1
    assert checked >= 40


def test_2d_and_1d_fields_ignore_axes(orc):
    gx, gy = np.linspace(0, 4, 5), np.array([0.0, 0.5, 2.0, 3.0])
    A2 = RNG.normal(size=(3, 5, 4, 1))
    f2 = _mesh(orc, A2, (gx, gy, [0.0]))
    for _ in range(20):
        x = np.array([RNG.uniform(0, 4), RNG.uniform(0, 3), RNG.uniform(-100, 100)])
        np.testing.assert_array_equal(orc.field_eval(f2, x), orc.field_eval(f2, np.r_[x[:2], 0.0]))
    # 1D along y (dir = 2): values depend on y only and match linear interpolation
    g = np.array([0.0, 1.0, 3.0, 4.0])
    A1 = RNG.normal(size=(3, 1, 4, 1))
    f1 = _mesh(orc, A1, ([0.0], g, [0.0]), bc=(CLAMP,) * 3)
    for y in RNG.uniform(0, 4, size=20):
        want = np.array([np.interp(y, g, A1[c, 0, :, 0]) for c in range(3)])
        np.testing.assert_allclose(orc.field_eval(f1, [RNG.normal(), y, RNG.normal()]), want,
                                   rtol=1e-13, atol=1e-13)


def test_uniform_mesh_solve_equals_grid_solve(orc):
    n = 10
    A = RNG.normal(size=(3, n, n, n)) * 1e-8
    h = 1e5
    axes = [h * np.arange(n)] * 3
    f = _mesh(orc, A, axes, bc=(PER,) * 3)
    g = orc.grid(_nodes(A), (0, 0, 0), (h, h, h), PER)
    xv0 = np.hstack([RNG.uniform(0, n * h, size=(16, 3)), RNG.normal(size=(16, 3)) * 1e5])
    kw = dict(q2m=9.58e7, tspan=(0.0, 3.0), dt=0.01, savestepinterval=20)
    np.testing.assert_array_equal(orc.solve(xv0, orc.zero(), f, **kw).u, orc.solve(xv0, orc.zero(), g, **kw).u)

"""General meshes on the device (rectilinear, cubic order 3, spherical
StructuredGrid, 2D/1D fields; src/utility/interpolation.jl:15-117,140-178,
237-319): the reference's own assertions through the device evaluator;
Cartesian meshes bit-identical to the oracle in evaluations, solves and
records; spherical meshes within 1e-13 of the oracle (device acos / atan2 /
sincos against the host's libm)."""
import numpy as np
import pytest

import testparticle_amd as TP
from parity import assert_bitwise, assert_close, oracle_field, run_oracle

import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(99)


def test_reference_kats_on_device():
    # test_utility.jl:95-121
    x, y, z = np.linspace(-10, 10, 4), np.linspace(-10, 10, 6), np.linspace(-10, 10, 8)
    I, J, K = np.meshgrid(np.arange(1, 5), np.arange(1, 7), np.arange(1, 9), indexing="ij")
    A = np.zeros((3, 4, 6, 8))
    A[0] = I + J + K
    for order, bc, pt, want in [(3, TP.FillExtrap(), (9, 0, 0), 11.85),
                                (3, TP.PeriodicBC(endpoint=":exclusive"), (20, 0, 0), 9.25),
                                (3, TP.ClampExtrap(), (20, 0, 0), 12.0)]:
        f = TP.build_interpolator(A, x, y, z, order, bc)
        assert isinstance(f, TP.MeshField)
        np.testing.assert_allclose(f(np.array(pt, float))[0], want, rtol=1e-14)
    # test_utility.jl:261-311
    nu = np.array([0.0, 1.0, 4.0, 9.0])
    Ii, Jj, Kk = np.meshgrid(nu, nu, nu, indexing="ij")
    An = np.zeros((3, 4, 4, 4))
    An[0] = np.sqrt(Ii) + np.sqrt(Jj) + np.sqrt(Kk)
    fn = TP.build_interpolator("RectilinearGrid", An, nu, nu, nu)
    np.testing.assert_allclose(fn([4.0, 4.0, 4.0])[0], 6.0, rtol=1e-14)
    np.testing.assert_allclose(fn([1.0, 4.0, 9.0])[0], 6.0, rtol=1e-14)
    # runtests.jl:150-183
    r = np.geomspace(0.1, 10.0, 4)
    th, ph = np.linspace(0, np.pi, 8), np.linspace(0, 2 * np.pi, 8)
    B = np.zeros((3, 4, 8, 8))
    for it, tv in enumerate(th):
        B[0, :, it, :] = 1e-8 * np.cos(tv)
        B[1, :, it, :] = -1e-8 * np.sin(tv)
    p = TP.prepare(r, th, ph, TP.ZeroField(), B, gridtype="StructuredGrid")
    np.testing.assert_allclose(p[3]([1.0, 1.0, 1.0])[2], 9.888387888463716e-9, rtol=1e-14)
    # test_utility.jl:124-165
    r2 = np.linspace(0.1, 10, 11)
    th2, ph2 = np.linspace(0, np.pi, 11), np.linspace(0, 2 * np.pi, 11)
    Br = np.zeros((3, 11, 11, 11))
    Br[0] = 1.0
    for order in (1, 3):
        fs = TP.build_interpolator("StructuredGrid", Br, r2, th2, ph2, order)
        np.testing.assert_allclose(fs([1.0, 1.0, 1.0]), [0.57735] * 3, atol=1e-5)
        assert np.all(np.isnan(fs([0.0, 0.0, 0.0])))


def _rect_axes(n, scale):
    return [scale * np.cumsum(np.r_[0.0, RNG.uniform(0.3, 1.7, size=k - 1)]) for k in n]


def _meshes():
    """(name, field, sampling box) for every Cartesian mesh variant"""
    out = []
    n = (9, 8, 10)
    ax = _rect_axes(n, 1e6)
    A = RNG.normal(size=(3,) + n) * 3e-9 + np.array([0, 0, 1e-8])[:, None, None, None]
    for bc in (TP.FillExtrap(), TP.ClampExtrap(), TP.PeriodicBC(endpoint=":inclusive")):
        out.append((f"rect-{type(bc).__name__}", TP.build_interpolator("RectilinearGrid", A, *ax, bc=bc), ax))
    ua = [1.1e6 * np.arange(k) for k in n]
    for bc in (TP.FillExtrap(), TP.ClampExtrap(), TP.PeriodicBC()):
        out.append((f"cubic-{type(bc).__name__}", TP.build_interpolator(A, *ua, order=3, bc=bc), ua))
    A32 = A.astype(np.float32)
    out.append(("rect-f32", TP.build_interpolator("RectilinearGrid", A32, *ax, bc=TP.ClampExtrap()), ax))
    A2 = A[:, :, :, 0]
    out.append(("two_d", TP.build_interpolator(A2, ax[0], ax[1], bc=TP.ClampExtrap()),
                [ax[0], ax[1], np.array([-1e6, 1e6])]))
    A1 = A[:, 0, :, 0]
    out.append(("one_d", TP.build_interpolator(A1, ax[1], dir=2), [np.array([-1e6, 1e6]), ax[1],
                                                                   np.array([-1e6, 1e6])]))
    return out


MESHES = _meshes()


@pytest.mark.parametrize("idx", range(len(MESHES)), ids=[m[0] for m in MESHES])
def test_cartesian_mesh_eval_and_jacobian_bitwise(idx):
    name, f, ax = MESHES[idx]
    lo = np.array([a[0] for a in ax])
    hi = np.array([a[-1] for a in ax])
    x = lo + RNG.uniform(-0.1, 1.1, size=(400, 3)) * (hi - lo)
    got = f(x)
    orc = oracle_field(f)
    want = np.array([O.field_eval(orc, xi) for xi in x])
    np.testing.assert_array_equal(got, want)
    # Jacobians through the record evaluator (the save_work path)
    prob = TP.TraceProblem(np.zeros(6), (0.0, 1.0), (TP.qi / TP.mi, TP.mi, TP.ZeroField(), f, TP.ZeroField()))
    xv = np.hstack([x, RNG.normal(size=(400, 3)) * 1e5])
    r = TP.record_eval(prob, xv, np.zeros(400))
    ro = O.record_eval(O.zero(), orc, q2m=TP.qi / TP.mi, m=TP.mi, xv=xv, t=np.zeros(400))
    np.testing.assert_array_equal(r, ro)


@pytest.mark.parametrize("idx", [0, 1, 2, 4, 5, 7], ids=[MESHES[i][0] for i in (0, 1, 2, 4, 5, 7)])
def test_cartesian_mesh_solves_bitwise(idx):
    name, B, ax = MESHES[idx]
    lo = np.array([a[0] for a in ax])
    hi = np.array([a[-1] for a in ax])
    N = 192
    x0 = lo + RNG.uniform(0.2, 0.8, size=(N, 3)) * (hi - lo)
    v0 = RNG.normal(size=(N, 3)) * 2e5
    E = TP.UxBField([1e5, 0.0, 0.0], B)
    prob = TP.TraceProblem(np.hstack([x0, v0]), (0.0, 20.0), (TP.qi / TP.mi, TP.mi, E, B, TP.ZeroField()))
    kw = dict(dt=0.05, savestepinterval=40)
    for alg in (TP.Boris(), TP.MultistepBoris(4, n=2)):
        assert_bitwise(TP.solve(prob, alg, **kw), run_oracle(prob, alg, **kw))
    kr = dict(kw, save_fields=True, save_work=True)
    assert_bitwise(TP.solve(prob, TP.Boris(), **kr), run_oracle(prob, TP.Boris(), **kr))
    pa = TP.TraceProblem(prob.u0[:48], (0.0, 20.0), prob.p)
    ka = dict(savestepinterval=8, max_steps=3000, nout_cap=600)
    assert_bitwise(TP.solve(pa, TP.AdaptiveBoris(safety=0.05), **ka),
                   run_oracle(pa, TP.AdaptiveBoris(safety=0.05), **ka))
    # OutsideGrid: the node box of a Cartesian mesh
    if name.startswith("rect") or name.startswith("cubic"):
        ko = dict(kw, isoutside=TP.OutsideGrid())
        assert_bitwise(TP.solve(prob, TP.Boris(), **ko), run_oracle(prob, TP.Boris(), **ko))


def test_fp32c_and_batch_with_meshes():
    _, B, ax = MESHES[1]
    lo = np.array([a[0] for a in ax])
    hi = np.array([a[-1] for a in ax])
    N = 256
    xv = np.hstack([lo + RNG.uniform(0.3, 0.7, size=(N, 3)) * (hi - lo), RNG.normal(size=(N, 3)) * 2e5])
    prob = TP.TraceProblem(xv, (0.0, 2.0), (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    kw = dict(dt=0.05, savestepinterval=1)
    a = TP.solve(prob, TP.Boris(), precision="fp32c", **kw)
    b = TP.solve(prob, TP.Boris(), **kw)
    # 40 steps: FP32C keeps ~1e-7 relative positions; fields are evaluated in FP64
    np.testing.assert_allclose(a.u_all[:, :, :3], b.u_all[:, :, :3], rtol=0, atol=1e-6 * (hi - lo).max())
    np.testing.assert_allclose(a.u_all[:, :, 3:], b.u_all[:, :, 3:], rtol=0, atol=1e-4 * 2e5 * 4)
    # a batch mixing a uniform grid (first instance) and a mesh
    g = np.linspace(0, 1e7, 12, endpoint=False)
    Ag = RNG.normal(size=(3, 12, 12, 12)) * 3e-9
    Bg = TP.GridField(Ag, g, g, g, bc=TP.PeriodicBC())
    pg = TP.TraceProblem(np.hstack([RNG.uniform(0, 1e7, size=(100, 3)), RNG.normal(size=(100, 3)) * 2e5]),
                         (0.0, 2.0), (TP.qi / TP.mi, TP.mi, TP.ZeroField(), Bg, TP.ZeroField()))
    sols = TP.solve_batch([pg, prob], TP.Boris(), **kw)
    for pr, s in zip((pg, prob), sols):
        np.testing.assert_array_equal(s.u_all, TP.solve(pr, TP.Boris(), **kw).u_all)


@pytest.mark.parametrize("order", [1, 3])
def test_spherical_mesh_close_to_oracle(order):
    # cardinal cubics need uniform axes: log-spaced r at order 1 only
    r = np.geomspace(1.0, 8.0, 12) if order == 1 else np.linspace(1.0, 8.0, 12)
    th, ph = np.linspace(0, np.pi, 15), np.linspace(0, 2 * np.pi, 19)
    A = RNG.normal(size=(3, 12, 15, 19))
    f = TP.build_interpolator("StructuredGrid", A, r, th, ph, order)
    orc = oracle_field(f)
    rr = RNG.uniform(0.5, 9.0, size=500)
    d = RNG.normal(size=(500, 3))
    x = rr[:, None] * d / np.linalg.norm(d, axis=1)[:, None]
    got = f(x)
    want = np.array([O.field_eval(orc, xi) for xi in x])
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=0, atol=1e-13 * np.abs(want[ok]).max())
    # short solves: device transcendental functions differ from libm in the last bits
    x0 = x[(rr > 2) & (rr < 6)][:64]
    prob = TP.TraceProblem(np.hstack([x0, RNG.normal(size=(len(x0), 3)) * 1e-1]), (0.0, 1.0),
                           (1.0, 1.0, TP.ZeroField(), f, TP.ZeroField()))
    kw = dict(dt=1e-3, savestepinterval=100)
    assert_close(TP.solve(prob, TP.Boris(), **kw), run_oracle(prob, TP.Boris(), **kw), rtol=1e-10,
                 atol_pos=1e-12, atol_vel=1e-12)

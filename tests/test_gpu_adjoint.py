"""Adjoint (reverse-mode) gradients of the trajectory loss
L = (1/N) sum |x_i(T) - x*_i|^2 through the fixed-step Boris pusher, checked
against forward-mode dual numbers through the CPU oracle (dL/d(x0, v0)) and
against central finite differences of the FP64 device solve (dL/dB)."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W

import oracle as O

pytestmark = pytest.mark.gpu

L_BOX = 1.0e7


def _setup(n=16, N=256, nsteps=128, seed=5, dB=0.3):
    A = W.turbulent_field(n, L_BOX, 5e-9, dB, seed, dtype=np.float64)
    ax = W.grid_axes(n, L_BOX)
    rng = np.random.default_rng(seed + 1)
    x = rng.uniform(0, L_BOX, size=(N, 3))
    v = W.isotropic_directions(rng, N) * np.sqrt(2 * 1.0e4 * TP.qi / TP.mi)
    xv0 = np.hstack([x, v])
    Tg = 2 * np.pi * TP.mi / (TP.qi * 5e-9)
    dt = Tg / 40
    targets = rng.uniform(0, L_BOX, size=(N, 3))
    return A, ax, xv0, dt, (0.0, nsteps * dt), targets


def _prob(A, ax, xv0, tspan, storage="f64"):
    # the FP32C adjoint gathers from float node storage
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC(), storage=storage)
    return TP.TraceProblem(xv0, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField())), B


def _loss_fp64(A, ax, xv0, tspan, dt, targets):
    prob, _ = _prob(A, ax, xv0, tspan)
    xT = TP.solve(prob, TP.Boris(), dt=dt, save_everystep=False).final_states()[:, :3]
    return float(np.mean(np.sum((xT - targets) ** 2, axis=1)))


def _final64(A, ax, xv0, tspan, dt):
    prob, _ = _prob(A, ax, xv0, tspan)
    return TP.solve(prob, TP.Boris(), dt=dt, save_everystep=False).final_states()[:, :3]


def _fd_node(A, ax, xv0, tspan, dt, targets, c, i, j, k, h=1e-6 * 5e-9):
    """Central difference of the FP64 device loss at one grid value, formed
    per particle: L+ - L- = mean_i (x+_i - x-_i) . (x+_i + x-_i - 2 x*_i).
    Differencing the loss itself (~1e16 m^2, a sum of N squares) would lose
    most digits of the change to rounding."""
    Ap, Am = A.copy(), A.copy()
    Ap[c, i, j, k] += h
    Am[c, i, j, k] -= h
    xp, xm = _final64(Ap, ax, xv0, tspan, dt), _final64(Am, ax, xv0, tspan, dt)
    return float(np.mean(np.sum((xp - xm) * (xp + xm - 2 * targets), axis=1))) / (2 * h)


def _oracle_grid(B):
    return O.grid(B.data.astype(np.float64), B.origin, B.spacing, O.EXTRAP_PERIODIC)


def test_forward_loss_matches_solver():
    A, ax, xv0, dt, tspan, tg = _setup()
    prob, _ = _prob(A, ax, xv0, tspan)
    r = TP.loss_gradient(prob, dt=dt, targets=tg, precision="fp64", ckpt_interval=32)
    assert r.loss == pytest.approx(_loss_fp64(A, ax, xv0, tspan, dt, tg), rel=1e-13)


# FP32 rounding grows along the chaotic orbits, so the FP32C cases run fewer steps
@pytest.mark.parametrize("precision,rtol,nsteps", [("fp64", 1e-8, 128), ("fp32c", 1e-4, 48)])
def test_initial_state_gradient_vs_oracle_duals(precision, rtol, nsteps):
    A, ax, xv0, dt, tspan, tg = _setup(N=64, nsteps=nsteps)
    prob, B = _prob(A, ax, xv0, tspan, "f64" if precision == "fp64" else "f32")
    r = TP.loss_gradient(prob, dt=dt, targets=tg, precision=precision, ckpt_interval=16)
    Bo = O.grid(B.data.astype(np.float64), B.origin, B.spacing, O.EXTRAP_PERIODIC)
    N = xv0.shape[0]
    ref = np.zeros((N, 6))
    for i in range(N):
        xT, J = O.boris_tangent(xv0[i], Bo, q2m=TP.qi / TP.mi, tspan=tspan, dt=dt)
        ref[i] = 2.0 / N * (xT - tg[i]) @ J
    scale = np.abs(ref).max(axis=0)          # per-component scale (positions vs velocities)
    err = np.abs(r.dL_dxv0 - ref) / scale
    print(f"{precision}: max error of dL/d(x0,v0) relative to the component scale {err.max():.2e}")
    assert err.max() < rtol


@pytest.mark.parametrize("precision,rtol,nsteps", [("fp64", 1e-6, 64), ("fp32c", 1e-4, 48)])
def test_grid_gradient_vs_central_differences(precision, rtol, nsteps):
    # 32 touched grid values, central differences of the FP64 solve with
    # h = 1e-6 B0 (the metric's check), where the flow is still linear enough
    # for differences to hold (~1e-8 of the dual-number derivative here)
    A, ax, xv0, dt, tspan, tg = _setup(n=12, N=2048, nsteps=nsteps)
    prob, _ = _prob(A, ax, xv0, tspan, "f64" if precision == "fp64" else "f32")
    r = TP.loss_gradient(prob, dt=dt, targets=tg, precision=precision, ckpt_interval=24)
    rng = np.random.default_rng(11)
    g = np.abs(r.dL_dB)
    # 32 random grid values among those the particles actually touch
    cand = np.argwhere(g > 1e-3 * g.max())
    picks = cand[rng.choice(len(cand), 32, replace=False)]
    worst = 0.0
    for c, i, j, k in picks:
        fd = _fd_node(A, ax, xv0, tspan, dt, tg, c, i, j, k)
        worst = max(worst, abs(r.dL_dB[c, i, j, k] - fd) / abs(fd))
    print(f"{precision}: worst relative error of dL/dB vs central FD on 32 nodes {worst:.2e}")
    assert worst < rtol


@pytest.mark.parametrize("nsteps", [64, 512])
def test_grid_gradient_vs_oracle_duals(nsteps):
    # dL/dB against forward-mode duals through the oracle: single grid values
    # (32 nodes) and dense random directions <dL/dB, d>.  Duals stay exact
    # where central differences fail (past ~200 steps the flow is too
    # nonlinear for any difference step), so this pins the long horizons.
    A, ax, xv0, dt, tspan, tg = _setup(n=12, N=1024, nsteps=nsteps)
    prob, B = _prob(A, ax, xv0, tspan)
    r = TP.loss_gradient(prob, dt=dt, targets=tg, precision="fp64", ckpt_interval=32)
    Bo = _oracle_grid(B)
    g = r.dL_dB                                    # [c, i, j, k]
    gnode = g.transpose(3, 2, 1, 0).reshape(-1)    # node layout, component fastest
    cand = np.flatnonzero(np.abs(gnode) > 1e-3 * np.abs(gnode).max())
    flat = np.random.default_rng(13).choice(cand, 32, replace=False)
    Lo, dl = O.loss_tangent(xv0, tg, Bo, q2m=TP.qi / TP.mi, tspan=tspan, dt=dt, flat=flat)
    assert r.loss == pytest.approx(Lo, rel=1e-13)
    rel = np.abs(gnode[flat] - dl) / np.abs(dl)
    print(f"{nsteps} steps: dL/dB vs duals on 32 nodes, max rel {rel.max():.2e}")
    assert rel.max() < 1e-9
    dirs = np.random.default_rng(14).normal(size=(2, gnode.size))
    _L, dd = O.loss_tangent(xv0, tg, Bo, q2m=TP.qi / TP.mi, tspan=tspan, dt=dt, dense=dirs)
    np.testing.assert_allclose(dirs @ gnode, dd, rtol=1e-9)


def test_fp32c_adjoint_stays_finite_past_the_float_range():
    # 2,048 steps of the 64^3 turbulent box (the differentiation workload's
    # geometry, fewer particles): the exact gradient grows past 1e38, the
    # float range; the FP32C adjoint's scaled state keeps every entry finite,
    # with magnitudes like the FP64 adjoint's
    A, ax, xv0, dt, tspan, tg = _setup(n=64, N=4096, nsteps=2048)
    r64 = TP.loss_gradient(_prob(A, ax, xv0, tspan)[0], dt=dt, targets=tg, precision="fp64",
                           ckpt_interval=64)
    r32 = TP.loss_gradient(_prob(A, ax, xv0, tspan, "f32")[0], dt=dt, targets=tg,
                           precision="fp32c", ckpt_interval=64)
    big64 = np.abs(r64.dL_dB).max()
    print(f"max |dL/dxv0|: fp64 {np.abs(r64.dL_dxv0).max():.2e}, fp32c {np.abs(r32.dL_dxv0).max():.2e}; "
          f"max |dL/dB|: fp64 {big64:.2e}, fp32c {np.abs(r32.dL_dB).max():.2e}")
    assert big64 > 1e38
    for r in (r32, r64):
        assert np.all(np.isfinite(r.dL_dB)) and np.all(np.isfinite(r.dL_dxv0))
    # chaotic orbits: FP32C and FP64 follow different trajectories by now, but
    # the gradients grow alike (exp(lambda T), lambda a property of the flow)
    assert abs(np.log10(np.abs(r32.dL_dB).max() / big64)) < 10


def test_nonfinite_gradient_is_an_error():
    # a fill-NaN grid: particles that leave the box see NaN fields, the loss is
    # NaN, and the call reports it instead of returning NaN as a result
    n, N, nsteps = 8, 64, 200
    A = W.turbulent_field(n, L_BOX, 5e-9, 0.3, 9, dtype=np.float64)
    ax = W.grid_axes(n, L_BOX, periodic=False)
    rng = np.random.default_rng(10)
    xv0 = np.hstack([rng.uniform(0.1 * L_BOX, 0.9 * L_BOX, size=(N, 3)),
                     W.isotropic_directions(rng, N) * np.sqrt(2 * 1.0e6 * TP.qi / TP.mi)])
    dt = 2 * np.pi * TP.mi / (TP.qi * 5e-9) / 40
    B = TP.GridField(A, *ax, bc=TP.FillExtrap(np.nan))
    prob = TP.TraceProblem(xv0, (0.0, nsteps * dt), (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    for precision, storage in (("fp64", "f64"), ("fp32c", "f32")):
        Bs = TP.GridField(A, *ax, bc=TP.FillExtrap(np.nan), storage=storage)
        pr = TP.TraceProblem(xv0, prob.tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(), Bs, TP.ZeroField()))
        with pytest.raises(FloatingPointError):
            TP.loss_gradient(pr, dt=dt, targets=np.zeros((N, 3)), precision=precision)


def test_grid_gradient_many_bins_radix_path():
    # 132^3 periodic cells = 4,492 bins of 512 cells: past the 2,048 bins the
    # block-local partition holds, so the records take the radix sort
    A, ax, xv0, dt, tspan, tg = _setup(n=132, N=512, nsteps=24)
    prob, _ = _prob(A, ax, xv0, tspan)
    r = TP.loss_gradient(prob, dt=dt, targets=tg, precision="fp64", ckpt_interval=8)
    rng = np.random.default_rng(3)
    g = np.abs(r.dL_dB)
    cand = np.argwhere(g > 1e-2 * g.max())
    picks = cand[rng.choice(len(cand), 8, replace=False)]
    worst = 0.0
    for c, i, j, k in picks:
        fd = _fd_node(A, ax, xv0, tspan, dt, tg, c, i, j, k)
        worst = max(worst, abs(r.dL_dB[c, i, j, k] - fd) / abs(fd))
    print(f"radix path: worst relative error of dL/dB vs central FD on 8 nodes {worst:.2e}")
    assert worst < 1e-6


def test_fp32c_adjoint_rejects_double_storage():
    A, ax, xv0, dt, tspan, tg = _setup(N=8, nsteps=4)
    prob, _ = _prob(A, ax, xv0, tspan, "f64")
    with pytest.raises(NotImplementedError):
        TP.loss_gradient(prob, dt=dt, targets=tg, precision="fp32c")


def test_checkpoint_interval_does_not_change_the_gradient():
    A, ax, xv0, dt, tspan, tg = _setup(N=512, nsteps=100)
    prob, _ = _prob(A, ax, xv0, tspan)
    rs = [TP.loss_gradient(prob, dt=dt, targets=tg, precision="fp64", ckpt_interval=K)
          for K in (1, 7, 32, 100, 1000)]
    for r in rs[1:]:
        np.testing.assert_array_equal(r.dL_dxv0, rs[0].dL_dxv0)
        np.testing.assert_allclose(r.dL_dB, rs[0].dL_dB, rtol=1e-5, atol=1e-7 * np.abs(rs[0].dL_dB).max())


@pytest.mark.parametrize("precision,storage", [("fp64", "f64"), ("fp64", "f32"), ("fp32c", "f32")])
def test_grid_gradient_clamped_grid(precision, storage):
    # a non-periodic grid (n - 1 cells per axis, clamped outside) and FP64
    # arithmetic over float-stored nodes
    n, N, nsteps = 10, 1024, 40
    A = W.turbulent_field(n, L_BOX, 5e-9, 0.3, 9, dtype=np.float64)
    if storage == "f32":
        A = A.astype(np.float32).astype(np.float64)     # the values float storage holds exactly
    ax = W.grid_axes(n, L_BOX, periodic=False)
    rng = np.random.default_rng(10)
    x = rng.uniform(0.2 * L_BOX, 0.8 * L_BOX, size=(N, 3))
    v = W.isotropic_directions(rng, N) * np.sqrt(2 * 1.0e3 * TP.qi / TP.mi)
    xv0 = np.hstack([x, v])
    dt = 2 * np.pi * TP.mi / (TP.qi * 5e-9) / 40
    tspan = (0.0, nsteps * dt)
    tg = rng.uniform(0, L_BOX, size=(N, 3))

    def prob_of(Av, st):
        B = TP.GridField(Av, *ax, bc=TP.ClampExtrap(), storage=st)
        return TP.TraceProblem(xv0, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField()))

    def loss64(Av):
        xT = TP.solve(prob_of(Av, "f64"), TP.Boris(), dt=dt, save_everystep=False).final_states()[:, :3]
        return float(np.mean(np.sum((xT - tg) ** 2, axis=1)))

    r = TP.loss_gradient(prob_of(A, storage), dt=dt, targets=tg, precision=precision, ckpt_interval=16)
    assert r.loss == pytest.approx(loss64(A), rel=1e-5 if precision == "fp32c" else 1e-13)
    g = np.abs(r.dL_dB)
    cand = np.argwhere(g > 1e-3 * g.max())
    picks = cand[np.random.default_rng(12).choice(len(cand), 16, replace=False)]
    h = 1e-6 * 5e-9
    worst = 0.0
    for c, i, j, k in picks:
        Ap, Am = A.copy(), A.copy()
        Ap[c, i, j, k] += h
        Am[c, i, j, k] -= h
        fd = (loss64(Ap) - loss64(Am)) / (2 * h)
        worst = max(worst, abs(r.dL_dB[c, i, j, k] - fd) / abs(fd))
    print(f"{precision}/{storage} clamped: worst relative error of dL/dB vs central FD {worst:.2e}")
    assert worst < (1e-4 if precision == "fp32c" else 1e-5)


@pytest.mark.parametrize("precision,storage", [("fp64", "f64"), ("fp32c", "f32")])
def test_cell_layout_adjoint_matches_nodes(precision, storage):
    A, ax, xv0, dt, tspan, tg = _setup(N=1000, nsteps=70)
    rs = []
    for layout in ("nodes", "cells"):
        B = TP.GridField(A, *ax, bc=TP.PeriodicBC(), storage=storage, layout=layout)
        prob = TP.TraceProblem(xv0, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
        rs.append(TP.loss_gradient(prob, dt=dt, targets=tg, precision=precision, ckpt_interval=16))
    # the loss is a sum of per-wave partial sums added atomically: last-bit order effects
    assert rs[1].loss == pytest.approx(rs[0].loss, rel=1e-13)
    np.testing.assert_array_equal(rs[0].dL_dxv0, rs[1].dL_dxv0)
    scale = np.abs(rs[0].dL_dB).max()
    np.testing.assert_allclose(rs[1].dL_dB, rs[0].dL_dB, rtol=1e-5, atol=1e-6 * scale)

"""The benchmark's workloads at their full sizes on the device.  FP64 must be
bit-identical to the oracle: the oracle runs a sample of particles spread over
the whole index range on its own (trajectories are independent), and the
device's rows for those particles must match it.  FP32C is held to its
measured error against FP64 (short horizon, before chaos separates them) or
to its energy drift (the dipole, where Boris conserves |v| to rounding)."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_rows_bitwise, rel_err, run_oracle

pytestmark = pytest.mark.gpu


def test_turbulent_box_256_full_size():
    # 256^3 periodic grid, 4,194,304 protons, MultistepBoris4(n=2), 5,000 steps,
    # E = -u x B derived in-kernel, start and end records
    N = 1 << 22
    w, A = W.turbulent_box(N=N, n=256, nsteps=5000, dtype=np.float64)
    ax = W.grid_axes(256, w.extra["L"])
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC(), layout="cells")
    E = TP.UxBField(w.extra["u"], B)
    alg = TP.MultistepBoris(4, n=2)
    kw = dict(dt=w.dt, save_everystep=False)
    prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, E, B, TP.ZeroField()))
    g = TP.solve(prob, alg, **kw)
    assert g.u_all.shape[0] == N and np.all(g.lengths == 2)
    idx = np.linspace(0, N - 1, 48).astype(int)
    sub = TP.TraceProblem(w.xv0[idx], w.tspan, prob.p)
    assert_rows_bitwise(g, idx, run_oracle(sub, alg, **kw))
    # FP32C (float cells) against FP64 after 10 steps: 1.7e-7 median measured
    B32 = TP.GridField(A.astype(np.float32), *ax, bc=TP.PeriodicBC(), layout="cells")
    p32 = TP.TraceProblem(w.xv0, (0.0, 10 * w.dt), (w.q2m, TP.mi, TP.UxBField(w.extra["u"], B32), B32,
                                                     TP.ZeroField()))
    p64 = TP.TraceProblem(w.xv0, (0.0, 10 * w.dt), prob.p)
    a = TP.solve(p32, alg, precision="fp32c", **kw).u_all[:, -1, :3]
    b = TP.solve(p64, alg, **kw).u_all[:, -1, :3]
    e = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    assert np.median(e) < 1e-6 and np.percentile(e, 95) < 1e-5


def test_dipole_belt_full_size():
    # 1,048,576 protons at L = 4-7, standard Boris, dt = T_g/32 per particle,
    # 10,000 steps, removal inside 1 R_E
    w = W.dipole_belt(N=1 << 20, nsteps=10000)
    N = len(w.xv0)
    param = (w.q2m, TP.mi, TP.ZeroField(), TP.DipoleField(w.extra["b0"], w.extra["R"]), TP.ZeroField())
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), param)
    kw = dict(nsteps=w.nsteps, save_everystep=False, isoutside=TP.InsideSphere(TP.Re))
    g = TP.solve(prob, TP.Boris(), dts=w.dts, **kw)
    idx = np.linspace(0, N - 1, 48).astype(int)
    sub = TP.TraceProblem(w.xv0[idx], (0.0, 0.0), param)
    assert_rows_bitwise(g, idx, run_oracle(sub, TP.Boris(), dts=w.dts[idx], **kw))
    # FP32C: |v| drift over 10,000 steps (2.9e-6 median measured; FP64 ~5e-15)
    f = TP.solve(prob, TP.Boris(), dts=w.dts, precision="fp32c", **kw)
    ok = (g.retcodes == 0) & (f.retcodes == 0)
    v0 = np.sum(w.xv0[ok, 3:] ** 2, axis=1)
    drift32 = np.abs(np.sum(f.u_all[ok, -1, 3:] ** 2, axis=1) - v0) / v0
    drift64 = np.abs(np.sum(g.u_all[ok, -1, 3:] ** 2, axis=1) - v0) / v0
    assert np.median(drift32) < 1e-5 and np.max(drift32) < 1e-3
    assert np.median(drift64) < 1e-13


def test_batched_turbulence_full_size():
    # 64 open turbulent boxes (32^3-128^3, float storage, cell layout), ~2e7
    # protons, AdaptiveBoris(0.1), exits at the grid faces, <= 2,000 steps,
    # in one persistent launch
    inst, tspan = W.batched_turbulence(n_inst=64, total=20_000_000)
    probs = [TP.TraceProblem(xv, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(),
                                         TP.GridField(A, *axes, layout="cells"), TP.ZeroField()))
             for A, axes, xv in inst]
    assert sum(len(p.u0) for p in probs) > 19_000_000
    alg = TP.AdaptiveBoris(safety=0.1)
    kw = dict(isoutside=TP.OutsideGrid(), max_steps=2000, save_everystep=False)
    sols = TP.solve_batch(probs, alg, **kw)
    for k in (0, 21, 42, 63):
        pr, s = probs[k], sols[k]
        idx = np.linspace(0, len(pr.u0) - 1, 12).astype(int)
        sub = TP.TraceProblem(pr.u0[idx], tspan, pr.p)
        assert_rows_bitwise(s, idx, run_oracle(sub, alg, **kw))
    # FP32C: final positions of particles that end alike (1.9e-8 median measured)
    f = TP.solve_batch(probs, alg, precision="fp32c", **kw)
    errs = []
    for a, b in zip(f, sols):
        same = (a.lengths == b.lengths) & (a.retcodes == b.retcodes)
        last = b.lengths[same] - 1
        xa = a.u_all[same, last, :3]
        xb = b.u_all[same, last, :3]
        errs.append(np.linalg.norm(xa - xb, axis=1) / np.linalg.norm(xb, axis=1))
    e = np.concatenate(errs)
    assert e.size > 0.9 * sum(len(p.u0) for p in probs)
    assert np.median(e) < 1e-6


def test_adjoint_64_full_size_vs_oracle_duals():
    # 64^3 periodic grid (dB = 0.3 B0), 262,144 protons at 10 keV, standard
    # Boris, 2,048 steps, checkpoints every 64: dL/d(x0, v0) of sampled
    # particles against forward-mode dual numbers through the oracle.  The
    # gradients reach ~1e35 on this chaotic flow; FP64 still agrees to ~1e-13
    # (FP32C overflows at this horizon and is checked at short ones elsewhere).
    import oracle as O

    n, L, B0, N = 64, 1.0e7, 5.0e-9, 1 << 18
    A = W.turbulent_field(n, L, B0, 0.3, 55, dtype=np.float64)
    ax = W.grid_axes(n, L)
    rng = np.random.default_rng(56)
    x = rng.uniform(0, L, size=(N, 3))
    v = W.isotropic_directions(rng, N) * np.sqrt(2 * 1.0e4 * TP.qi / TP.mi)
    xv0 = np.hstack([x, v])
    tg = rng.uniform(0, L, size=(N, 3))
    dt = 2 * np.pi * TP.mi / (TP.qi * B0) / 40
    tspan = (0.0, 2048 * dt)
    B = TP.GridField(A, *ax, bc=TP.PeriodicBC())
    prob = TP.TraceProblem(xv0, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField()))
    r = TP.loss_gradient(prob, dt=dt, targets=tg, precision="fp64", ckpt_interval=64)
    Bo = O.grid(B.data.astype(np.float64), B.origin, B.spacing, O.EXTRAP_PERIODIC)
    idx = np.linspace(0, N - 1, 16).astype(int)
    for i in idx:
        xT, J = O.boris_tangent(xv0[i], Bo, q2m=TP.qi / TP.mi, tspan=tspan, dt=dt)
        ref = 2.0 / N * (xT - tg[i]) @ J
        assert np.linalg.norm(r.dL_dxv0[i] - ref) <= 1e-11 * np.linalg.norm(ref), i

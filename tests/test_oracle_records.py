"""The oracle's saved-record augmentation (save_fields / save_work,
src/boris/boris.jl:405-440; work rates src/equations.jl:394-441; magnetic
properties src/utility/utility.jl:379-422), pinned by the reference's own
assertions (test/test_boris.jl:359-459), by closed forms worked out by hand
for fields whose gradient and curvature are known, and by central differences
of the oracle's field evaluation for the Jacobians."""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(41)


def _fd_jac(orc, f, x, t=0.0, h=1e-6):
    J = np.zeros((3, 3))
    for j in range(3):
        xp, xm = np.array(x, float), np.array(x, float)
        step = h * max(1.0, abs(x[j]))
        xp[j] += step
        xm[j] -= step
        J[:, j] = (orc.field_eval(f, xp, t) - orc.field_eval(f, xm, t)) / (2 * step)
    return J


def test_reference_save_fields_and_work_kats(orc):
    me, qe = 9.1093837015e-31, -1.602176634e-19
    q2m = qe / me
    # test_boris.jl:359-372: electron, E = 0, B = (0, 0, 0.01), K = 10
    u0 = [0.0, 0.0, 0.0, 0.0, 1e5, 0.0]
    r = orc.solve(u0, orc.zero(), orc.uniform([0, 0, 0.01]), q2m=q2m, tspan=(0.0, 3e-8), dt=3e-11,
                  savestepinterval=10, save_fields=True, mass=me)
    t, u = r.traj(0)
    assert u.shape[1] == 12
    for row in (u[0], u[-1]):
        assert list(row[6:9]) == [0.0, 0.0, 0.0]
        assert list(row[9:12]) == [0.0, 0.0, 0.01]
    # test_boris.jl:374-459: E = (0, 1e5, 0), B = (0, 0, 0.01 (1 + x)), v0 = (1e5, 0, 0)
    E = orc.uniform([0.0, 1e5, 0.0])
    gradB = orc.affine([0, 0, 0.01], np.array([[0, 0, 0], [0, 0, 0], [0.01, 0, 0]]), [0, 0, 0])
    kw = dict(q2m=q2m, tspan=(0.0, 1e-9), dt=1e-11, savestepinterval=1, mass=me)
    s_w = orc.solve([0, 0, 0, 1e5, 0, 0], E, gradB, save_work=True, **kw)
    w = s_w.traj(0)[1]
    assert w.shape[1] == 10
    assert abs(w[0, 6]) < 1e-10 and abs(w[0, 7]) < 1e-10 and abs(w[0, 9]) < 1e-10
    assert abs(w[0, 8]) > 0.0
    # both: the work columns do not depend on save_fields (:441-447)
    s_b = orc.solve([0, 0, 0, 1e5, 0, 0], E, gradB, save_work=True, save_fields=True, **kw)
    b = s_b.traj(0)[1]
    assert b.shape[1] == 16
    np.testing.assert_array_equal(b[:, 12:16], w[:, 6:10])
    # betatron: B = (0, 0, 0.01 (1 + 1e7 t)) gives P_betatron > 0
    beta = orc.affine([0, 0, 0.01], np.zeros((3, 3)), [0, 0, 0.01 * 1e7])
    s_t = orc.solve([0, 0, 0, 1e5, 0, 0], orc.zero(), beta, save_work=True, **kw)
    assert s_t.traj(0)[1][0, 9] > 0.0
    # multistep and adaptive solvers carry the same columns (:433-459)
    s_m = orc.solve([0, 0, 0, 1e5, 0, 0], E, gradB, save_work=True, alg=orc.ALG_MULTISTEP,
                    order_N=2, nsub=2, **kw)
    assert s_m.traj(0)[1].shape[1] == 10 and abs(s_m.traj(0)[1][0, 8]) > 0
    s_a = orc.solve([0, 0, 0, 1e5, 0, 0], E, gradB, save_work=True, adaptive=True, safety=0.1,
                    nout_cap=2000, q2m=q2m, tspan=(0.0, 1e-9), mass=me)
    assert s_a.traj(0)[1].shape[1] == 10 and abs(s_a.traj(0)[1][0, 8]) > 0


def test_gradient_drift_work_closed_form(orc):
    # B = B0 (1 + g x) z, E = Ey y, x = 0, v = (v, 0, 0):
    # grad|B| = B0 g x, kappa = 0, mu = m v^2 / (2 B0), P_grad = mu g Ey
    B0, g, Ey, v, m, q2m = 0.01, 1.0, 1e5, 1e5, 9.1093837015e-31, -1.75882001076e11
    B = orc.affine([0, 0, B0], np.array([[0, 0, 0], [0, 0, 0], [B0 * g, 0, 0]]), [0, 0, 0])
    E = orc.uniform([0, Ey, 0])
    out = orc.record_eval(E, B, q2m=q2m, m=m, xv=[0, 0, 0, v, 0, 0], t=0.0)[0]
    np.testing.assert_array_equal(out[:6], [0, Ey, 0, 0, 0, B0])
    mu = m * v * v / (2 * B0)
    np.testing.assert_allclose(out[6:], [0.0, 0.0, mu * g * Ey, 0.0], rtol=1e-14, atol=0)


def test_parallel_and_betatron_closed_forms(orc):
    # E parallel to B: P_par = q v_par E_par; dB/dt along b: P_betatron = mu dB/dt
    q2m, m = 2.0, 3.0
    B = orc.affine([0, 0, 4.0], np.zeros((3, 3)), [0, 0, 0.5])
    E = orc.uniform([0, 0, 7.0])
    out = orc.record_eval(E, B, q2m=q2m, m=m, xv=[1.0, 2.0, 3.0, 0.3, 0.4, 5.0], t=2.0)[0]
    Bz = 4.0 + 0.5 * 2.0
    np.testing.assert_allclose(out[3:6], [0, 0, Bz], rtol=0)
    mu = m * (0.3 ** 2 + 0.4 ** 2) / (2 * Bz)
    np.testing.assert_allclose(out[6:], [q2m * m * 5.0 * 7.0, 0.0, 0.0, mu * 0.5], rtol=1e-14)


def test_curved_field_lines_fermi_and_grad(orc):
    # B = c (-y, x, 0): circular field lines, |B| = c r, at (r, 0, 0):
    # b = y, kappa = -x / r, grad|B| = c x.  With E = (0, 0, Ez):
    # P_fermi = m v_par^2 Ez / (c r^2), P_grad = -mu Ez / r
    c, r, Ez, m = 0.5, 2.0, 3.0, 1.5
    B = orc.affine([0, 0, 0], np.array([[0, -c, 0], [c, 0, 0], [0, 0, 0]]), [0, 0, 0])
    E = orc.uniform([0, 0, Ez])
    vpar, vx, vz = 0.7, 0.2, -0.1
    out = orc.record_eval(E, B, q2m=1.0, m=m, xv=[r, 0, 0, vx, vpar, vz], t=0.0)[0]
    mu = m * (vx * vx + vz * vz) / (2 * c * r)
    np.testing.assert_allclose(out[6:], [1.0 * m * vpar * 0.0, m * vpar ** 2 * Ez / (c * r * r),
                                         -mu * Ez / r, 0.0], rtol=1e-13, atol=1e-15)


def test_dipole_gradient_and_curvature_at_equator(orc):
    # equator of a dipole: |B| = b0 (R/r)^3 so grad|B| = -3 |B| / r x, and the
    # field-line curvature radius is r / 3, toward the dipole
    b0, R, r = 3.12e-5, 6.371e6, 4 * 6.371e6
    B = orc.dipole(b0, R)
    Bv, J, dBdt = orc.field_jac(B, [r, 0, 0])
    Bmag = np.linalg.norm(Bv)
    bh = Bv / Bmag
    grad = J.T @ bh
    kappa = (J @ bh - bh * (grad @ bh)) / Bmag
    np.testing.assert_allclose(Bmag, b0 * (R / r) ** 3, rtol=1e-14)
    np.testing.assert_allclose(grad, [-3 * Bmag / r, 0, 0], rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(kappa, [-3 / r, 0, 0], rtol=1e-12, atol=1e-25)
    assert list(dBdt) == [0.0, 0.0, 0.0]


def test_jacobians_match_central_differences(orc):
    # analytic Jacobians against central differences of the oracle's field evaluation
    dip = orc.dipole(3.12e-5, 6.371e6)
    for _ in range(20):
        x = RNG.normal(size=3) * 2e7 + np.array([3e7, 0, 0])
        _, J, _ = orc.field_jac(dip, x)
        np.testing.assert_allclose(J, _fd_jac(orc, dip, x, h=1e-7), rtol=1e-6,
                                   atol=1e-6 * np.abs(J).max())
    aff = orc.affine([1.0, 2.0, 3.0], RNG.normal(size=(3, 3)), [0.1, 0.2, 0.3])
    _, J, dBdt = orc.field_jac(aff, [0.3, -0.2, 0.5], 1.5)
    np.testing.assert_allclose(J, _fd_jac(orc, aff, [0.3, -0.2, 0.5], 1.5), rtol=1e-8)
    np.testing.assert_array_equal(dBdt, [0.1, 0.2, 0.3])
    n, h = (6, 5, 7), (0.5, 0.4, 0.3)
    data = RNG.normal(size=(n[2], n[1], n[0], 3))
    for ex in (orc.EXTRAP_PERIODIC, orc.EXTRAP_CLAMP, orc.EXTRAP_FILL_NAN):
        g = orc.grid(data, (0.0, 0.0, 0.0), h, ex)
        for _ in range(20):
            # interior points away from cell faces, where the interpolant is smooth
            cell = [RNG.integers(0, n[d] - 1) for d in range(3)]
            x = np.array([(cell[d] + RNG.uniform(0.1, 0.9)) * h[d] for d in range(3)])
            _, J, _ = orc.field_jac(g, x)
            np.testing.assert_allclose(J, _fd_jac(orc, g, x, h=1e-6), rtol=1e-6, atol=1e-8)
    # clamped outside the box: flat along the clamped axis
    g = orc.grid(data, (0.0, 0.0, 0.0), h, orc.EXTRAP_CLAMP)
    x = np.array([-1.0, 0.85, 0.95])
    _, J, _ = orc.field_jac(g, x)
    np.testing.assert_array_equal(J[:, 0], [0.0, 0.0, 0.0])
    np.testing.assert_allclose(J, _fd_jac(orc, g, x, h=1e-6), rtol=1e-6, atol=1e-8)


def test_zero_field_gives_zero_work(orc):
    # Bmag == 0 returns zeros (equations.jl:414-416)
    out = orc.record_eval(orc.uniform([1.0, 2.0, 3.0]), orc.zero(), q2m=1.0, m=1.0,
                          xv=[0, 0, 0, 1, 2, 3], t=0.0)[0]
    np.testing.assert_array_equal(out[6:], [0.0, 0.0, 0.0, 0.0])
    assert math.isfinite(out[0])


def test_work_off_equator_matches_differenced_geometry(orc):
    # away from the equator grad|B| . b != 0, so every term of the curvature
    # (utility.jl:383) matters; kappa = (b . grad) b and grad|B| by differences
    b0, R = 3.12e-5, 6.371e6
    B = orc.dipole(b0, R)
    x = np.array([2.0e7, 1.0e7, 1.5e7])
    bhat = lambda p: orc.field_eval(B, p, 0.0) / np.linalg.norm(orc.field_eval(B, p, 0.0))
    bmag = lambda p: np.linalg.norm(orc.field_eval(B, p, 0.0))
    b = bhat(x)
    eps = 1e2
    kappa = (bhat(x + eps * b) - bhat(x - eps * b)) / (2 * eps)
    grad = np.array([(bmag(x + eps * e) - bmag(x - eps * e)) / (2 * eps) for e in np.eye(3)])
    assert abs(grad @ b) > 0.1 * np.linalg.norm(grad)
    E = orc.uniform([1e-3, -2e-3, 5e-4])
    v = np.array([3e5, -1e5, 2e5])
    q2m, m = 9.58e7, 1.67e-27
    out = orc.record_eval(E, B, q2m=q2m, m=m, xv=np.r_[x, v], t=0.0)[0]
    Bm = bmag(x)
    vpar = v @ b
    mu = m * np.sum((v - vpar * b) ** 2) / (2 * Bm)
    Ev = orc.field_eval(E, x, 0.0)
    expect = [q2m * m * vpar * (Ev @ b), m * vpar ** 2 / Bm * (np.cross(b, kappa) @ Ev),
              mu / Bm * (np.cross(b, grad) @ Ev), 0.0]
    np.testing.assert_allclose(out[6:], expect, rtol=1e-6, atol=1e-40)

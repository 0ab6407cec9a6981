"""Pin the oracle where the reference's known answers leave it open.

Each check is an independent statement of the mathematics (a closed form, an
invariant, a brute-force count or a convergence order), chosen so that a
plausible slip in tp_oracle.c -- a dropped term, a flipped sign, a swapped
index -- fails at least one of them.
"""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(12345)


def rodrigues(v, axis, angle):
    k = axis / np.linalg.norm(axis)
    return (v * math.cos(angle) + np.cross(k, v) * math.sin(angle)
            + k * np.dot(k, v) * (1 - math.cos(angle)))


def test_boris_is_exact_rotation(orc):
    # E = 0: Boris rotates v about B by -2 atan(|h B|) (positive q2m gyrates
    # clockwise about B); |v| is conserved to rounding.
    for _ in range(200):
        v = RNG.normal(size=3)
        B = RNG.normal(size=3)
        h = RNG.uniform(0.01, 1.5) * (1 if RNG.random() < 0.5 else -1)
        out = orc.boris_velocity_update(v, [0, 0, 0], B, h)
        tmag = abs(h) * np.linalg.norm(B)
        ang = -2 * math.atan(tmag) * np.sign(h)
        np.testing.assert_allclose(out, rodrigues(v, B, ang), rtol=0, atol=1e-13 * np.linalg.norm(v))
        assert abs(np.linalg.norm(out) - np.linalg.norm(v)) <= 1e-14 * np.linalg.norm(v)


def test_boris_exb_fixed_point(orc):
    # v = E x B / |B|^2 is a fixed point of the full update when E is perpendicular to B
    for _ in range(100):
        B = RNG.normal(size=3)
        E = np.cross(RNG.normal(size=3), B)
        vd = np.cross(E, B) / np.dot(B, B)
        out = orc.boris_velocity_update(vd, E, B, RNG.uniform(0.05, 1.0))
        np.testing.assert_allclose(out, vd, rtol=0, atol=1e-13 * (np.linalg.norm(vd) + 1))


def test_boris_electric_acceleration(orc):
    # B = 0: v -> v + 2 h E (two half kicks)
    v, E, h = RNG.normal(size=3), RNG.normal(size=3), 0.37
    np.testing.assert_allclose(orc.boris_velocity_update(v, E, [0, 0, 0], h), v + 2 * h * E,
                               rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 12])
def test_rational_polynomials_are_the_trig_coefficients(orc, n):
    D, P1, P2, P3, P6 = orc.rational_polys(n)
    # integer coefficients; (1+u)^n
    assert np.all(D == [math.comb(n, j) for j in range(n + 1)])
    for tau in (1e-3, 0.05, 0.3, 0.9):
        u = tau * tau
        a = math.atan(tau)
        ev = lambda p: np.polyval(p[::-1], u)
        Dv = ev(D)
        np.testing.assert_allclose(ev(P1) / Dv, math.cos(2 * n * a), rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(ev(P2) / Dv, math.sin(2 * n * a) / tau, rtol=1e-12)
        np.testing.assert_allclose(ev(P3) / Dv, 2 * math.sin(n * a) ** 2 / u, rtol=1e-10)
        np.testing.assert_allclose(ev(P6) / Dv, (2 * n - math.sin(2 * n * a) / tau) / u, rtol=1e-6)
    # tau -> 0 limits are the Taylor-branch values of multistep_boris.jl:101-108
    assert P1[0] / D[0] == 1 and P2[0] == 2 * n and P3[0] == 2 * n * n
    assert P6[0] == pytest.approx((2 * n + 4 * n ** 3) / 3)


@pytest.mark.parametrize("n", [1, 2, 4, 6])
def test_rational_and_reference_forms_agree(orc, n):
    for tau in np.concatenate([np.geomspace(1e-7, 1e-4, 7), np.geomspace(1e-4, 1.5, 25)]):
        cr = orc.multistep_coeffs(tau * tau, n, orc.FORM_REFERENCE)
        cq = orc.multistep_coeffs(tau * tau, n, orc.FORM_RATIONAL)
        # c6 carries (2n - c2)/tau^2 cancellation in the reference form
        np.testing.assert_allclose(cq[:3], cr[:3], rtol=2e-13, atol=2e-14)
        tol = 1e-12 if tau > 1e-2 else (1e-5 if tau > 1e-4 else 1e-7)
        np.testing.assert_allclose(cq[3], cr[3], rtol=tol)


@pytest.mark.parametrize("n", [1, 2, 3, 5])
@pytest.mark.parametrize("form", [0, 1])
def test_multistep_n2_is_n_boris_substeps(orc, n, form):
    # N = 2, n subcycles in uniform fields equals n Boris updates with dt/n
    q2m, dt = 1.7, 0.23
    for _ in range(30):
        v, E, B = RNG.normal(size=3), RNG.normal(size=3), RNG.normal(size=3)
        w = v.copy()
        for _ in range(n):
            w = orc.boris_velocity_update(w, E, B, q2m * 0.5 * dt / n)
        got = orc.multistep_velocity(v, E, B, q2m, dt, n, 2, form)
        np.testing.assert_allclose(got, w, rtol=1e-12, atol=1e-12)


def _gyro_phase_error(orc, N, dt, form):
    # one update with E = 0 about B = z: exact gyration angle is -q2m |B| dt
    v = np.array([1.0, 0.0, 0.0])
    out = orc.multistep_velocity(v, [0, 0, 0], [0, 0, 1.0], 1.0, dt, 1, N, form)
    return abs(math.atan2(out[1], out[0]) - (-dt))


@pytest.mark.parametrize("form", [0, 1])
def test_hyper_boris_gyrophase_order(orc, form):
    # Zenitani & Kato 2025: phase error per step O(dt^(N+1)) for N = 2, 4, 6
    for N, order in ((2, 3), (4, 5), (6, 7)):
        e1 = _gyro_phase_error(orc, N, 0.2, form)
        e2 = _gyro_phase_error(orc, N, 0.1, form)
        assert math.log2(e1 / e2) == pytest.approx(order, abs=0.15), (N, e1, e2)


def test_hyper_boris_exb_exact(orc):
    # hyper correction keeps the E x B drift a fixed point
    for N in (4, 6):
        for form in (0, 1):
            E, B = np.array([0, 1.0, 0]), np.array([0, 0, 1.0])
            vd = np.cross(E, B)
            out = orc.multistep_velocity(vd, E, B, 1.0, 0.3, 2, N, form)
            np.testing.assert_allclose(out, vd, atol=1e-14)


def _grid_from_func(f, n, origin, h, dtype=np.float64):
    nx, ny, nz = n
    zz, yy, xx = np.meshgrid(*(origin[d] + h[d] * np.arange(n[d]) for d in (2, 1, 0)),
                             indexing="ij")
    return f(xx, yy, zz).astype(dtype)


def test_trilinear_exact_for_multilinear_data(orc):
    # data multilinear in (x, y, z) is reproduced exactly by trilinear interpolation
    n, origin, h = (5, 7, 6), (-1.0, 2.0, 0.5), (0.5, 0.25, 1.0)

    def f(x, y, z):
        return np.stack([1 + 2 * x - 3 * y + 0.5 * z + x * y * z,
                         -2 + x * y, 0.25 * z - y * z], axis=-1)

    g = orc.grid(_grid_from_func(f, n, origin, h), origin, h, orc.EXTRAP_FILL_NAN)
    for _ in range(300):
        p = [origin[d] + RNG.uniform(0, (n[d] - 1) * h[d]) for d in range(3)]
        np.testing.assert_allclose(orc.field_eval(g, p), f(*p), rtol=1e-13, atol=1e-13)
    # nodes are reproduced exactly, including the upper faces
    top = [origin[d] + (n[d] - 1) * h[d] for d in range(3)]
    np.testing.assert_allclose(orc.field_eval(g, top), f(*top), rtol=1e-15)


def test_periodic_and_clamp_extrapolation(orc):
    n, origin, h = (4, 5, 6), (0.0, 0.0, 0.0), (1.0, 2.0, 0.5)
    data = RNG.normal(size=(n[2], n[1], n[0], 3))
    per = orc.grid(data, origin, h, orc.EXTRAP_PERIODIC)
    cla = orc.grid(data, origin, h, orc.EXTRAP_CLAMP)
    L = [n[d] * h[d] for d in range(3)]
    for _ in range(100):
        p = np.array([RNG.uniform(0, L[d]) for d in range(3)])
        shift = np.array([L[d] * RNG.integers(-3, 4) for d in range(3)])
        np.testing.assert_allclose(orc.field_eval(per, p + shift), orc.field_eval(per, p),
                                   rtol=1e-10, atol=1e-12)
    # wrap cell: halfway between node N-1 and node 0 along x
    p = [(n[0] - 0.5) * h[0], 0.0, 0.0]
    np.testing.assert_allclose(orc.field_eval(per, p), 0.5 * (data[0, 0, n[0] - 1] + data[0, 0, 0]),
                               rtol=1e-14)
    # clamp: far outside equals the corner node
    np.testing.assert_allclose(orc.field_eval(cla, [-50.0, 99.0, -3.0]), data[0, n[1] - 1, 0],
                               rtol=1e-15)


def test_uxb_field(orc):
    B = orc.uniform([0.3, -1.0, 2.0])
    u = np.array([4e5, 1.0, -2.0])
    E = orc.uxb(u, B)
    np.testing.assert_allclose(orc.field_eval(E, [1, 2, 3]), -np.cross(u, [0.3, -1.0, 2.0]),
                               rtol=1e-15)


def test_dipole_matches_vector_formula(orc):
    b0, R = 3.12e-5, 6371.0e3
    f = orc.dipole(b0, R)
    m = np.array([0.0, 0.0, -1.0])
    for _ in range(100):
        r = RNG.normal(size=3) * 3 * R
        rn = np.linalg.norm(r)
        ref = b0 * R ** 3 * (3 * np.dot(m, r) * r - m * rn ** 2) / rn ** 5
        np.testing.assert_allclose(orc.field_eval(f, r), ref, rtol=1e-12, atol=1e-25)
    # equator at 1 R_E: (0, 0, +b0)
    np.testing.assert_allclose(orc.field_eval(f, [R, 0, 0]), [0, 0, b0], rtol=1e-15)


def test_nout_brute_force(orc):
    # count the records a full (never terminated) run writes, by walking the
    # schedule: the start record, the in-loop saves at steps it-1 = K, 2K, ...
    # (boris.jl:463), and the record at nt only when save_end (the reference's
    # scenario 5, test_boris.jl:325-334, stops at step 990 of 1000)
    for nt in range(0, 40):
        for K in range(1, 9):
            for ss in (0, 1):
                for se in (0, 1):
                    for sev in (0, 1):
                        cnt = ss + (sum(1 for it in range(1, nt + 1)
                                        if sev and (it - 1) > 0 and (it - 1) % K == 0))
                        if se:
                            cnt += 1
                        assert orc.calculate_nout(nt, K, ss, se, sev) == cnt, (nt, K, ss, se, sev)


def test_fixed_step_uniform_closed_form(orc):
    # uniform B, E = 0: the leapfrog orbit is a circle whose per-step phase is
    # -2 atan(q2m B dt / 2); positions follow from the exact rotation.
    q2m, Bz, dt, nt = 2.0, 1.5, 0.05, 400
    v0 = np.array([0.7, -0.2, 0.3])
    r = orc.solve(np.r_[np.zeros(3), v0], orc.zero(), orc.uniform([0, 0, Bz]), q2m=q2m,
                  tspan=(0.0, nt * dt), dt=dt, savestepinterval=1)
    t, u = r.traj(0)
    assert len(t) == nt + 1
    th = 2 * math.atan(q2m * Bz * dt / 2)
    # v_{-1/2}: half back-step, then each full step rotates by -th
    vh = orc.boris_velocity_update(v0, [0, 0, 0], [0, 0, Bz], q2m * 0.5 * (-0.5 * dt))
    x = np.zeros(3)
    for k in range(1, nt + 1):
        vh = rodrigues(vh, np.array([0, 0, 1.0]), -th)
        x = x + vh * dt
    np.testing.assert_allclose(u[-1][:3], x, rtol=0, atol=1e-11)
    assert abs(np.linalg.norm(u[-1][3:]) - np.linalg.norm(v0)) < 1e-14


def test_adaptive_uniform_matches_fixed(orc):
    # in uniform B the adaptive step is constant, dt = safety * T_gyro, so the
    # first records agree with the fixed-step solver at that dt
    q2m, Bz, safety = 1.0, 2.0, 0.05
    dt = safety * 2 * math.pi / (q2m * Bz)
    u0 = [0.1, 0.2, 0.0, 1.0, 0.5, 0.2]
    a = orc.solve(u0, orc.zero(), orc.uniform([0, 0, Bz]), q2m=q2m, tspan=(0.0, 123.4 * dt),
                  adaptive=True, safety=safety, nout_cap=1000)
    f = orc.solve(u0, orc.zero(), orc.uniform([0, 0, Bz]), q2m=q2m, tspan=(0.0, 100 * dt), dt=dt)
    ta, ua = a.traj(0)
    tf, uf = f.traj(0)
    np.testing.assert_allclose(ta[:100], tf[:100], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(ua[:100], uf[:100], rtol=1e-9, atol=1e-11)
    assert ta[-1] == pytest.approx(123.4 * dt, rel=1e-14)


def test_boris_tangent_matches_solve_and_finite_differences(orc):
    # forward-mode duals through the fixed-step Boris solve: the primal is the
    # solver's own final position, dx(T)/dx0 is the identity in a uniform B,
    # and in a periodic gridded B the Jacobian matches central differences
    q2m, dt, nt = 1.0, 0.05, 200
    u0 = np.array([0.3, -0.2, 0.1, 0.8, -0.4, 0.3])
    const = orc.grid(np.broadcast_to([0.1, -0.3, 1.2], (3, 3, 3, 3)).copy(), (0.0, 0.0, 0.0),
                     (1.0, 1.0, 1.0), orc.EXTRAP_PERIODIC)
    xT, J = orc.boris_tangent(u0, const, q2m=q2m, tspan=(0.0, nt * dt), dt=dt)
    r = orc.solve(u0, orc.zero(), const, q2m=q2m, tspan=(0.0, nt * dt), dt=dt,
                  save_everystep=False)
    np.testing.assert_array_equal(xT, r.traj(0)[1][-1][:3])
    np.testing.assert_allclose(J[:, :3], np.eye(3), atol=1e-13)

    n, L = (6, 5, 7), 4.0
    h = tuple(L / n[d] for d in range(3))
    data = RNG.normal(scale=0.4, size=(n[2], n[1], n[0], 3)) + np.array([0.0, 0.0, 1.0])
    g = orc.grid(data, (0.0, 0.0, 0.0), h, orc.EXTRAP_PERIODIC)
    xT, J = orc.boris_tangent(u0, g, q2m=q2m, tspan=(0.0, nt * dt), dt=dt)
    r = orc.solve(u0, orc.zero(), g, q2m=q2m, tspan=(0.0, nt * dt), dt=dt, save_everystep=False)
    np.testing.assert_array_equal(xT, r.traj(0)[1][-1][:3])
    eps = 1e-6
    fd = np.zeros((3, 6))
    for j in range(6):
        up, um = u0.copy(), u0.copy()
        up[j] += eps
        um[j] -= eps
        xp = orc.solve(up, orc.zero(), g, q2m=q2m, tspan=(0.0, nt * dt), dt=dt,
                       save_everystep=False).traj(0)[1][-1][:3]
        xm = orc.solve(um, orc.zero(), g, q2m=q2m, tspan=(0.0, nt * dt), dt=dt,
                       save_everystep=False).traj(0)[1][-1][:3]
        fd[:, j] = (xp - xm) / (2 * eps)
    np.testing.assert_allclose(J, fd, rtol=0, atol=1e-6 * np.abs(fd).max())
    # a dropped field-gradient term would make the position block the identity
    assert np.abs(J[:, :3] - np.eye(3)).max() > 1e-3


def test_loss_tangent_closed_form_in_uniform_field(orc):
    # In a uniform B = B0 z every Boris step is an exact rotation about z by
    # -theta, theta = 2 atan(h B0) (h = q2m dt / 2), and the back half-step at
    # t0 (boris.jl:514) rotates by +phi, phi = 2 atan(h B0 / 2).  So
    #   x(T) = x0 + dt sum_{n=1..nt} R(phi - n theta) v0,
    # and along the direction "every node's z component + 1" (d/dB0)
    #   dx(T)/dB0 = dt sum_n (phi' - n theta') z x R(phi - n theta) v0.
    q2m, dt, nt, B0 = 1.0, 0.07, 150, 1.3
    rng = np.random.default_rng(21)
    N = 5
    xv0 = np.hstack([rng.uniform(0, 3, (N, 3)), rng.normal(size=(N, 3))])
    tg = rng.uniform(0, 3, (N, 3))
    n = (4, 3, 5)
    data = np.zeros((n[2], n[1], n[0], 3))
    data[..., 2] = B0
    g = orc.grid(data, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), orc.EXTRAP_PERIODIC)
    dz = np.zeros((1, data.size))
    dz[0, 2::3] = 1.0
    L, dL = orc.loss_tangent(xv0, tg, g, q2m=q2m, tspan=(0.0, nt * dt), dt=dt, dense=dz, nthreads=2)
    h = q2m * dt / 2
    th, ph = 2 * np.arctan(h * B0), 2 * np.arctan(h * B0 / 2)
    dth, dph = 2 * h / (1 + (h * B0) ** 2), h / (1 + (h * B0 / 2) ** 2)
    a = ph - np.arange(1, nt + 1) * th                      # rotation angle of v_{n-1/2}
    da = dph - np.arange(1, nt + 1) * dth
    ca, sa = np.cos(a), np.sin(a)
    xT = np.empty((N, 3))
    dxT = np.zeros((N, 3))
    for i in range(N):
        vx, vy, vz = xv0[i, 3:]
        rx, ry = ca * vx - sa * vy, sa * vx + ca * vy       # R(a) v0 (counter-clockwise)
        xT[i] = xv0[i, :3] + dt * np.array([rx.sum(), ry.sum(), nt * vz])
        dxT[i, :2] = dt * np.array([np.sum(-da * ry), np.sum(da * rx)])   # z x (R v0)
    L_ref = np.mean(np.sum((xT - tg) ** 2, axis=1))
    dL_ref = np.mean(np.sum(2 * (xT - tg) * dxT, axis=1))
    assert L == pytest.approx(L_ref, rel=1e-11)
    assert dL[0] == pytest.approx(dL_ref, rel=1e-9)


def test_loss_tangent_gridded_field(orc):
    # periodic random grid: the directional derivatives are linear in the
    # direction, a single-value direction equals its dense one-hot bit for bit,
    # the primal loss is the solver's, and central differences of the solver's
    # loss (per-particle differences, h = 1e-7) agree
    q2m, dt, nt = 1.0, 0.05, 60
    n, Lbox = (6, 5, 7), 4.0
    sp = tuple(Lbox / n[d] for d in range(3))
    data = RNG.normal(scale=0.4, size=(n[2], n[1], n[0], 3)) + np.array([0.0, 0.0, 1.0])
    rng = np.random.default_rng(22)
    N = 40
    xv0 = np.hstack([rng.uniform(0, Lbox, (N, 3)), rng.normal(scale=0.5, size=(N, 3))])
    tg = rng.uniform(0, Lbox, (N, 3))
    tspan = (0.0, nt * dt)
    g = orc.grid(data, (0.0, 0.0, 0.0), sp, orc.EXTRAP_PERIODIC)
    d1, d2 = rng.normal(size=(2, data.size))
    L, dL = orc.loss_tangent(xv0, tg, g, q2m=q2m, tspan=tspan, dt=dt,
                             dense=np.stack([d1, d2, 0.5 * d1 - 2.0 * d2]), nthreads=3)
    assert dL[2] == pytest.approx(0.5 * dL[0] - 2.0 * dL[1], rel=1e-10)
    r = orc.solve(xv0, orc.zero(), g, q2m=q2m, tspan=tspan, dt=dt, save_everystep=False,
                  save_start=False)
    xT = np.array([r.traj(i)[1][-1][:3] for i in range(N)])
    assert L == pytest.approx(np.mean(np.sum((xT - tg) ** 2, axis=1)), rel=1e-14)

    def final(dat):
        gg = orc.grid(dat, (0.0, 0.0, 0.0), sp, orc.EXTRAP_PERIODIC)
        rr = orc.solve(xv0, orc.zero(), gg, q2m=q2m, tspan=tspan, dt=dt, save_everystep=False,
                       save_start=False)
        return np.array([rr.traj(i)[1][-1][:3] for i in range(N)])

    eps = 1e-7
    for k, dvec in enumerate((d1, d2)):
        xp = final(data + eps * dvec.reshape(data.shape))
        xm = final(data - eps * dvec.reshape(data.shape))
        fd = np.mean(np.sum((xp - xm) * (xp + xm - 2 * tg), axis=1)) / (2 * eps)
        assert dL[k] == pytest.approx(fd, rel=1e-6), k
    flat = np.array([5, 3 * 17 + 2, 3 * 100 + 1])
    Ls, dLs = orc.loss_tangent(xv0, tg, g, q2m=q2m, tspan=tspan, dt=dt, flat=flat, nthreads=3)
    onehot = np.zeros((3, data.size))
    onehot[np.arange(3), flat] = 1.0
    Ld, dLd = orc.loss_tangent(xv0, tg, g, q2m=q2m, tspan=tspan, dt=dt, dense=onehot, nthreads=3)
    assert Ls == Ld
    np.testing.assert_array_equal(dLs, dLd)
    assert np.all(dLs != 0.0)


def _ulp_err(got, exact):
    import math

    import mpmath
    e = float(abs(mpmath.mpf(got) - exact))
    return e / math.ulp(float(exact))


def test_fdlibm_atan_within_one_ulp(orc):
    # the reference form's atan (src/boris/multistep_boris.jl:110), as the
    # device computes it: within 1 ulp of atan in 60-digit arithmetic on every
    # reduction interval (|x| < 7/16, 7/16-11/16, 11/16-19/16, 19/16-39/16,
    # above), and exact at the special values
    import mpmath
    mpmath.mp.dps = 60
    rng = np.random.default_rng(31)
    xs = np.concatenate([rng.uniform(1e-9, 0.4375, 3000), rng.uniform(0.4375, 2.4375, 3000),
                         10 ** rng.uniform(0.39, 19.8, 2000), -rng.uniform(0, 3, 1000),
                         [0.4375, 0.6875, 1.1875, 2.4375, 1e-4, 0.039, 7.450580596923828e-9]])
    worst = max(_ulp_err(orc.fdlibm_atan(float(x)), mpmath.atan(mpmath.mpf(float(x)))) for x in xs)
    assert worst < 1.0, worst
    assert orc.fdlibm_atan(0.0) == 0.0 and np.signbit(orc.fdlibm_atan(-0.0))
    assert orc.fdlibm_atan(np.inf) == orc.fdlibm_atan(1e30) == np.pi / 2
    assert np.isnan(orc.fdlibm_atan(np.nan))


def test_fdlibm_sincos_within_one_ulp(orc):
    # sin and cos of the reference form (multistep_boris.jl:112) with the
    # medium-size argument reduction: within 1 ulp of the 60-digit values up
    # to |x| ~ 2^19 pi/2, including arguments next to multiples of pi/2
    import mpmath
    mpmath.mp.dps = 60
    rng = np.random.default_rng(32)
    near = np.array([k * np.pi / 2 for k in range(1, 400)])
    xs = np.concatenate([rng.uniform(-0.8, 0.8, 2000), rng.uniform(-40, 40, 2000),
                         rng.uniform(-8.2e5, 8.2e5, 1000), near, np.nextafter(near, 0), 1e-9 * near])
    ws = wc = 0.0
    for x in xs:
        s, c = orc.fdlibm_sincos(float(x))
        xm = mpmath.mpf(float(x))
        ws = max(ws, _ulp_err(s, mpmath.sin(xm)))
        wc = max(wc, _ulp_err(c, mpmath.cos(xm)))
    assert ws < 1.0 and wc < 1.0, (ws, wc)
    s, c = orc.fdlibm_sincos(2.0e6)          # past the medium reduction: NaN, never asked for
    assert np.isnan(s) and np.isnan(c)

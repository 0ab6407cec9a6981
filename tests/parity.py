"""Shared test helpers: one field/algorithm description, run through both the
device library (testparticle_amd) and the CPU oracle (oracle/), on the same
inputs."""
from __future__ import annotations

import numpy as np

import oracle as O
import testparticle_amd as TP
from testparticle_amd import _lib


def oracle_field(f, B_oracle=None):
    """Translate a device field description into the oracle's."""
    if isinstance(f, TP.ZeroField):
        return O.zero()
    if isinstance(f, TP.UniformField):
        return O.uniform(f.value)
    if isinstance(f, TP.AffineField):
        return O.affine(f.f0, f.jac, f.ft)
    if isinstance(f, TP.DipoleField):
        return O.dipole(f.b0, f.radius)
    if isinstance(f, TP.GridField):
        data = f.data
        if f.storage == "f32" and data.dtype != np.float32:
            data = data.astype(np.float32)
        if f.storage == "f64" and data.dtype != np.float64:
            data = data.astype(np.float64)
        return O.grid(data, f.origin, f.spacing, f.extrap)
    if isinstance(f, TP.UxBField):
        return O.uxb(f.u, B_oracle)
    if isinstance(f, TP.MeshField):
        data = f.data.astype(np.float32 if f.storage == "f32" else np.float64, copy=False)
        coord = O.MESH_SPHERICAL if f.coord == "spherical" else O.MESH_CARTESIAN
        return O.mesh(data, f.axes, coord, f.order, f.extrap)
    if isinstance(f, TP.LazyTimeInterpolator):
        dt = np.float32 if f.storage == "f32" else np.float64
        snaps = np.stack([g.data.astype(dt, copy=False) for _, g in f.snapshots()])
        return O.time_grid(snaps, f.times, f.origin, f.spacing, f.extrap)
    raise TypeError(type(f))


def run_oracle(prob, alg, *, dt=None, savestepinterval=1, isoutside=None, save_start=True,
               save_end=True, save_everystep=True, maxiters=1_000_000, form="reference",
               max_steps=0, nout_cap=None, dts=None, nsteps=0, nthreads=8, trajectories=1,
               save_fields=False, save_work=False):
    xv0 = TP.initial_states(prob, trajectories)     # host-side prob_func, no device
    B = oracle_field(prob.p[3])
    E = oracle_field(prob.p[2], B)
    kw = {}
    if isinstance(alg, TP.MultistepBoris):
        kw = dict(alg=O.ALG_MULTISTEP, order_N=alg.N, nsub=alg.n,
                  form=O.FORM_RATIONAL if form == "rational" else O.FORM_REFERENCE)
    b = isoutside.condition if hasattr(isoutside, "condition") else isoutside
    if b is not None:
        flags, lo, hi = b.flags, list(b.box_lo), list(b.box_hi)
        if flags & _lib.BND_GRID:
            # the grid's node span, computed as the library does: origin + (n-1) / (1/h)
            g = prob.p[3]
            for d in range(3):
                if isinstance(g, TP.MeshField):      # a mesh's node box; one node: unbounded
                    a = g.axes[d]
                    glo, ghi = (-np.inf, np.inf) if len(a) == 1 else (float(a[0]), float(a[-1]))
                else:
                    glo = float(g.origin[d])
                    ghi = glo + float(g.shape[d] - 1) / (1.0 / float(g.spacing[d]))
                if flags & _lib.BND_BOX:
                    lo[d], hi[d] = max(lo[d], glo), min(hi[d], ghi)
                else:
                    lo[d], hi[d] = glo, ghi
            flags = (flags & ~_lib.BND_GRID) | _lib.BND_BOX
        kw.update(bnd_flags=flags, r_in=b.r_in, r_out=b.r_out, box_lo=lo, box_hi=hi)
    if alg.adaptive:
        kw.update(adaptive=True, safety=alg.safety, max_steps=max_steps,
                  nout_cap=nout_cap if nout_cap is not None else 4096)
    return O.solve(xv0, E, B, q2m=prob.p[0], tspan=prob.tspan, dt=dt or 0.0,
                   savestepinterval=savestepinterval, save_start=save_start, save_end=save_end,
                   save_everystep=save_everystep, maxiters=maxiters, dt_pp=dts,
                   nsteps_pp=nsteps, nthreads=nthreads, save_fields=save_fields,
                   save_work=save_work, mass=float(prob.p[1]), **kw)


def assert_bitwise(gpu: TP.EnsembleSolution, orc: O.OracleResult, n_check=None):
    N = len(gpu)
    assert orc.len.shape[0] == N
    np.testing.assert_array_equal(gpu.lengths, orc.len)
    np.testing.assert_array_equal(gpu.retcodes, orc.retcode)
    assert gpu.u_all.shape[2] == orc.u.shape[2], "record widths differ"
    cap = min(gpu.u_all.shape[1], orc.u.shape[1])
    idx = range(N) if n_check is None else np.linspace(0, N - 1, n_check).astype(int)
    for i in idx:
        n = min(int(gpu.lengths[i]), cap)
        gu, ou = gpu.u_all[i, :n], orc.u[i, :n]
        gt, ot = gpu.t_all[i, :n], orc.t[i, :n]
        # a record holds NaN where the field was NaN (FillExtrap outside the
        # grid); the adaptive step then turns t into NaN too, on both sides
        if not (np.array_equal(gu, ou, equal_nan=True) and np.array_equal(gt, ot, equal_nan=True)):
            d = np.nanmax(np.abs(gu - ou) / np.maximum(np.abs(ou), 1e-300))
            nan_u = int(np.sum(np.isnan(gu) != np.isnan(ou)))
            dt = np.nanmax(np.abs(gt - ot)) if gt.size else 0.0
            raise AssertionError(f"trajectory {i}: GPU and oracle differ (max rel {d:.3e}, "
                                 f"{nan_u} NaN mismatches, max |dt| {dt:.3e})")


def assert_rows_bitwise(gpu: TP.EnsembleSolution, idx, orc: O.OracleResult):
    """Rows idx of a device solution against an oracle run of those particles
    alone (row k of the oracle is particle idx[k]): trajectories are
    independent, so a sample checks a solve of any size."""
    idx = np.asarray(idx)
    np.testing.assert_array_equal(gpu.lengths[idx], orc.len)
    np.testing.assert_array_equal(gpu.retcodes[idx], orc.retcode)
    for k, i in enumerate(idx):
        n = int(orc.len[k])
        assert np.array_equal(gpu.u_all[i, :n], orc.u[k, :n], equal_nan=True), f"trajectory {i}: states differ"
        assert np.array_equal(gpu.t_all[i, :n], orc.t[k, :n], equal_nan=True), f"trajectory {i}: times differ"


def rel_err(a, b, floor=0.0):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def assert_close(gpu: TP.EnsembleSolution, orc: O.OracleResult, rtol, atol_pos=0.0, atol_vel=0.0):
    np.testing.assert_array_equal(gpu.lengths, orc.len)
    cap = min(gpu.u_all.shape[1], orc.u.shape[1])
    for i in range(len(gpu)):
        n = min(int(gpu.lengths[i]), cap)
        np.testing.assert_allclose(gpu.u_all[i, :n, :3], orc.u[i, :n, :3], rtol=rtol, atol=atol_pos)
        np.testing.assert_allclose(gpu.u_all[i, :n, 3:], orc.u[i, :n, 3:], rtol=rtol, atol=atol_vel)


__all__ = ["oracle_field", "run_oracle", "assert_bitwise", "assert_rows_bitwise", "assert_close", "rel_err",
           "_lib"]

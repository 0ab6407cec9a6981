"""The Python wrapper's high-level tracer on the device path: `trace()` and
`TraceResult` (bindings/python/src/testparticle/__init__.py:245-439).

The reference wrapper hands Python callables to Julia, which calls them at
every field evaluation.  A GPU kernel cannot call back into Python, so here a
field is a device field description (`UniformField`, `GridField`, ...), a
constant 3-vector, or a Python callable that `trace()` tabulates on the host
once, at the nodes of a grid the caller names (`grid=(x, y, z)`) and, for a
time-dependent callable, at the snapshot times `times` (a
`LazyTimeInterpolator` over the tabulated snapshots).  A callable is
time-dependent when it takes two positional arguments `(x, t)`, the rule of
the reference wrapper's `_wrap_field` (__init__.py:173-211).
"""
from __future__ import annotations

import inspect
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import algorithms as _alg
from .constants import Electron, Ion, Proton, Species
from .fields import (AbstractField, ClampExtrap, LazyTimeInterpolator, UniformField, ZeroField,
                     build_interpolator)
from .problem import TraceProblem
from .solve import solve

_SPECIES = {"proton": Proton, "electron": Electron, "ion": Ion}


@dataclass
class TraceResult:
    """Trajectory of a full particle traced with the Boris method
    (__init__.py:245-277): state columns x (3), v (3), then E (3) and B (3)
    with `save_fields`, then the work rates (4) with `save_work`."""

    t: np.ndarray
    state: np.ndarray
    save_fields: bool = False
    save_work: bool = False

    @property
    def x(self) -> np.ndarray:
        return self.state[:, 0:3]

    @property
    def v(self) -> np.ndarray:
        return self.state[:, 3:6]

    @property
    def E(self) -> np.ndarray:
        if not self.save_fields:
            raise AttributeError("E field was not saved; pass save_fields=True")
        return self.state[:, 6:9]

    @property
    def B(self) -> np.ndarray:
        if not self.save_fields:
            raise AttributeError("B field was not saved; pass save_fields=True")
        return self.state[:, 9:12]

    @property
    def work(self) -> np.ndarray:
        # the work columns follow x, v (and E, B when saved); the reference
        # reads columns 12:16 unconditionally, which are absent without
        # save_fields
        if not self.save_work:
            raise AttributeError("work was not saved; pass save_work=True")
        c = 12 if self.save_fields else 6
        return self.state[:, c:c + 4]


def _time_dependent(func) -> bool:
    """Two positional parameters (or *args): f(x, t) (__init__.py:186-200)."""
    try:
        params = inspect.signature(func).parameters.values()
    except (ValueError, TypeError):
        return False
    n_pos = sum(1 for p in params if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD))
    return n_pos >= 2 or any(p.kind == p.VAR_POSITIONAL for p in params)


def _tabulate(func, grid, t=None) -> np.ndarray:
    """A[c, i, j, k] = func((x_i, y_j, z_k)[, t]), evaluated on the host."""
    xs, ys, zs = (np.asarray(g, dtype=np.float64) for g in grid)
    A = np.empty((3, len(xs), len(ys), len(zs)))
    for i, x in enumerate(xs):
        for j, y in enumerate(ys):
            for k, z in enumerate(zs):
                v = func((x, y, z)) if t is None else func((x, y, z), t)
                A[:, i, j, k] = np.asarray(v, dtype=np.float64).reshape(3)
    return A


def _as_device_field(f, grid, times, bc):
    if f is None:
        return None
    if isinstance(f, AbstractField):
        return f
    if not callable(f):
        a = np.asarray(f, dtype=np.float64)
        if a.shape == (3,):
            return UniformField(a)
        raise TypeError("a field must be a device field, a 3-vector or a callable")
    if grid is None:
        raise TypeError("a Python callable cannot run on the GPU: pass grid=(x, y, z) to tabulate "
                        "it on the host, or describe the field with a device field "
                        "(UniformField, AffineField, DipoleField, GridField, ...)")
    if not _time_dependent(f):
        return build_interpolator(_tabulate(f, grid), *grid, bc=bc)
    if times is None:
        raise TypeError("a time-dependent callable f(x, t) needs the snapshot times `times` "
                        "to tabulate it")
    tt = np.asarray(times, dtype=np.float64)
    snaps = [build_interpolator(_tabulate(f, grid, float(t)), *grid, bc=bc) for t in tt]
    return LazyTimeInterpolator(tt, lambda i: snaps[i - 1])


def _as_alg(alg):
    if alg is None:
        return _alg.Boris()
    if isinstance(alg, _alg.AbstractBoris):
        return alg
    if isinstance(alg, str):
        a = getattr(_alg, alg, None)
        if a is None:
            raise _lib.ArgumentError(f"unknown solver {alg!r}")
        return a()
    raise TypeError("alg must be None, a solver name or a solver object")


def trace(x0, v0, tspan, dt, B, E=None, *, species="proton", trajectories=1, savestepinterval=1,
          save_fields=False, save_work=False, parallel=False, alg=None, init_func=None, q=None,
          m=None, grid=None, times=None, bc=None, precision="fp64", device=None):
    """Trace charged particles with the Boris pusher on the GPU
    (`trace`, __init__.py:330-439): same arguments and results.

    Extensions: `grid` / `times` / `bc` tabulate callable fields (see the
    module docstring; default boundary ClampExtrap, so a constant callable
    stays constant everywhere), `precision` ("fp64" | "fp32c"), `device`.
    `parallel` is accepted for compatibility: every trajectory runs on the
    GPU regardless.

    Returns a TraceResult when trajectories == 1 and init_func is None,
    otherwise a list of them."""
    x0 = np.asarray(x0, dtype=float)
    v0 = np.asarray(v0, dtype=float)
    if x0.shape != (3,) or v0.shape != (3,):
        raise ValueError("x0 and v0 must both have length 3")
    sp = _SPECIES[species.lower()] if isinstance(species, str) else species
    if not isinstance(sp, Species):
        raise TypeError("species must be 'proton', 'electron', or a Species")
    qq = sp.q if q is None else float(q)
    mm = sp.m if m is None else float(m)
    bc = ClampExtrap() if bc is None else bc
    fB = _as_device_field(B, grid, times, bc)
    fE = _as_device_field(E, grid, times, bc) or ZeroField()
    n = int(trajectories)
    if init_func is not None:
        # init_func(i) -> [x, y, z, vx, vy, vz] for trajectory i (1-based),
        # evaluated on the host (__init__.py:222-243)
        u0 = np.array([np.asarray(init_func(i), dtype=np.float64).reshape(6) for i in range(1, n + 1)])
    else:
        u0 = np.tile(np.concatenate([x0, v0]), (n, 1))
    prob = TraceProblem(u0, tuple(float(t) for t in tspan), (qq / mm, mm, fE, fB, ZeroField()))
    sol = solve(prob, _as_alg(alg), dt=float(dt), savestepinterval=int(savestepinterval),
                save_fields=bool(save_fields), save_work=bool(save_work), precision=precision,
                device=device)
    res = [TraceResult(s.t.copy(), s.u.copy(), bool(save_fields), bool(save_work)) for s in sol.u]
    if n == 1 and init_func is None:
        return res[0]
    return res


def trace_gc(*args, **kwargs):
    """The wrapper's guiding-centre tracer: not part of the device path."""
    raise NotImplementedError("the guiding-centre solver is not part of the device path; use the "
                              "reference's trace_gc")


def trace_fieldline(*args, **kwargs):
    """The wrapper's field-line tracer: not part of the device path."""
    raise NotImplementedError("field-line tracing is not part of the device path; use the "
                              "reference's trace_fieldline")

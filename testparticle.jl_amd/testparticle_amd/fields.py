"""Device field menu and `prepare` (src/prepare.jl, src/utility/interpolation.jl).

The reference carries fields as Julia callables inside p = (q2m, m, E, B, F).
A GPU cannot call a host closure, so fields here are descriptions the device
library evaluates: ZeroField, UniformField, AffineField, DipoleField,
GridField (a FieldInterpolator over a uniform CartesianGrid, order 1) and
UxBField (E = -u x B).  Each is still callable, E(r, t), evaluated on the
device, so code written against the reference's callables keeps working.
"""
from __future__ import annotations

import ctypes as C
import math
import sys

import numpy as np

from . import _lib
from .constants import Proton, Species


# ---------------------------------------------------------------------------
# Boundary-condition markers for gridded fields (interpolation.jl:203-206)
# ---------------------------------------------------------------------------
class FillExtrap:
    def __init__(self, fill_value=math.nan):
        if not (isinstance(fill_value, float) and math.isnan(fill_value)):
            raise NotImplementedError("the device path supports FillExtrap(NaN) only")
        self.fill_value = fill_value


class ClampExtrap:
    pass


class WrapExtrap:
    """Exclusive periodic wrapping, period N * dx."""


class PeriodicBC:
    def __init__(self, endpoint=":exclusive", check=True):
        if endpoint not in (":exclusive", "exclusive"):
            raise NotImplementedError("only PeriodicBC(endpoint=:exclusive) is supported")


def _extrap_code(bc) -> int:
    if bc is None or isinstance(bc, FillExtrap) or bc is FillExtrap:
        return _lib.EXTRAP_FILL_NAN
    if isinstance(bc, ClampExtrap) or bc is ClampExtrap:
        return _lib.EXTRAP_CLAMP
    if isinstance(bc, (WrapExtrap, PeriodicBC)) or bc in (WrapExtrap, PeriodicBC):
        return _lib.EXTRAP_PERIODIC
    raise _lib.ArgumentError(f"unsupported boundary condition {bc!r}")


# ---------------------------------------------------------------------------
# Field descriptions
# ---------------------------------------------------------------------------
class AbstractField:
    """A field the device library can evaluate.  Time-independent unless noted
    (is_time_dependent, src/prepare.jl:8-11)."""

    time_dependent = False

    def __init__(self):
        self._handles = {}

    def _create(self, ctx: _lib.Context):  # pragma: no cover - abstract
        raise NotImplementedError

    def handle(self, ctx: _lib.Context) -> C.c_void_p:
        h = self._handles.get(ctx.device)
        if h is None:
            h = C.c_void_p()
            with ctx.lock:
                ctx.check(self._create(ctx, h))
            self._handles[ctx.device] = (ctx, h)
            return h
        return h[1]

    def release(self):
        for ctx, h in self._handles.values():
            ctx.L.tpamd_field_free(ctx.h, h)
        self._handles = {}

    def __del__(self):
        # never call into the HIP runtime while the interpreter shuts down
        if sys.is_finalizing():
            return
        try:
            self.release()
        except Exception:
            pass

    def __call__(self, x, t=0.0, *, precision="fp64", device=0):
        """Evaluate at points x (shape (3,) or (n, 3)) on the device."""
        ctx = _lib.context(device)
        xa = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float64)))
        if xa.shape[-1] != 3:
            raise _lib.ArgumentError("field points must have 3 coordinates")
        out = np.empty_like(xa)
        B = self.src if isinstance(self, UxBField) else None
        hb = B.handle(ctx) if B is not None else None
        hf = self.handle(ctx)
        prec = _lib.FP64 if precision == "fp64" else _lib.FP32C
        with ctx.lock:
            ctx.check(ctx.L.tpamd_field_eval(ctx.h, hf, hb, _lib.dptr(xa), xa.shape[0],
                                             float(t), prec, _lib.dptr(out)))
        return out[0] if np.ndim(x) == 1 else out


class ZeroField(AbstractField):
    """ZeroField (src/utility/zero.jl:1,35-51)."""

    def _create(self, ctx, h):
        return ctx.L.tpamd_field_zero(ctx.h, C.byref(h))


class UniformField(AbstractField):
    def __init__(self, value):
        super().__init__()
        self.value = np.asarray(value, dtype=np.float64).reshape(3).copy()

    def _create(self, ctx, h):
        return ctx.L.tpamd_field_uniform(ctx.h, _lib.dptr(self.value), C.byref(h))


class AffineField(AbstractField):
    """F(x, t) = f0 + J x + dfdt t."""

    def __init__(self, f0, jacobian=None, dfdt=None):
        super().__init__()
        self.f0 = np.asarray(f0, dtype=np.float64).reshape(3).copy()
        self.jac = np.zeros(9) if jacobian is None else np.asarray(jacobian, np.float64).reshape(9).copy()
        self.ft = np.zeros(3) if dfdt is None else np.asarray(dfdt, np.float64).reshape(3).copy()
        self.time_dependent = bool(np.any(self.ft != 0))

    def _create(self, ctx, h):
        return ctx.L.tpamd_field_affine(ctx.h, _lib.dptr(self.f0), _lib.dptr(self.jac),
                                        _lib.dptr(self.ft), C.byref(h))


class DipoleField(AbstractField):
    """Centred dipole with B = (0, 0, b0) on the equator at r = radius and
    moment along -z (Earth: BMoment_Earth = (0, 0, -7.94e22))."""

    def __init__(self, b0: float = 3.12e-5, radius: float = 6371.0e3):
        super().__init__()
        self.b0, self.radius = float(b0), float(radius)

    def _create(self, ctx, h):
        return ctx.L.tpamd_field_dipole(ctx.h, self.b0, self.radius, C.byref(h))


class GridField(AbstractField):
    """Trilinear FieldInterpolator over a uniform CartesianGrid
    (build_interpolator(CartesianGrid, A::4D, x, y, z, 1, bc),
    src/utility/interpolation.jl:217-235).

    `A` follows the reference's layout: A[c, i, j, k] is component c at node
    (x[i], y[j], z[k]) (Julia's 3 x nx x ny x nz array).  `storage` chooses
    the device storage precision ("f64" or "f32").  `layout="cells"` adds a
    cell-replicated copy on the device (8 corners x 3 components per cell in
    one block: one cache line per gather for float storage, ~10x the memory);
    results are identical to `layout="nodes"`."""

    def __init__(self, A, x, y, z, bc=None, storage=None, layout="nodes"):
        super().__init__()
        if layout not in ("nodes", "cells"):
            raise _lib.ArgumentError("layout must be 'nodes' or 'cells'")
        self.layout = layout
        A = np.asarray(A)
        if A.ndim != 4 or A.shape[0] != 3:
            raise _lib.ArgumentError("Incompatible 3D force field and grid!")
        if A.dtype not in (np.float64, np.float32):
            A = A.astype(np.float64)
        nx, ny, nz = A.shape[1:]
        grids = [np.asarray(g, dtype=np.float64) for g in (x, y, z)]
        for g, n in zip(grids, (nx, ny, nz)):
            if g.shape != (n,):
                raise _lib.ArgumentError("Incompatible 3D force field and grid!")
            if n < 2:
                raise _lib.ArgumentError("each grid vector needs at least 2 nodes")
            if np.any(np.diff(g) <= 0):
                raise _lib.ArgumentError("Grid vectors must be sorted.")
            d = np.diff(g)
            if not np.allclose(d, d[0], rtol=1e-9, atol=0):
                raise NotImplementedError("the device path supports uniform (range) grids only")
        self.shape = (nx, ny, nz)
        self.origin = np.array([g[0] for g in grids])
        self.spacing = np.array([(g[-1] - g[0]) / (len(g) - 1) for g in grids])
        self.extrap = _extrap_code(bc)
        # device layout: [k][j][i][c], component fastest
        self.data = np.ascontiguousarray(A.transpose(3, 2, 1, 0))
        self.storage = storage or ("f32" if A.dtype == np.float32 else "f64")

    def _create(self, ctx, h):
        dt = _lib.F64 if self.data.dtype == np.float64 else _lib.F32
        st = _lib.F64 if self.storage == "f64" else _lib.F32
        rc = ctx.L.tpamd_field_grid(ctx.h, self.data.ctypes.data, dt, *self.shape,
                                    _lib.dptr(self.origin), _lib.dptr(self.spacing),
                                    self.extrap, st, C.byref(h))
        if rc == _lib.OK and self.layout == "cells":
            rc = ctx.L.tpamd_field_grid_cells(ctx.h, h, 1)
            if rc != _lib.OK:
                msg = ctx.L.tpamd_last_error(ctx.h)
                ctx.L.tpamd_field_free(ctx.h, h)
                h.value = None
                ctx.L.tpamd_last_error(ctx.h)
                raise NotImplementedError(msg.decode())
        return rc


FieldInterpolator = GridField


class LazyTimeInterpolator(AbstractField):
    """Linear interpolation in time between gridded snapshots
    (LazyTimeInterpolator(times, loader), src/utility/interpolation.jl:
    337-430).  `loader(i)` (1-based, as in Julia) returns the GridField of
    snapshot i; every snapshot must share the first one's grid, boundary
    condition and storage.  At time t, with k = searchsortedlast(times, t):
    k == 0 gives snapshot 1, k == len(times) the last snapshot, otherwise
    (1 - w) A_k(x) + w A_{k+1}(x), w = (t - t_k) / (t_{k+1} - t_k).

    The reference loads snapshots on demand and keeps at most three; here the
    first upload to a device calls the loader for every snapshot once and
    keeps them all resident in device memory, which the kernels read at any
    t without a host round trip.  The loader is called for snapshot 1 at
    construction, like the reference's constructor."""

    def __init__(self, times, loader):
        super().__init__()
        t = np.asarray(times, dtype=np.float64).reshape(-1).copy()
        if t.size < 1:
            raise _lib.ArgumentError("LazyTimeInterpolator needs at least one time")
        if not np.all(np.isfinite(t)):
            raise _lib.ArgumentError("snapshot times must be finite")
        if np.any(np.diff(t) <= 0):
            raise _lib.ArgumentError("snapshot times must be strictly increasing")
        self.times = t
        self.loader = loader
        self.time_dependent = t.size > 1
        first = self._load(1)
        self.shape, self.extrap, self.storage = first.shape, first.extrap, first.storage
        self.origin, self.spacing = first.origin.copy(), first.spacing.copy()
        self._first = first

    def _load(self, i: int) -> GridField:
        f = self.loader(i)
        if not isinstance(f, GridField):
            raise TypeError("the device path needs the loader to return gridded fields "
                            "(build_interpolator / GridField)")
        return f

    def snapshots(self):
        """Yields (index, GridField) for every snapshot, 1-based, checking that
        each shares the first snapshot's grid."""
        for i in range(1, self.times.size + 1):
            f = self._first if i == 1 else self._load(i)
            if (f.shape != self.shape or f.extrap != self.extrap or f.storage != self.storage
                    or not np.array_equal(f.origin, self.origin)
                    or not np.array_equal(f.spacing, self.spacing)):
                raise _lib.ArgumentError(f"snapshot {i} does not share the grid of snapshot 1")
            yield i, f

    def _create(self, ctx, h):
        datas = []
        for _, f in self.snapshots():
            d = f.data
            if d.dtype != self._first.data.dtype:
                d = d.astype(self._first.data.dtype)
            datas.append(d)
        ptrs = (C.c_void_p * len(datas))(*[d.ctypes.data for d in datas])
        dt = _lib.F64 if datas[0].dtype == np.float64 else _lib.F32
        st = _lib.F64 if self.storage == "f64" else _lib.F32
        return ctx.L.tpamd_field_time_grid(ctx.h, ptrs, dt, *self.shape, len(datas),
                                           _lib.dptr(self.times), _lib.dptr(self.origin),
                                           _lib.dptr(self.spacing), self.extrap, st, C.byref(h))


class UxBField(AbstractField):
    """E = -u x B(x) for a uniform drift velocity u (ideal-MHD electric field)."""

    def __init__(self, u, B: AbstractField):
        super().__init__()
        self.u = np.asarray(u, dtype=np.float64).reshape(3).copy()
        self.src = B

    def _create(self, ctx, h):
        return ctx.L.tpamd_field_uxb(ctx.h, _lib.dptr(self.u), self.src.handle(ctx), C.byref(h))


def build_interpolator(A, x, y, z, order: int = 1, bc=None, *, gridtype="CartesianGrid",
                       storage=None, layout="nodes") -> GridField:
    """build_interpolator(A, x, y, z, order=1, bc=FillExtrap(NaN))
    (src/utility/interpolation.jl:192-235) for the device: CartesianGrid,
    order 1."""
    if gridtype != "CartesianGrid":
        raise NotImplementedError(f"{gridtype} is not provided by the device path")
    if order != 1:
        raise NotImplementedError("the device path provides order=1 (trilinear) only")
    return GridField(A, x, y, z, bc=bc, storage=storage, layout=layout)


def _as_field(f, grids, order, bc, storage):
    if f is None:
        return ZeroField()
    if isinstance(f, AbstractField):
        return f
    if isinstance(f, np.ndarray) or (isinstance(f, (list, tuple)) and np.ndim(f) == 4):
        if grids is None:
            raise _lib.ArgumentError("gridded field needs grid vectors x, y, z")
        return build_interpolator(np.asarray(f), *grids, order=order, bc=bc, storage=storage)
    if isinstance(f, (list, tuple)) and len(f) == 3 and np.ndim(f) == 1:
        return UniformField(f)
    if callable(f):
        raise TypeError("a Python callable cannot run on the GPU: describe the field with "
                        "UniformField/AffineField/DipoleField/GridField/UxBField")
    raise TypeError(f"cannot use {type(f).__name__} as a field")


def prepare(*args, species: Species = Proton, q=None, m=None, order: int = 1, bc=None,
            gridtype="CartesianGrid", storage=None, E=None, F=None):
    """prepare(...) -> (q2m, m, E, B, F)  (src/prepare.jl:205-282).

    Supported call shapes:
      prepare(E, B[, F]; species, q, m)
      prepare(x, y, z, E, B[, F]; order=1, bc=FillExtrap(NaN))
      prepare(B; E=ZeroField(), F=ZeroField())
    """
    grids = None
    if len(args) >= 5 and all(np.ndim(a) == 1 for a in args[:3]):
        grids = tuple(np.asarray(g, dtype=np.float64) for g in args[:3])
        rest = args[3:]
        if len(rest) not in (2, 3):
            raise _lib.ArgumentError("prepare(x, y, z, E, B[, F])")
        Ef, Bf = rest[0], rest[1]
        Ff = rest[2] if len(rest) == 3 else F
    elif len(args) == 1:
        Ef, Bf, Ff = E, args[0], F
    elif len(args) in (2, 3):
        Ef, Bf = args[0], args[1]
        Ff = args[2] if len(args) == 3 else F
    else:
        raise _lib.ArgumentError("unsupported prepare() call")
    qq = species.q if q is None else float(q)
    mm = species.m if m is None else float(m)
    q2m = qq / mm
    fE = _as_field(Ef, grids, order, bc, storage)
    fB = _as_field(Bf, grids, order, bc, storage)
    fF = ZeroField() if Ff is None else Ff  # the Boris path ignores F (boris.jl:38-45)
    return (q2m, mm, fE, fB, fF)

"""Device field menu and `prepare` (src/prepare.jl, src/utility/interpolation.jl).

The reference carries fields as Julia callables inside p = (q2m, m, E, B, F).
A GPU cannot call a host closure, so fields here are descriptions the device
library evaluates: ZeroField, UniformField, AffineField, DipoleField,
GridField (a FieldInterpolator over a uniform CartesianGrid, order 1),
MeshField (rectilinear, cubic, spherical, 2D and 1D grids),
LazyTimeInterpolator (gridded snapshots blended in time) and UxBField
(E = -u x B).  Each is still callable, E(r, t), evaluated on the device, so
code written against the reference's callables keeps working.
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
    """PeriodicBC(endpoint=:exclusive): period N dx (node N+1 is node 1);
    endpoint=:inclusive: period x_N - x_1 (node N is node 1), meshes only."""

    def __init__(self, endpoint=":exclusive", check=True):
        if endpoint not in (":exclusive", "exclusive", ":inclusive", "inclusive"):
            raise NotImplementedError("PeriodicBC endpoint must be :exclusive or :inclusive")
        self.inclusive = endpoint in (":inclusive", "inclusive")


def _extrap_code(bc) -> int:
    if bc is None or isinstance(bc, FillExtrap) or bc is FillExtrap:
        return _lib.EXTRAP_FILL_NAN
    if isinstance(bc, ClampExtrap) or bc is ClampExtrap:
        return _lib.EXTRAP_CLAMP
    if isinstance(bc, PeriodicBC) and bc.inclusive:
        return _lib.EXTRAP_PERIODIC_INCLUSIVE
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
        if self.extrap == _lib.EXTRAP_PERIODIC_INCLUSIVE:
            raise NotImplementedError("the uniform grid wraps exclusively; use MeshField")
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


class MeshField(AbstractField):
    """Gridded field on a general mesh: the build_interpolator variants
    beyond the uniform trilinear grid (src/utility/interpolation.jl:15-117,
    140-178, 237-319).

      coord="cartesian": axes (x, y, z), any strictly increasing node
        vectors (RectilinearGrid, or CartesianGrid with vectors); order 1, or
        order 3 on uniform axes; an axis given as a single node is ignored,
        which gives the 2D and 1D interpolators (FieldInterpolator2D / 1D);
      coord="spherical": StructuredGrid with axes (r, theta, phi) and values
        (B_r, B_theta, B_phi); r and theta FillExtrap(NaN), phi periodic with
        period phi_N - phi_1, as the reference fixes them.

    `A` is (3, n1, n2, n3) in the reference's layout; `bc` one boundary for
    every Cartesian axis or a tuple of three.  Order 3 is the cardinal cubic
    (Catmull-Rom), with the linear extrapolation 2 f_1 - f_2 standing in for
    a neighbour past a non-periodic end."""

    def __init__(self, A, axes, coord="cartesian", order: int = 1, bc=None, storage=None):
        super().__init__()
        if coord not in ("cartesian", "spherical"):
            raise _lib.ArgumentError("coord must be 'cartesian' or 'spherical'")
        if order not in (0, 1, 3):
            raise _lib.ArgumentError(f"Interpolation order {order} is not supported. "
                                     "Supported orders are 0, 1, and 3.")
        if order == 0:
            raise NotImplementedError("order 0 (constant) interpolation is not provided by the device path")
        A = np.asarray(A)
        if A.ndim != 4 or A.shape[0] != 3:
            raise _lib.ArgumentError("Incompatible 3D force field and grid!")
        if A.dtype not in (np.float64, np.float32):
            A = A.astype(np.float64)
        self.axes = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1)) for a in axes]
        if len(self.axes) != 3:
            raise _lib.ArgumentError("a mesh needs three axes (a single node ignores an axis)")
        for a, n in zip(self.axes, A.shape[1:]):
            if a.shape != (n,):
                raise _lib.ArgumentError("Incompatible 3D force field and grid!")
            if np.any(np.diff(a) <= 0):
                raise _lib.ArgumentError("Grid vectors must be sorted.")
        self.shape = tuple(int(n) for n in A.shape[1:])
        self.coord = coord
        self.order = int(order)
        if coord == "spherical":
            self.extrap = (_lib.EXTRAP_FILL_NAN, _lib.EXTRAP_FILL_NAN, _lib.EXTRAP_PERIODIC_INCLUSIVE)
        elif isinstance(bc, (tuple, list)):
            self.extrap = tuple(_extrap_code(b) for b in bc)
        else:
            self.extrap = (_extrap_code(bc),) * 3
        self.data = np.ascontiguousarray(A.transpose(3, 2, 1, 0))
        self.storage = storage or ("f32" if A.dtype == np.float32 else "f64")

    @property
    def origin(self):
        return np.array([a[0] for a in self.axes])

    def _create(self, ctx, h):
        dt = _lib.F64 if self.data.dtype == np.float64 else _lib.F32
        st = _lib.F64 if self.storage == "f64" else _lib.F32
        n = (C.c_int64 * 3)(*self.shape)
        ax = (C.POINTER(C.c_double) * 3)(*[_lib.dptr(a) for a in self.axes])
        ex = (C.c_int32 * 3)(*self.extrap)
        coord = _lib.MESH_SPHERICAL if self.coord == "spherical" else _lib.MESH_CARTESIAN
        return ctx.L.tpamd_field_mesh(ctx.h, self.data.ctypes.data, dt, n, ax, coord, self.order,
                                      ex, st, C.byref(h))


def _uniform(g) -> bool:
    g = np.asarray(g, dtype=np.float64)
    if g.size < 2:
        return False
    d = np.diff(g)
    return bool(np.allclose(d, d[0], rtol=1e-9, atol=0))


def build_interpolator(A, *grids, order: int = 1, bc=None, gridtype=None, dir: int = 1,
                       storage=None, layout="nodes") -> AbstractField:
    """build_interpolator([gridtype,] A, grids..., order=1, bc=FillExtrap(NaN))
    (src/utility/interpolation.jl:192-319) for the device.

    gridtype "CartesianGrid" (default), "RectilinearGrid" or "StructuredGrid";
    3, 2 or 1 grid vectors (1D: `dir` picks the axis).  A uniform Cartesian
    3D grid at order 1 gives a GridField (the fast trilinear path); every
    other combination a MeshField."""
    if isinstance(A, str):               # build_interpolator(gridtype, A, grids...)
        gridtype, A, grids = A, grids[0], grids[1:]
    gridtype = gridtype or "CartesianGrid"
    grids = list(grids)
    trailing = []                        # positional order and bc after the grids
    while grids and np.ndim(grids[-1]) == 0:
        trailing.insert(0, grids.pop())
    if trailing:
        order = int(trailing[0])
        if len(trailing) > 1:
            bc = trailing[1]
    if gridtype not in ("CartesianGrid", "RectilinearGrid", "StructuredGrid"):
        raise NotImplementedError(f"{gridtype} is not provided by the device path")
    if order not in (0, 1, 3):
        raise _lib.ArgumentError(f"Interpolation order {order} is not supported. "
                                 "Supported orders are 0, 1, and 3.")
    A = np.asarray(A)
    if gridtype == "RectilinearGrid" and order != 1:
        raise _lib.ArgumentError("RectilinearGrid (CartesianNonUniform) only supports order=1 (Linear) "
                                 "interpolation.")
    if gridtype == "StructuredGrid":
        if len(grids) != 3:
            raise _lib.ArgumentError("StructuredGrid needs grids r, θ, ϕ")
        r, th, ph = (np.asarray(g, dtype=np.float64) for g in grids)
        if not r.min() >= 0:
            raise _lib.ArgumentError("r must be non-negative.")
        if not (th.min() >= 0 and th.max() <= math.pi):
            raise _lib.ArgumentError("θ must be within [0, π].")
        if not (ph.min() >= 0 and ph.max() <= 2 * math.pi):
            raise _lib.ArgumentError("ϕ must be within [0, 2π].")
        return MeshField(A, (r, th, ph), coord="spherical", order=order, storage=storage)
    if len(grids) == 3:
        if (gridtype == "CartesianGrid" and order == 1 and all(_uniform(g) for g in grids)
                and _extrap_code(bc) != _lib.EXTRAP_PERIODIC_INCLUSIVE):
            return GridField(A, *grids, bc=bc, storage=storage, layout=layout)
        return MeshField(A, grids, order=order, bc=bc, storage=storage)
    if len(grids) == 2:                  # FieldInterpolator2D: (3, nx, ny), z ignored
        if A.ndim != 3 or A.shape[0] != 3:
            raise _lib.ArgumentError("Incompatible 2D force field and grid!")
        return MeshField(A[..., None], (grids[0], grids[1], [0.0]), order=order, bc=bc,
                         storage=storage)
    if len(grids) == 1:                  # FieldInterpolator1D: (3, n) along axis `dir`
        if A.ndim != 2 or A.shape[0] != 3:
            raise _lib.ArgumentError("Incompatible 1D force field and grid!")
        if dir not in (1, 2, 3):
            raise _lib.ArgumentError("dir must be 1, 2 or 3")
        shape = [1, 1, 1]
        shape[dir - 1] = A.shape[1]
        axes = [[0.0], [0.0], [0.0]]
        axes[dir - 1] = grids[0]
        return MeshField(A.reshape(3, *shape), axes, order=order, bc=bc, storage=storage)
    raise _lib.ArgumentError("build_interpolator needs 1, 2 or 3 grid vectors")


def _as_field(f, grids, order, bc, storage, gridtype="CartesianGrid", dir=1):
    if f is None:
        return ZeroField()
    if isinstance(f, AbstractField):
        return f
    if grids is not None and np.ndim(f) == len(grids) + 1:
        return build_interpolator(np.asarray(f), *grids, order=order, bc=bc, storage=storage,
                                  gridtype=gridtype, dir=dir)
    if isinstance(f, np.ndarray) and f.ndim >= 2:
        raise _lib.ArgumentError("gridded field needs grid vectors")
    if isinstance(f, (list, tuple)) and len(f) == 3 and np.ndim(f) == 1:
        return UniformField(f)
    if callable(f):
        raise TypeError("a Python callable cannot run on the GPU: describe the field with "
                        "UniformField/AffineField/DipoleField/GridField/UxBField")
    raise TypeError(f"cannot use {type(f).__name__} as a field")


def _n_grids(args) -> int:
    """Leading grid vectors of a prepare() call: the largest k <= 3 with k
    1-D vectors followed by a field array of k + 1 dimensions."""
    for k in (3, 2, 1):
        if len(args) >= k + 2 and all(np.ndim(a) == 1 for a in args[:k]):
            if any(np.ndim(f) == k + 1 for f in args[k:k + 3] if not isinstance(f, AbstractField)):
                return k
    return 0


def prepare(*args, species: Species = Proton, q=None, m=None, order: int = 1, bc=None,
            gridtype="CartesianGrid", storage=None, E=None, F=None, dir: int = 1):
    """prepare(...) -> (q2m, m, E, B, F)  (src/prepare.jl:105-170,205-282).

    Supported call shapes:
      prepare(E, B[, F]; species, q, m)
      prepare(x, y, z, E, B[, F]; order=1, bc=FillExtrap(NaN), gridtype)
      prepare(x, y, E, B[, F]; order=1, bc=FillExtrap(NaN))
      prepare(x, E, B[, F]; order=1, bc=ClampExtrap(), dir=1)
      prepare(B; E=ZeroField(), F=ZeroField())
    """
    grids = None
    k = _n_grids(args)
    if k:
        grids = tuple(np.asarray(g, dtype=np.float64) for g in args[:k])
        for g in grids:
            if np.any(np.diff(g) < 0):
                raise _lib.ArgumentError("Grid vector must be sorted.")
        rest = args[k:]
        if len(rest) not in (2, 3):
            raise _lib.ArgumentError("prepare(grids..., E, B[, F])")
        Ef, Bf = rest[0], rest[1]
        Ff = rest[2] if len(rest) == 3 else F
        if k == 1 and bc is None:
            bc = ClampExtrap()           # the 1D method's default (prepare.jl:165)
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
    gt = gridtype if k == 3 else "CartesianGrid"
    fE = _as_field(Ef, grids, order, bc, storage, gt, dir)
    fB = _as_field(Bf, grids, order, bc, storage, gt, dir)
    fF = ZeroField() if Ff is None else Ff  # the Boris path ignores F (boris.jl:38-45)
    return (q2m, mm, fE, fB, fF)

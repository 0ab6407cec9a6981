"""TraceProblem (src/types.jl:23-49) and the isoutside menu."""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib


def DEFAULT_PROB_FUNC(prob, ctx):
    return prob


@dataclass
class EnsembleContext:
    """What prob_func receives (SciMLBase EnsembleContext: sim_id is 1-based)."""
    sim_id: int
    repeat: int = 1
    rng: object = None
    seed: object = None


@dataclass
class TraceProblem:
    """TraceProblem(u0, tspan, p; prob_func) with p = (q2m, m, E, B, F).

    u0 is the 6-vector (x, y, z, vx, vy, vz).  As an extension for large
    ensembles u0 may also be an (N, 6) array of initial states, one row per
    trajectory (prob_func is then not applied)."""
    u0: object
    tspan: tuple
    p: tuple
    prob_func: object = DEFAULT_PROB_FUNC

    def __post_init__(self):
        self.tspan = tuple(self.tspan)
        if len(self.tspan) != 2:
            raise _lib.ArgumentError("tspan must be (t0, t1)")


def remake(prob: TraceProblem, **kw) -> TraceProblem:
    """SciMLBase.remake for TraceProblem."""
    return replace(prob, **kw)


# ---------------------------------------------------------------------------
# isoutside(u, p, t) menu (boris.jl:95,457-461; adaptive_boris.jl:26,227-231)
# ---------------------------------------------------------------------------
@dataclass
class Boundary:
    """A device-evaluable isoutside predicate; combine with `|`."""
    flags: int = 0
    r_in: float = 0.0
    r_out: float = math.inf
    box_lo: tuple = (-math.inf, -math.inf, -math.inf)
    box_hi: tuple = (math.inf, math.inf, math.inf)

    def __or__(self, other: "Boundary") -> "Boundary":
        lo = tuple(max(a, b) for a, b in zip(self.box_lo, other.box_lo))
        hi = tuple(min(a, b) for a, b in zip(self.box_hi, other.box_hi))
        return Boundary(self.flags | other.flags, max(self.r_in, other.r_in),
                        min(self.r_out, other.r_out), lo, hi)

    def __call__(self, u, p=None, t=None) -> bool:
        """Host evaluation of the same predicate (for checks and tests)."""
        u = np.asarray(u, dtype=np.float64)
        x = u[:3]
        if self.flags & _lib.BND_NAN and np.any(np.isnan(u[:6])):
            return True
        r = math.sqrt(float((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]))
        if self.flags & _lib.BND_SPHERE_IN and r < self.r_in:
            return True
        if self.flags & _lib.BND_SPHERE_OUT and r > self.r_out:
            return True
        if self.flags & _lib.BND_BOX:
            if any(x[c] < self.box_lo[c] or x[c] > self.box_hi[c] for c in range(3)):
                return True
        return False


def OutsideNaN() -> Boundary:
    """any(isnan, u)."""
    return Boundary(flags=_lib.BND_NAN)


def InsideSphere(radius: float) -> Boundary:
    """norm(x) < radius (e.g. loss to the planet, r < R_E)."""
    return Boundary(flags=_lib.BND_SPHERE_IN, r_in=float(radius))


def OutsideSphere(radius: float) -> Boundary:
    """norm(x) > radius."""
    return Boundary(flags=_lib.BND_SPHERE_OUT, r_out=float(radius))


def OutsideBox(lo, hi) -> Boundary:
    """x outside [lo, hi] in any component (open-boundary exit)."""
    return Boundary(flags=_lib.BND_BOX, box_lo=tuple(float(a) for a in lo),
                    box_hi=tuple(float(a) for a in hi))


def OutsideGrid() -> Boundary:
    """x outside the node span of the solve's gridded B field (each instance's
    own grid in a batched solve): particles exit at the grid faces."""
    return Boundary(flags=_lib.BND_GRID)


@dataclass
class TerminateOutside:
    """TerminateOutside(isoutofdomain) (src/utility/utility.jl:589-591): wraps a
    predicate; `.condition` is what solve(...; isoutside=) takes."""
    condition: Boundary = field(default_factory=Boundary)

"""Post-processing of saved trajectories: get_fields / get_work
(src/boris/boris.jl:599-636).  The reference maps its field callables and
get_work_rates (src/equations.jl:394-441) over (sol.u, sol.t) on the host; here
one tpamd_record_eval call evaluates every record on the device, with the same
formulas the solver uses for save_fields / save_work records."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def record_eval(prob, xv, t, *, want_fields=True, want_work=True, device: int = 0) -> np.ndarray:
    """The extra record columns at states xv (n, 6) and times t (n,) for the
    fields and q/m, m of `prob`: (n, 6) E, B and/or (n, 4) work rates."""
    from .solve import _fields

    E, B = _fields(prob)
    xv = np.ascontiguousarray(np.atleast_2d(np.asarray(xv, np.float64))[:, :6])
    n = xv.shape[0]
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, np.float64), (n,)))
    w = (6 if want_fields else 0) + (4 if want_work else 0)
    out = np.zeros((n, w))
    ctx = _lib.context(device)
    hE, hB = E.handle(ctx), B.handle(ctx)
    q2m = float(prob.p[0])
    m = float(prob.p[1]) if len(prob.p) > 1 else 0.0
    with ctx.lock:
        rc = ctx.L.tpamd_record_eval(ctx.h, hE, hB, q2m, m, _lib.dptr(xv), _lib.dptr(t), n,
                                     int(want_fields), int(want_work), _lib.dptr(out))
    ctx.check(rc)
    return out


def get_fields(sol):
    """(E, B) at every saved record of a TraceSolution: two (n, 3) arrays."""
    r = record_eval(sol.prob, np.asarray(sol.u), np.asarray(sol.t), want_fields=True, want_work=False)
    return r[:, :3], r[:, 3:6]


def get_work(sol):
    """(P_par, P_fermi, P_grad, P_betatron) at every saved record: (n, 4)."""
    return record_eval(sol.prob, np.asarray(sol.u), np.asarray(sol.t), want_fields=False, want_work=True)

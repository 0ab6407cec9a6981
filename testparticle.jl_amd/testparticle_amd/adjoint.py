"""Reverse-mode gradients of a trajectory loss (tpamd_adjoint).

L = (1/N) sum_i |x_i(T) - x*_i|^2 over the final positions of a fixed-step
standard Boris solve with E = 0 and a gridded B.  The reference obtains such
derivatives with ForwardDiff through its CPU solve (test/test_derivatives.jl);
here one forward pass with FP64 checkpoints every K steps and one reverse
sweep give dL/dB at every grid node and dL/d(x0, v0) for every particle.

`trajectory_loss` wraps the same call as a torch.autograd.Function (the
Python counterpart of a ChainRulesCore rrule): gradients flow to the B-grid
tensor and to the initial states.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .algorithms import Boris
from .fields import GridField, ZeroField
from .solve import initial_states, make_desc


@dataclass
class AdjointResult:
    loss: float
    dL_dB: np.ndarray        # A[c, i, j, k] layout, like the field array given to GridField
    dL_dxv0: np.ndarray      # (N, 6)
    forward_ms: float
    reverse_ms: float
    stats: object


def loss_gradient(prob, *, dt, targets, precision: str = "fp64", ckpt_interval: int = 64,
                  device: int = 0) -> AdjointResult:
    """One forward pass and one reverse sweep.  precision "fp64" differentiates
    the FP64 solve (what ForwardDiff through the reference computes);
    "fp32c" runs the forward, recompute and reverse steps in float over float
    storage, with the adjoint state scaled by a per-particle power of two.
    Raises FloatingPointError when the loss or a gradient entry is not finite."""
    q2m, _m, E, B = prob.p[:4]
    if not isinstance(E, ZeroField):
        raise NotImplementedError("the adjoint covers E = 0")
    if not isinstance(B, GridField):
        raise NotImplementedError("the adjoint needs a gridded B (GridField)")
    xv0 = initial_states(prob, 1)
    N = xv0.shape[0]
    tg = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, np.float64), (N, 3)))
    desc = make_desc(prob, Boris(), dt=dt, precision=precision, save_everystep=False)
    ctx = _lib.context(device)
    hB = B.handle(ctx)
    loss = np.zeros(1)
    gB = np.zeros(B.data.shape)            # device node layout [k][j][i][c]
    gx = np.zeros((N, 6))
    st = _lib.Stats()
    with ctx.lock:
        rc = ctx.L.tpamd_adjoint(ctx.h, C.byref(desc), hB, _lib.dptr(xv0), N, _lib.dptr(tg),
                                 int(ckpt_interval), _lib.dptr(loss), _lib.dptr(gB),
                                 _lib.dptr(gx), C.byref(st))
    ctx.check(rc, st)
    return AdjointResult(float(loss[0]), np.ascontiguousarray(gB.transpose(3, 2, 1, 0)), gx,
                         st.kernel_ms, st.aux_ms, st)


def trajectory_loss(B_values, xv0, *, axes, bc, q2m, tspan, dt, targets, precision="fp64",
                    ckpt_interval: int = 64):
    """torch autograd entry: loss(B_values (3, nx, ny, nz), xv0 (N, 6)) with
    gradients to both tensors from the adjoint sweep."""
    import torch

    from .problem import TraceProblem

    class _TrajectoryLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, Bv, x0):
            A = Bv.detach().cpu().numpy()
            field = GridField(A, *axes, bc=bc)
            prob = TraceProblem(x0.detach().cpu().numpy(), tspan, (q2m, 1.0, ZeroField(), field, ZeroField()))
            r = loss_gradient(prob, dt=dt, targets=targets, precision=precision,
                              ckpt_interval=ckpt_interval)
            ctx.save_for_backward(torch.from_numpy(r.dL_dB).to(Bv), torch.from_numpy(r.dL_dxv0).to(x0))
            return torch.tensor(r.loss, dtype=torch.float64, device=Bv.device)

        @staticmethod
        def backward(ctx, g):
            gB, gx = ctx.saved_tensors
            return g * gB, g * gx

    return _TrajectoryLoss.apply(B_values, xv0)

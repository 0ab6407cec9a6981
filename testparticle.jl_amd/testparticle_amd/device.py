"""Device-resident ensembles: initial states and records live in HBM (torch
tensors used as plain device buffers), and each run is one call of
tpamd_solve_device on the caller's stream.  For repeated solves (benchmarks,
optimisation loops) where host<->device copies are not part of the work."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solve import _fields, make_desc, plan


class DeviceEnsemble:
    def __init__(self, prob, alg, xv0: np.ndarray, *, device: int = 0, nout_cap=None,
                 dts=None, **desc_kw):
        import torch

        self.torch = torch
        self.ctx = _lib.context(device)
        self.dev = torch.device(f"cuda:{device}")
        self.E, self.B = _fields(prob)
        self.N = int(xv0.shape[0])
        self.desc = make_desc(prob, alg, dts=dts, **desc_kw)
        self.dts = None
        if dts is not None:
            # per-particle steps: the schedule comes from nsteps, the pointer
            # (device memory here) is only tested for presence by the planner
            self.dts = torch.from_numpy(np.ascontiguousarray(dts, dtype=np.float64)).to(self.dev)
            self.desc.dt_pp = self.dts.data_ptr()
        if alg.adaptive:
            cap = int(nout_cap if nout_cap is not None else 2)
        else:
            cap = plan(self.desc)[1]
        self.desc.nout_cap = cap
        self.cap = cap
        self.xv0 = torch.from_numpy(np.ascontiguousarray(xv0, dtype=np.float64)).to(self.dev)
        self.u = torch.empty((self.N, max(cap, 1), self.desc.record_width), dtype=torch.float64,
                             device=self.dev)
        self.t = torch.empty((self.N, max(cap, 1)), dtype=torch.float64, device=self.dev)
        self.len = torch.empty(self.N, dtype=torch.int32, device=self.dev)
        self.ret = torch.empty(self.N, dtype=torch.int32, device=self.dev)
        self.hE, self.hB = self.E.handle(self.ctx), self.B.handle(self.ctx)
        torch.cuda.synchronize(self.dev)

    def run(self) -> _lib.Stats:
        """One solve on torch's current stream; returns the library's stats
        (kernel_ms from HIP events on that stream)."""
        st = _lib.Stats()
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        with self.ctx.lock:
            rc = self.ctx.L.tpamd_solve_device(
                self.ctx.h, C.byref(self.desc), self.hE, self.hB, self.xv0.data_ptr(), self.N,
                self.u.data_ptr(), self.t.data_ptr(), self.len.data_ptr(), self.ret.data_ptr(),
                C.c_void_p(stream), C.byref(st))
        self.ctx.check(rc, st)
        return st

    def final_states(self) -> np.ndarray:
        idx = (self.len.long() - 1).clamp(min=0)
        ar = self.torch.arange(self.N, device=self.dev)
        return self.u[ar, idx].cpu().numpy()

    def host(self):
        return (self.u.cpu().numpy(), self.t.cpu().numpy(), self.len.cpu().numpy(),
                self.ret.cpu().numpy())


class DeviceBatch:
    """Many field instances in one persistent, graph-captured launch
    (tpamd_batch_*), with states and records resident in HBM."""

    def __init__(self, problems, alg, *, device: int = 0, nout_cap=None, **desc_kw):
        import torch

        from .solve import _adaptive_cap, initial_states

        self.torch = torch
        self.ctx = _lib.context(device)
        self.dev = torch.device(f"cuda:{device}")
        self.problems = problems
        fields = [_fields(pr) for pr in problems]
        states = [initial_states(pr, 1) for pr in problems]
        self.offsets = np.zeros(len(problems) + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum([s.shape[0] for s in states])
        self.N = int(self.offsets[-1])
        self.desc = make_desc(problems[0], alg, **desc_kw)
        if alg.adaptive:
            cap = nout_cap if nout_cap is not None else _adaptive_cap(
                desc_kw.get("save_start", True), desc_kw.get("save_everystep", True),
                desc_kw.get("max_steps", 0), desc_kw.get("savestepinterval", 1))
        else:
            cap = plan(self.desc)[1]
        self.cap = int(cap)
        self.desc.nout_cap = self.cap
        self.xv0 = torch.from_numpy(np.ascontiguousarray(np.concatenate(states))).to(self.dev)
        self.u = torch.empty((self.N, max(self.cap, 1), self.desc.record_width), dtype=torch.float64,
                             device=self.dev)
        self.t = torch.empty((self.N, max(self.cap, 1)), dtype=torch.float64, device=self.dev)
        self.len = torch.empty(self.N, dtype=torch.int32, device=self.dev)
        self.ret = torch.empty(self.N, dtype=torch.int32, device=self.dev)
        self._keep = fields
        hE = (C.c_void_p * len(problems))(*[f[0].handle(self.ctx).value for f in fields])
        hB = (C.c_void_p * len(problems))(*[f[1].handle(self.ctx).value for f in fields])
        torch.cuda.synchronize(self.dev)
        h = C.c_void_p()
        with self.ctx.lock:
            rc = self.ctx.L.tpamd_batch_create(
                self.ctx.h, C.byref(self.desc), len(problems), hE, hB,
                self.offsets.ctypes.data_as(C.POINTER(C.c_int64)), self.xv0.data_ptr(),
                self.u.data_ptr(), self.t.data_ptr(), self.len.data_ptr(), self.ret.data_ptr(),
                C.byref(h))
        self.ctx.check(rc)
        self.h = h

    def autotune(self):
        bs, ppt = C.c_int32(), C.c_int32()
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        with self.ctx.lock:
            rc = self.ctx.L.tpamd_batch_autotune(self.h, C.c_void_p(stream), C.byref(bs), C.byref(ppt))
        self.ctx.check(rc)
        return int(bs.value), int(ppt.value)

    def run(self) -> _lib.Stats:
        st = _lib.Stats()
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        with self.ctx.lock:
            rc = self.ctx.L.tpamd_batch_run(self.h, C.c_void_p(stream), C.byref(st))
        self.ctx.check(rc, st)
        return st

    def final_states(self) -> np.ndarray:
        idx = (self.len.long() - 1).clamp(min=0)
        ar = self.torch.arange(self.N, device=self.dev)
        return self.u[ar, idx].cpu().numpy()

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.tpamd_batch_free(self.h)
            self.h = None

    def __del__(self):
        import sys
        if not sys.is_finalizing():
            self.close()


def solve_batch_noalloc(problems, alg, *, device: int = 0, nout_cap=None, particles_per_lane: int = 0,
                        **desc_kw):
    """One batched solve through tpamd_solve_batch_noalloc: the kernels of
    DeviceBatch, with the library's scratch in its fixed device array, so the
    call allocates no device memory (the output tensors here are torch's).
    Returns (u, t, lengths, retcodes) device tensors, the offsets and the
    stats."""
    import torch

    from .solve import _adaptive_cap, initial_states

    ctx = _lib.context(device)
    dev = torch.device(f"cuda:{device}")
    fields = [_fields(pr) for pr in problems]
    states = [initial_states(pr, 1) for pr in problems]
    offsets = np.zeros(len(problems) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([s.shape[0] for s in states])
    N = int(offsets[-1])
    desc = make_desc(problems[0], alg, **desc_kw)
    if alg.adaptive:
        cap = nout_cap if nout_cap is not None else _adaptive_cap(
            desc_kw.get("save_start", True), desc_kw.get("save_everystep", True),
            desc_kw.get("max_steps", 0), desc_kw.get("savestepinterval", 1))
    else:
        cap = plan(desc)[1]
    desc.nout_cap = int(cap)
    desc.particles_per_lane = int(particles_per_lane)
    xv0 = torch.from_numpy(np.ascontiguousarray(np.concatenate(states))).to(dev)
    u = torch.empty((N, max(int(cap), 1), desc.record_width), dtype=torch.float64, device=dev)
    t = torch.empty((N, max(int(cap), 1)), dtype=torch.float64, device=dev)
    ln = torch.empty(N, dtype=torch.int32, device=dev)
    ret = torch.empty(N, dtype=torch.int32, device=dev)
    hE = (C.c_void_p * len(problems))(*[f[0].handle(ctx).value for f in fields])
    hB = (C.c_void_p * len(problems))(*[f[1].handle(ctx).value for f in fields])
    st = _lib.Stats()
    stream = torch.cuda.current_stream(dev).cuda_stream
    with ctx.lock:
        rc = ctx.L.tpamd_solve_batch_noalloc(
            ctx.h, C.byref(desc), len(problems), hE, hB, offsets.ctypes.data_as(C.POINTER(C.c_int64)),
            xv0.data_ptr(), u.data_ptr(), t.data_ptr(), ln.data_ptr(), ret.data_ptr(), C.c_void_p(stream),
            C.byref(st))
    ctx.check(rc, st)
    return u, t, ln, ret, offsets, st
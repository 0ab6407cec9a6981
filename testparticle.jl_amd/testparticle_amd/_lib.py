"""ctypes binding of the gfx950 C-ABI library (include/tpamd.h).

The library is the only compute path: if it is missing, every entry point
raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "_build", "libtpamd.so")

ABI_VERSION = 1

# tpamd_status
OK, ERR_ARGUMENT, ERR_DT_TOO_SMALL, ERR_MAXITERS, ERR_ORDER = 0, 1, 2, 3, 4
ERR_DEVICE, ERR_OOM, ERR_UNSUPPORTED, ERR_CAPACITY, ERR_NONFINITE = 5, 6, 7, 8, 9
# enums
FIELD_ZERO, FIELD_UNIFORM, FIELD_AFFINE, FIELD_DIPOLE, FIELD_GRID, FIELD_UXB = range(6)
FIELD_TIME_GRID, FIELD_MESH = 6, 7
EXTRAP_FILL_NAN, EXTRAP_CLAMP, EXTRAP_PERIODIC, EXTRAP_PERIODIC_INCLUSIVE = 0, 1, 2, 3
MESH_CARTESIAN, MESH_SPHERICAL = 0, 1
F64, F32 = 0, 1
ALG_BORIS, ALG_MULTISTEP = 0, 1
FORM_REFERENCE, FORM_RATIONAL = 0, 1
FP64, FP32C = 0, 1
BND_NAN, BND_SPHERE_IN, BND_SPHERE_OUT, BND_BOX, BND_GRID = 1, 2, 4, 8, 16
RET_SUCCESS, RET_TERMINATED, RET_MAXITERS = 0, 1, 2
FO_ENTER, FO_CHECKED = 1, 2                     # tpamd_fo_segment flags in
FO_CHECK, FO_END, FO_OUTSIDE = 0, 1, 2          # and status out
FLAG_GENERIC_KERNELS = 1
RETCODE_NAMES = {RET_SUCCESS: "Success", RET_TERMINATED: "Terminated", RET_MAXITERS: "MaxIters"}


class ArgumentError(ValueError):
    """Mirror of Julia's ArgumentError raised by the reference's checks."""


class DeviceError(RuntimeError):
    pass


class CapacityError(RuntimeError):
    def __init__(self, msg, needed):
        super().__init__(msg)
        self.needed = needed


class SolveDesc(C.Structure):
    _fields_ = [
        ("alg", C.c_int32), ("adaptive", C.c_int32), ("order_N", C.c_int32),
        ("nsub", C.c_int32), ("form", C.c_int32), ("precision", C.c_int32),
        ("save_start", C.c_int32), ("save_end", C.c_int32), ("save_everystep", C.c_int32),
        ("bnd_flags", C.c_int32), ("safety", C.c_double), ("q2m", C.c_double),
        ("t0", C.c_double), ("t1", C.c_double), ("dt", C.c_double),
        ("maxiters", C.c_int64), ("savestepinterval", C.c_int64), ("max_steps", C.c_int64),
        ("nout_cap", C.c_int64), ("r_in", C.c_double), ("r_out", C.c_double),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
        ("dt_pp", C.c_void_p), ("nsteps_pp", C.c_int64), ("block_size", C.c_int32),
        ("flags", C.c_int32), ("mass", C.c_double), ("save_fields", C.c_int32),
        ("save_work", C.c_int32), ("particles_per_lane", C.c_int32), ("reserved", C.c_int32),
    ]

    @property
    def record_width(self) -> int:
        return 6 + (6 if self.save_fields else 0) + (4 if self.save_work else 0)


class Stats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("particle_steps", C.c_int64),
        ("nt", C.c_int64), ("nout", C.c_int64), ("max_len", C.c_int64),
        ("aux_ms", C.c_double),
    ]


EXPORTS = [
    "tpamd_abi_version", "tpamd_init", "tpamd_finalize", "tpamd_last_error",
    "tpamd_field_zero", "tpamd_field_uniform", "tpamd_field_affine", "tpamd_field_dipole",
    "tpamd_field_grid", "tpamd_field_time_grid", "tpamd_field_mesh", "tpamd_field_grid_cells",
    "tpamd_field_uxb",
    "tpamd_field_eval",
    "tpamd_field_free", "tpamd_record_eval",
    "tpamd_plan", "tpamd_solve", "tpamd_solve_device", "tpamd_adjoint",
    "tpamd_batch_create", "tpamd_batch_autotune", "tpamd_batch_run", "tpamd_batch_free",
    "tpamd_solve_batch", "tpamd_fo_segment", "tpamd_solve_batch_noalloc",
]

_lib = None
_lock = threading.Lock()


def _bind_one_hip_runtime():
    """PyTorch-ROCm wheels ship their own libamdhip64 whose soname
    (libamdhip64.so.7) equals the system one the library links against, but
    torch requests it by a different file name.  Loaded first, torch's copy
    satisfies the library's dependency and the process holds ONE HIP runtime;
    loaded after the library, torch would bring a second HIP/HSA runtime that
    finds no device.  So torch, when installed, is imported before the
    library is opened.  Set TPAMD_NO_TORCH=1 to skip this."""
    if os.environ.get("TPAMD_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(path: str = LIB_PATH):
    """Load and type the library; raises if it was not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"HIP library {path} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
        _bind_one_hip_runtime()
        L = C.CDLL(path)
        vp, d3, i64 = C.c_void_p, C.POINTER(C.c_double), C.c_int64
        pp = C.POINTER(C.c_void_p)
        L.tpamd_abi_version.restype = C.c_int32
        L.tpamd_init.argtypes = [C.c_int32, pp]
        L.tpamd_finalize.argtypes = [vp]
        L.tpamd_last_error.argtypes = [vp]
        L.tpamd_last_error.restype = C.c_char_p
        L.tpamd_field_zero.argtypes = [vp, pp]
        L.tpamd_field_uniform.argtypes = [vp, d3, pp]
        L.tpamd_field_affine.argtypes = [vp, d3, d3, d3, pp]
        L.tpamd_field_dipole.argtypes = [vp, C.c_double, C.c_double, pp]
        L.tpamd_field_grid.argtypes = [vp, vp, C.c_int32, i64, i64, i64, d3, d3, C.c_int32,
                                       C.c_int32, pp]
        L.tpamd_field_time_grid.argtypes = [vp, pp, C.c_int32, i64, i64, i64, C.c_int32, d3, d3,
                                            d3, C.c_int32, C.c_int32, pp]
        L.tpamd_field_mesh.argtypes = [vp, vp, C.c_int32, C.POINTER(i64), C.POINTER(d3), C.c_int32,
                                       C.c_int32, C.POINTER(C.c_int32), C.c_int32, pp]
        L.tpamd_field_uxb.argtypes = [vp, d3, vp, pp]
        L.tpamd_field_grid_cells.argtypes = [vp, vp, C.c_int32]
        L.tpamd_adjoint.argtypes = [vp, C.POINTER(SolveDesc), vp, d3, i64, d3, C.c_int32, d3,
                                    d3, d3, C.POINTER(Stats)]
        L.tpamd_field_eval.argtypes = [vp, vp, vp, d3, i64, C.c_double, C.c_int32, d3]
        L.tpamd_field_free.argtypes = [vp, vp]
        L.tpamd_record_eval.argtypes = [vp, vp, vp, C.c_double, C.c_double, d3, d3, i64, C.c_int32,
                                        C.c_int32, d3]
        L.tpamd_plan.argtypes = [vp, C.POINTER(SolveDesc), C.POINTER(i64), C.POINTER(i64)]
        L.tpamd_solve.argtypes = [vp, C.POINTER(SolveDesc), vp, vp, vp, i64, vp, vp, vp, vp,
                                  C.POINTER(Stats)]
        L.tpamd_solve_device.argtypes = [vp, C.POINTER(SolveDesc), vp, vp, vp, i64, vp, vp, vp,
                                         vp, vp, C.POINTER(Stats)]
        vpa = C.POINTER(C.c_void_p)
        L.tpamd_batch_create.argtypes = [vp, C.POINTER(SolveDesc), C.c_int32, vpa, vpa,
                                         C.POINTER(i64), vp, vp, vp, vp, vp, pp]
        L.tpamd_batch_autotune.argtypes = [vp, vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.tpamd_batch_run.argtypes = [vp, vp, C.POINTER(Stats)]
        L.tpamd_batch_free.argtypes = [vp]
        L.tpamd_solve_batch.argtypes = [vp, C.POINTER(SolveDesc), C.c_int32, vpa, vpa,
                                        C.POINTER(i64), vp, vp, vp, vp, vp, C.POINTER(Stats)]
        L.tpamd_solve_batch_noalloc.argtypes = [vp, C.POINTER(SolveDesc), C.c_int32, vpa, vpa,
                                                C.POINTER(i64), vp, vp, vp, vp, vp, vp,
                                                C.POINTER(Stats)]
        L.tpamd_fo_segment.argtypes = [vp, C.POINTER(SolveDesc), vp, vp, i64, C.c_double, C.c_double,
                                       i64, d3, C.POINTER(i64), C.POINTER(C.c_int32), d3, d3, d3,
                                       C.POINTER(C.c_int32), C.POINTER(Stats)]
        for name in EXPORTS:
            if name not in ("tpamd_abi_version", "tpamd_last_error"):
                getattr(L, name).restype = C.c_int32
        v = L.tpamd_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libtpamd ABI {v} != expected {ABI_VERSION}")
        _lib = L
        return L


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def iptr64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Context:
    """A device context (tpamd_ctx): one per HIP device, owns its stream."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.tpamd_init(int(device), C.byref(h))
        if rc:
            raise DeviceError(self.L.tpamd_last_error(None).decode())
        self.h = h
        self.device = device
        # reentrant: creating a field handle can create its source field's
        # handle (UxB) while the lock is held
        self.lock = threading.RLock()

    def check(self, rc: int, stats: Stats | None = None):
        if rc == OK:
            return
        msg = self.L.tpamd_last_error(self.h).decode()
        if rc in (ERR_ARGUMENT, ERR_DT_TOO_SMALL, ERR_MAXITERS, ERR_ORDER):
            raise ArgumentError(msg)
        if rc == ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        if rc == ERR_CAPACITY:
            raise CapacityError(msg, int(stats.max_len) if stats is not None else -1)
        if rc == ERR_NONFINITE:
            raise FloatingPointError(msg)
        raise DeviceError(msg)

    def close(self):
        if self.h:
            self.L.tpamd_finalize(self.h)
            self.h = C.c_void_p()


_contexts: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    with _lock:
        c = _contexts.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _contexts.setdefault(device, c)
            c = _contexts[device]
    return c

"""testparticle_amd -- TestParticle.jl's ensemble Boris pusher on the MI355X.

Host-side mirror of the reference's API for this path:

    import testparticle_amd as TP
    param = TP.prepare(TP.ZeroField(), TP.UniformField([0, 0, 0.01]), species=TP.Electron)
    prob = TP.TraceProblem(u0, (0.0, 3e-8), param)
    sol = TP.solve(prob, TP.Boris(), dt=3e-11, savestepinterval=10).u[0]

All arithmetic runs in the gfx950 HIP kernels behind include/tpamd.h.
"""
from . import _lib
from ._lib import ArgumentError, CapacityError, DeviceError, load_library
from .algorithms import (AdaptiveBoris, AdaptiveMultistepBoris, Boris, EnsembleDistributed,
                         EnsembleSerial, EnsembleSplitThreads, EnsembleThreads, MI355X,
                         MultistepBoris, MultistepBoris2, MultistepBoris4, MultistepBoris6)
from .constants import (BMoment_Earth, Electron, Ion, Proton, Re, Species, SpeciesDict, c, eV,
                        get_gyrofrequency, get_gyroperiod, kB, me, mi, qe, qi)
from .fields import (AffineField, ClampExtrap, DipoleField, FieldInterpolator, FillExtrap,
                     GridField, LazyTimeInterpolator, MeshField, PeriodicBC, UniformField, UxBField,
                     WrapExtrap, ZeroField, build_interpolator, prepare)
from .problem import (Boundary, EnsembleContext, InsideSphere, OutsideBox, OutsideGrid,
                      OutsideNaN, OutsideSphere, TerminateOutside, TraceProblem, remake)
from .solution import EnsembleSolution, TraceSolution
from .solve import initial_states, make_desc, plan, solve, solve_batch
from .adjoint import AdjointResult, loss_gradient, trajectory_loss
from .hybrid import AdaptiveHybrid, FullOrbitResult, TraceHybridProblem, fo_segment, full_orbit, hybrid
from .postprocess import get_fields, get_work, record_eval
from .wrapper import TraceResult, trace, trace_fieldline, trace_gc

__all__ = [n for n in dir() if not n.startswith("_")]

"""Algorithm, ensemble and backend types (src/types.jl:51-107; SciMLBase
ensemble algorithms; KernelAbstractions backends)."""
from __future__ import annotations

from dataclasses import dataclass

from ._lib import ArgumentError


class AbstractBoris:
    adaptive = False
    safety = 0.0


@dataclass(frozen=True)
class Boris(AbstractBoris):
    """Boris{Adaptive}: Boris() is the fixed-step pusher; AdaptiveBoris(safety)
    the adaptive one (src/types.jl:51-55, :85-95)."""
    adaptive: bool = False
    safety: float = 0.0


def AdaptiveBoris(safety: float = 0.1) -> Boris:
    """dt = safety * 2 pi / |q B / m| (src/types.jl:85-95)."""
    return Boris(adaptive=True, safety=float(safety))


class _MultistepMeta(type):
    def __getitem__(cls, N):
        """MultistepBoris[N](n=...) mirrors Julia's MultistepBoris{N}(; n)."""
        def ctor(n: int = 1, safety: float = 0.0, adaptive: bool = False):
            return cls(N, n=n, adaptive=adaptive, safety=safety)
        return ctor


class MultistepBoris(AbstractBoris, metaclass=_MultistepMeta):
    """Multistep/Hyper Boris of order N (2, 4, 6) with n subcycles
    (src/types.jl:65-83; src/boris/multistep_boris.jl)."""

    def __init__(self, N: int, n: int = 1, adaptive: bool = False, safety: float = 0.0):
        if N not in (2, 4, 6):
            raise ArgumentError("Multistep Boris order N must be 2, 4, or 6.")
        if int(n) < 1:
            raise ArgumentError("Multistep Boris n must be >= 1")
        self.N = int(N)
        self.n = int(n)
        self.adaptive = bool(adaptive)
        self.safety = float(safety)

    def __repr__(self):
        kind = "AdaptiveMultistepBoris" if self.adaptive else "MultistepBoris"
        return f"{kind}{{{self.N}}}(n={self.n}, safety={self.safety})"

    def __eq__(self, other):
        return (isinstance(other, MultistepBoris) and
                (self.N, self.n, self.adaptive, self.safety) ==
                (other.N, other.n, other.adaptive, other.safety))

    def __hash__(self):
        return hash((self.N, self.n, self.adaptive, self.safety))


def MultistepBoris2(n: int = 1) -> MultistepBoris:
    return MultistepBoris(2, n=n)


def MultistepBoris4(n: int = 1) -> MultistepBoris:
    return MultistepBoris(4, n=n)


def MultistepBoris6(n: int = 1) -> MultistepBoris:
    return MultistepBoris(6, n=n)


class _AdaptiveMultistepMeta(type):
    def __getitem__(cls, N):
        def ctor(n: int = 1, safety: float = 0.1):
            return MultistepBoris(N, n=n, adaptive=True, safety=safety)
        return ctor


class AdaptiveMultistepBoris(metaclass=_AdaptiveMultistepMeta):
    """AdaptiveMultistepBoris{N}(; n=1, safety=0.1) (src/types.jl:97-107):
    write AdaptiveMultistepBoris[N](n=..., safety=...)."""

    def __new__(cls, N: int, n: int = 1, safety: float = 0.1):
        return MultistepBoris(N, n=n, adaptive=True, safety=safety)


# SciMLBase ensemble algorithms: accepted for signature compatibility.  Every
# one of them runs the whole ensemble as one device launch.
class EnsembleAlgorithm:
    def __repr__(self):
        return f"{type(self).__name__}()"


class EnsembleSerial(EnsembleAlgorithm):
    pass


class EnsembleThreads(EnsembleAlgorithm):
    pass


class EnsembleDistributed(EnsembleAlgorithm):
    pass


class EnsembleSplitThreads(EnsembleAlgorithm):
    pass


@dataclass(frozen=True)
class MI355X:
    """Device backend (the positional `backend` of
    solve(prob, Boris(), backend; ...), src/boris/boris_kernel.jl:312)."""
    device: int = 0

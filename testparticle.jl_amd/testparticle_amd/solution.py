"""Solution containers: the ODESolution / EnsembleSolution contract of the
reference (src/boris/boris.jl:553-556, :161), array-backed so that millions
of trajectories do not become millions of Python objects until asked for."""
from __future__ import annotations

import numpy as np

from ._lib import RETCODE_NAMES


class TraceSolution:
    """One trajectory: `t` (n,), `u` (n, w) rows (x, y, z, vx, vy, vz, then E and B
    with save_fields, then P_par, P_fermi, P_grad, P_betatron with save_work),
    `retcode` ("Success" / "Terminated" / "MaxIters").  Calling it
    interpolates linearly between saved records, as the reference's
    LinearInterpolation(tsave, traj) does."""

    def __init__(self, t, u, retcode: str, prob=None, alg=None):
        self.t = t
        self.u = u
        self.retcode = retcode
        self.prob = prob
        self.alg = alg

    def __len__(self):
        return len(self.t)

    def __getitem__(self, idx):
        """sol[i, j] = u[j][i] (Julia's ODESolution indexing, 0-based here)."""
        if isinstance(idx, tuple) and len(idx) == 2:
            i, j = idx
            return self.u[j][i]
        return self.u[idx]

    def __call__(self, tq):
        t, u = self.t, self.u
        if len(t) == 0:
            raise ValueError("empty solution")
        tq = float(tq)
        lo, hi = (t[0], t[-1]) if t[-1] >= t[0] else (t[-1], t[0])
        if tq < lo or tq > hi:
            raise ValueError(f"t = {tq} outside the saved span [{lo}, {hi}]")
        if len(t) == 1:
            return u[0].copy()
        if t[-1] >= t[0]:
            j = int(np.searchsorted(t, tq, side="left"))
        else:
            j = int(np.searchsorted(-t, -tq, side="left"))
        j = min(max(j, 1), len(t) - 1)
        w = (tq - t[j - 1]) / (t[j] - t[j - 1])
        return u[j - 1] + w * (u[j] - u[j - 1])


class EnsembleSolution:
    """EnsembleSolution(sols, elapsed_time, converged): `u` is the list of
    TraceSolution (built on first access); `t_all` (N, cap), `u_all`
    (N, cap, w), `lengths` (N,) and `retcodes` (N,) expose the raw arrays."""

    def __init__(self, t_all, u_all, lengths, retcodes, elapsed_time, stats=None, prob=None,
                 alg=None):
        self.t_all = t_all
        self.u_all = u_all
        self.lengths = lengths
        self.retcodes = retcodes
        self.elapsed_time = elapsed_time
        self.stats = stats
        self.prob = prob
        self.alg = alg
        self.converged = True
        self._sols = None

    def __len__(self):
        return len(self.lengths)

    def _make(self, i):
        n = int(min(self.lengths[i], self.t_all.shape[1]))
        return TraceSolution(self.t_all[i, :n], self.u_all[i, :n],
                             RETCODE_NAMES.get(int(self.retcodes[i]), "Unknown"), self.prob,
                             self.alg)

    @property
    def u(self):
        if self._sols is None:
            self._sols = [self._make(i) for i in range(len(self))]
        return self._sols

    def __getitem__(self, i):
        return self.u[i] if self._sols is not None else self._make(i if i >= 0 else len(self) + i)

    def final_states(self) -> np.ndarray:
        """(N, 6) array of each trajectory's last saved record."""
        idx = np.clip(self.lengths.astype(np.int64) - 1, 0, None)
        return self.u_all[np.arange(len(self)), idx]

"""Seeded synthetic workloads of the benchmark configurations.

Every generator is deterministic in its seed and returns plain numpy arrays
plus field descriptions, so the GPU and the CPU checker read identical inputs.

  uniform_gyration  1024 protons, B = (0,0,1e-8) T, E = 0, x ~ U[-1,1]^3 m,
                    Maxwellian v (thermal speed 1e5 m/s), dt = T_g/64, 6400 steps
  dipole_belt       protons on the equator of a dipole (B0 = 3.12e-5 T at R_E),
                    L ~ U[4,7], E_k log-uniform 10 keV-10 MeV, isotropic pitch,
                    per-particle dt = local T_g/32, 10,000 steps, loss at r < R_E
  turbulent_box     periodic 256^3 box of side 1e7 m: B = B0 z + dB with a
                    k^-5/3 random-phase spectrum, dB_rms = 0.5 B0, E = -u x B,
                    u = (4e5, 0, 0); protons 1-100 keV isotropic; dt = T_g(B0)/40
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .constants import Re, mi, qi

EV = qi


@dataclass
class Workload:
    name: str
    xv0: np.ndarray            # (N, 6)
    q2m: float
    tspan: tuple
    dt: float | None = None    # fixed step
    dts: np.ndarray | None = None   # per-particle step
    nsteps: int = 0
    savestepinterval: int = 1
    save_everystep: bool = True
    extra: dict | None = None


def maxwellian_velocities(rng, n, vth):
    """Isotropic Maxwellian with thermal speed vth = sqrt(2 kT/m): per-component
    sigma = vth / sqrt(2) (ext/TestParticleVDFExt.jl:38 convention)."""
    return rng.normal(0.0, vth / math.sqrt(2.0), size=(n, 3))


def isotropic_directions(rng, n):
    cth = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2 * math.pi, n)
    sth = np.sqrt(1.0 - cth * cth)
    return np.stack([sth * np.cos(ph), sth * np.sin(ph), cth], axis=1)


def uniform_gyration(N: int = 1024, seed: int = 1, nsteps: int = 6400, K: int = 10) -> Workload:
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=(N, 3))
    v = maxwellian_velocities(rng, N, 1.0e5)
    Bz = 1.0e-8
    Tg = 2 * math.pi * mi / (qi * Bz)
    dt = Tg / 64
    return Workload("uniform_gyration", np.ascontiguousarray(np.hstack([x, v])), qi / mi,
                    (0.0, nsteps * dt), dt=dt, nsteps=nsteps, savestepinterval=K,
                    extra={"B": (0.0, 0.0, Bz)})


def dipole_field_np(x, b0=3.12e-5, R=Re):
    r2 = np.sum(x * x, axis=-1)
    r = np.sqrt(r2)
    fac = b0 * R ** 3 / (r2 * r2 * r)
    return np.stack([fac * (-3 * x[..., 0] * x[..., 2]), fac * (-3 * x[..., 1] * x[..., 2]),
                     fac * (r2 - 3 * x[..., 2] ** 2)], axis=-1)


def dipole_belt(N: int = 1 << 20, seed: int = 2, nsteps: int = 10000, b0: float = 3.12e-5,
                steps_per_gyro: int = 32) -> Workload:
    rng = np.random.default_rng(seed)
    L = rng.uniform(4.0, 7.0, N)
    phi = rng.uniform(0.0, 2 * math.pi, N)
    x = np.stack([L * Re * np.cos(phi), L * Re * np.sin(phi), np.zeros(N)], axis=1)
    Ek = np.exp(rng.uniform(math.log(1.0e4), math.log(1.0e7), N)) * EV
    speed = np.sqrt(2 * Ek / mi)
    cpa = rng.uniform(-1.0, 1.0, N)
    psi = rng.uniform(0.0, 2 * math.pi, N)
    spa = np.sqrt(1 - cpa * cpa)
    # on the equator B is along +z: v_par along z, v_perp in the x-y plane
    v = np.stack([speed * spa * np.cos(psi), speed * spa * np.sin(psi), speed * cpa], axis=1)
    Bmag = np.linalg.norm(dipole_field_np(x, b0), axis=1)
    dts = 2 * math.pi * mi / (qi * Bmag) / steps_per_gyro
    return Workload("dipole_belt", np.ascontiguousarray(np.hstack([x, v])), qi / mi,
                    (0.0, 0.0), dts=dts, nsteps=nsteps, save_everystep=False,
                    extra={"b0": b0, "R": Re})


def turbulent_field(n: int = 256, L: float = 1.0e7, B0: float = 5.0e-9, dB_rel: float = 0.5,
                    seed: int = 3, slope: float = -5.0 / 3.0, dtype=np.float32) -> np.ndarray:
    """Divergence-free random-phase turbulence on an n^3 periodic grid.

    Modes on the integer lattice 1 <= |k| <= n/2 with amplitude |k|^((slope-2)/2)
    (shell spectrum ~ k^slope), transverse to k, built by inverse FFT and
    scaled to rms |dB| = dB_rel * B0.  Returns A[c, i, j, k] (3, n, n, n)."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = np.fft.rfftfreq(n, 1.0 / n)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    kk = np.sqrt(k2)
    amp = np.zeros_like(kk)
    sel = (kk >= 1.0) & (kk <= n / 2)
    amp[sel] = kk[sel] ** ((slope - 2.0) / 2.0)
    shape = kk.shape
    a = (rng.standard_normal((3,) + shape) + 1j * rng.standard_normal((3,) + shape))
    with np.errstate(invalid="ignore", divide="ignore"):
        khat = np.stack([KX, KY, KZ]) / np.where(kk > 0, kk, 1.0)
    adotk = np.sum(a * khat, axis=0)
    a = (a - adotk[None] * khat) * amp[None]
    del adotk, khat, KX, KY, KZ, k2
    dB = np.stack([np.fft.irfftn(a[c], s=(n, n, n), axes=(0, 1, 2)) for c in range(3)])
    del a
    rms = math.sqrt(float(np.mean(np.sum(dB * dB, axis=0))))
    dB *= dB_rel * B0 / rms
    dB[2] += B0
    return dB.astype(dtype, copy=False)


def turbulent_box(N: int = 1 << 22, n: int = 256, seed: int = 3, nsteps: int = 5000,
                  L: float = 1.0e7, B0: float = 5.0e-9, dtype=np.float32,
                  steps_per_gyro: int = 40) -> tuple[Workload, np.ndarray]:
    A = turbulent_field(n, L, B0, 0.5, seed, dtype=dtype)
    rng = np.random.default_rng(seed + 1000)
    x = rng.uniform(0.0, L, size=(N, 3))
    Ek = rng.uniform(1.0e3, 1.0e5, N) * EV
    v = isotropic_directions(rng, N) * np.sqrt(2 * Ek / mi)[:, None]
    dt = 2 * math.pi * mi / (qi * B0) / steps_per_gyro
    w = Workload("turbulent_box", np.ascontiguousarray(np.hstack([x, v])), qi / mi,
                 (0.0, nsteps * dt), dt=dt, nsteps=nsteps, save_everystep=False,
                 extra={"L": L, "n": n, "u": (4.0e5, 0.0, 0.0), "B0": B0})
    return w, A


def batched_turbulence(n_inst: int = 64, seed: int = 4, total: int = 20_000_000,
                       L: float = 1.0e7, B0: float = 5.0e-9, sizes=(32, 64, 96, 128),
                       dtype=np.float32, safety: float = 0.1, max_steps: int = 2000):
    """Uneven batch of independent turbulent boxes with open boundaries.

    Instance k: grid edge drawn from `sizes`, a seeded turbulent field as in
    turbulent_box on nodes spanning [0, L]^3, and a particle count drawn
    log-uniform in [1e3, 1e6]; the counts are then scaled by one common factor
    so that they sum to `total`.  Protons 1-100 keV, isotropic, uniform in the
    box.  The span covers `max_steps` adaptive steps at |B| = B0.
    Returns (instances, tspan) with instances = [(A, axes, xv0), ...]."""
    rng = np.random.default_rng(seed)
    edges = rng.choice(np.asarray(sizes), n_inst)
    raw = np.exp(rng.uniform(math.log(1.0e3), math.log(1.0e6), n_inst))
    counts = np.maximum(1, np.round(raw * (total / raw.sum()))).astype(np.int64)
    inst = []
    for k in range(n_inst):
        n = int(edges[k])
        A = turbulent_field(n, L, B0, 0.5, seed * 1000 + k, dtype=dtype)
        axes = grid_axes(n, L, periodic=False)
        N = int(counts[k])
        x = rng.uniform(0.0, L, size=(N, 3))
        Ek = rng.uniform(1.0e3, 1.0e5, N) * EV
        v = isotropic_directions(rng, N) * np.sqrt(2 * Ek / mi)[:, None]
        inst.append((A, axes, np.ascontiguousarray(np.hstack([x, v]))))
    Tg = 2 * math.pi * mi / (qi * B0)
    return inst, (0.0, max_steps * safety * Tg)


def grid_axes(n: int, L: float, periodic: bool = True):
    h = L / n if periodic else L / (n - 1)
    ax = np.arange(n) * h
    return ax, ax, ax

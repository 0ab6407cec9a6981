"""Batched ensembles (many field instances, uneven populations and
lifetimes, one persistent launch): bit-identical FP64 results against the
oracle run instance by instance, against single solves, and across graph
replays and autotuned launch shapes."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import workloads as W
from parity import assert_bitwise, run_oracle

pytestmark = pytest.mark.gpu


def _instances(sizes, counts, seed=21, dtype=np.float64, bc=None, L=1.0e7):
    rng = np.random.default_rng(seed)
    probs = []
    for k, (n, N) in enumerate(zip(sizes, counts)):
        A = W.turbulent_field(n, L, 5e-9, 0.5, seed * 100 + k, dtype=dtype)
        ax = W.grid_axes(n, L, periodic=False)
        B = TP.GridField(A, *ax, bc=bc[k] if isinstance(bc, list) else (bc or TP.ClampExtrap()))
        x = rng.uniform(0, L, size=(N, 3))
        v = W.isotropic_directions(rng, N) * np.sqrt(2 * rng.uniform(1e3, 1e5, N) * TP.qi / TP.mi)[:, None]
        probs.append((B, np.hstack([x, v])))
    return probs


def _problems(inst, E_of, tspan):
    return [TP.TraceProblem(xv, tspan, (TP.qi / TP.mi, TP.mi, E_of(B), B, TP.ZeroField()))
            for B, xv in inst]


@pytest.mark.parametrize("hetero", [False, True])
def test_batch_adaptive_matches_oracle(hetero):
    # hetero mixes wrap modes, which forces the generic policy
    sizes, counts = [8, 12, 16, 9, 20, 10, 8], [300, 5, 700, 1, 64, 129, 0]
    bc = [TP.ClampExtrap(), TP.FillExtrap(), TP.ClampExtrap(), TP.FillExtrap(), TP.ClampExtrap(),
          TP.FillExtrap(), TP.ClampExtrap()] if hetero else None
    inst = _instances(sizes, counts, bc=bc)
    Tg = 2 * np.pi * TP.mi / (TP.qi * 5e-9)
    probs = _problems(inst, lambda B: TP.ZeroField(), (0.0, 60 * Tg))
    alg = TP.AdaptiveBoris(safety=0.1)
    kw = dict(isoutside=TP.OutsideGrid(), max_steps=400, savestepinterval=25, nout_cap=24)
    sols = TP.solve_batch(probs, alg, **kw)
    for pr, g in zip(probs, sols):
        o = run_oracle(pr, alg, trajectories=1, **kw)
        assert_bitwise(g, o)
    # exits and caps both happen
    rets = np.concatenate([s.retcodes for s in sols])
    assert (rets == 1).any()


def test_batch_fixed_multistep_matches_single_solves():
    inst = _instances([16, 8, 24], [500, 77, 260], bc=[TP.PeriodicBC()] * 3)
    Tg = 2 * np.pi * TP.mi / (TP.qi * 5e-9)
    dt = Tg / 40
    probs = _problems(inst, lambda B: TP.UxBField([4e5, 0, 0], B), (0.0, 300 * dt))
    alg = TP.MultistepBoris(4, n=2)
    sols = TP.solve_batch(probs, alg, dt=dt, savestepinterval=50)
    for pr, g in zip(probs, sols):
        s = TP.solve(pr, alg, dt=dt, savestepinterval=50)
        np.testing.assert_array_equal(g.u_all, s.u_all)
        np.testing.assert_array_equal(g.t_all, s.t_all)
        assert_bitwise(g, run_oracle(pr, alg, dt=dt, savestepinterval=50))


def test_batch_graph_replays_and_autotune_are_consistent():
    torch = pytest.importorskip("torch")
    from testparticle_amd.device import DeviceBatch

    inst = _instances([16, 32, 8, 24], [4000, 20000, 300, 9000], dtype=np.float32)
    Tg = 2 * np.pi * TP.mi / (TP.qi * 5e-9)
    probs = _problems(inst, lambda B: TP.ZeroField(), (0.0, 200 * Tg))
    alg = TP.AdaptiveBoris(safety=0.1)
    for prec in ("fp32c", "fp64"):
        b = DeviceBatch(probs, alg, precision=prec, isoutside=TP.OutsideGrid(), max_steps=2000,
                        save_everystep=False)
        st = b.run()
        ref = b.final_states().copy()
        ref_len = b.len.cpu().numpy().copy()
        assert st.particle_steps > 0
        bs, ppt = b.autotune()
        assert bs in (64, 128, 256) and ppt in (1, 2, 4, 8, 16, 32)
        for _ in range(3):
            # eager device work between replays
            torch.zeros(1 << 20, device="cuda").add_(1.0)
            torch.cuda.synchronize()
            st2 = b.run()
            assert st2.particle_steps == st.particle_steps
            np.testing.assert_array_equal(b.final_states(), ref)
            np.testing.assert_array_equal(b.len.cpu().numpy(), ref_len)
        b.close()


@pytest.mark.parametrize("layout", ["nodes", "cells"])
@pytest.mark.parametrize("bs,ppt", [(64, 1), (128, 3), (256, 16)])
def test_batch_chunking_does_not_change_results(layout, bs, ppt):
    """Block size and chunk size (64 x ppt particles) only change which wave
    runs which particle: lane refill, a wave taking the next chunk of the same
    instance while long-lived lanes finish, a chunk of another instance held
    until the wave is empty.  Every particle's records, lengths and return
    codes stay bit-identical to the default launch."""
    pytest.importorskip("torch")
    from testparticle_amd.device import DeviceBatch

    rng = np.random.default_rng(77)
    L, B0 = 1.0e7, 5e-9
    probs = []
    for k, (n, N) in enumerate(zip([16, 8, 24, 12, 8], [3000, 70, 5000, 1, 900])):
        A = W.turbulent_field(n, L, B0, 0.5, 700 + k, dtype=np.float32)
        ax = W.grid_axes(n, L, periodic=False)
        B = TP.GridField(A, *ax, bc=TP.FillExtrap(), layout=layout)
        x = rng.uniform(0, L, size=(N, 3))
        v = W.isotropic_directions(rng, N) * np.sqrt(2 * rng.uniform(1e3, 1e5, N) * TP.qi / TP.mi)[:, None]
        probs.append(TP.TraceProblem(np.hstack([x, v]), (0.0, 200 * 2 * np.pi * TP.mi / (TP.qi * B0)),
                                     (TP.qi / TP.mi, TP.mi, TP.ZeroField(), B, TP.ZeroField())))
    alg = TP.AdaptiveBoris(safety=0.1)
    for prec in ("fp64", "fp32c"):
        kw = dict(precision=prec, isoutside=TP.OutsideGrid(), max_steps=2000, save_everystep=False)
        out = []
        for extra in ({}, dict(workgroup_size=bs, particles_per_lane=ppt)):
            b = DeviceBatch(probs, alg, **kw, **extra)
            st = b.run()
            ln = b.len.cpu().numpy()
            keep = np.arange(b.cap)[None, :] < ln[:, None]
            out.append((st.particle_steps, ln, b.ret.cpu().numpy(), b.u.cpu().numpy()[keep],
                        b.t.cpu().numpy()[keep]))
            b.close()
        (s0, l0, r0, u0, t0), (s1, l1, r1, u1, t1) = out
        assert s0 == s1 and s0 > 0
        np.testing.assert_array_equal(l1, l0)
        np.testing.assert_array_equal(r1, r0)
        np.testing.assert_array_equal(u1, u0)
        np.testing.assert_array_equal(t1, t0)
        assert (r0 == 1).any() and (l0 >= 2).any()

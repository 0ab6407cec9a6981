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

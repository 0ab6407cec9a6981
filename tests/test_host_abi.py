"""CPU tests of the drop-in boundary: the C-ABI library loads, exports every
symbol include/tpamd.h declares, and its host-side logic (step count, record
schedule, argument errors) matches the reference; the Python mirror of the
reference API converts layouts and arguments correctly.  No kernel runs here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "tpamd.h")) as f:
        hdr = f.read()
    declared = set(re.findall(r"^(?:int32_t|const char\*)\s+(tpamd_\w+)\s*\(", hdr, re.M))
    assert declared == set(_lib.EXPORTS)
    L = C.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert TP.load_library().tpamd_abi_version() == _lib.ABI_VERSION


def _desc(**kw):
    prob = TP.TraceProblem([0, 0, 0, 1, 0, 0], kw.pop("tspan", (0.0, 3.0e-8)),
                           (1.0, 1.0, TP.ZeroField(), TP.ZeroField(), TP.ZeroField()))
    alg = kw.pop("alg", TP.Boris())
    return TP.make_desc(prob, alg, **kw)


def test_plan_matches_reference_schedule(orc):
    # _prepare_checks + _calculate_nout through the C ABI vs the oracle
    for nt in (0, 1, 2, 7, 10, 99, 100, 1000):
        for K in (1, 3, 10):
            for ss in (0, 1):
                for se in (0, 1):
                    for sev in (0, 1):
                        d = _desc(tspan=(0.0, nt * 0.5), dt=0.5, savestepinterval=K,
                                  save_start=ss, save_end=se, save_everystep=sev)
                        assert TP.plan(d) == (nt, orc.calculate_nout(nt, K, ss, se, sev))


def test_plan_rounding_and_negative_span():
    # nt = abs(round(Int, (t1 - t0) / dt)), ties to even (boris.jl:353)
    assert TP.plan(_desc(tspan=(0.0, 2.5), dt=1.0))[0] == 2
    assert TP.plan(_desc(tspan=(0.0, 3.5), dt=1.0))[0] == 4
    assert TP.plan(_desc(tspan=(1.0, 0.0), dt=0.1))[0] == 10
    assert TP.plan(_desc(tspan=(0, 10), dt=1.0)) == (10, 11)


def test_argument_errors(kats):
    k = kats["solver_limits"]
    with pytest.raises(TP.ArgumentError, match="maxiters"):
        TP.plan(_desc(tspan=k["tspan"], dt=k["dt"], maxiters=50))
    with pytest.raises(TP.ArgumentError, match="too small"):
        TP.plan(_desc(dt=np.finfo(np.float64).eps))
    with pytest.raises(TP.ArgumentError):
        TP.plan(_desc(dt=1e-9, savestepinterval=0))
    for N in k["bad_orders"]:
        with pytest.raises(TP.ArgumentError, match="order N must be 2, 4, or 6"):
            TP.MultistepBoris(N, n=2)
    # the C ABI applies the same check to a raw descriptor
    d = _desc(dt=1e-9, alg=TP.MultistepBoris(4, n=2))
    d.order_N = 3
    with pytest.raises(TP.ArgumentError, match="order N"):
        TP.plan(d)
    # the rational coefficients are tabulated up to n = 12; the reference form
    # (the default) takes any n up to 2^18 (its sincos argument reduction)
    d = _desc(dt=1e-9, alg=TP.MultistepBoris(4, n=13), form="rational")
    with pytest.raises(NotImplementedError):
        TP.plan(d)
    TP.plan(_desc(dt=1e-9, alg=TP.MultistepBoris(4, n=13)))
    with pytest.raises(NotImplementedError):
        TP.plan(_desc(dt=1e-9, alg=TP.MultistepBoris(4, n=(1 << 18) + 1)))
    with pytest.raises(TypeError, match="dt"):
        _desc()


def test_algorithm_constructors():
    assert TP.Boris() == TP.Boris(adaptive=False, safety=0.0)
    a = TP.AdaptiveBoris()
    assert a.adaptive and a.safety == 0.1
    m = TP.MultistepBoris[4](n=2)
    assert (m.N, m.n, m.adaptive, m.safety) == (4, 2, False, 0.0)
    am = TP.AdaptiveMultistepBoris[6]()
    assert (am.N, am.n, am.adaptive, am.safety) == (6, 1, True, 0.1)
    am2 = TP.AdaptiveMultistepBoris[2](n=1, safety=0.2)
    assert am2.safety == 0.2 and am2.adaptive
    assert TP.MultistepBoris4(n=2) == m


def test_prepare_shapes():
    q2m, m, E, B, F = TP.prepare(TP.ZeroField(), TP.UniformField([0, 0, 1e-8]))
    assert q2m == pytest.approx(TP.qi / TP.mi) and m == TP.mi
    assert isinstance(E, TP.ZeroField) and isinstance(B, TP.UniformField)
    q2m, *_ = TP.prepare(TP.UniformField([0, 0, 1.0]), species=TP.Electron)
    assert q2m == pytest.approx(TP.qe / TP.me)
    q2m, m, *_ = TP.prepare(TP.ZeroField(), TP.ZeroField(), species=TP.Ion(16, 1))
    assert m == pytest.approx(16 * TP.mi)
    q2m, m, *_ = TP.prepare(TP.ZeroField(), TP.ZeroField(), q=2.0, m=4.0)
    assert q2m == 0.5
    with pytest.raises(TypeError, match="callable"):
        TP.prepare(TP.ZeroField(), lambda x: (0, 0, 1))


def test_grid_layout_matches_reference_indexing():
    # A[c, i, j, k] (Julia's 3 x nx x ny x nz) -> device memory [k][j][i][c]
    rng = np.random.default_rng(0)
    A = rng.normal(size=(3, 4, 5, 6))
    x, y, z = np.arange(4.0), np.arange(5.0) * 2, np.arange(6.0) * 0.5
    g = TP.GridField(A, x, y, z)
    flat = g.data.reshape(-1)
    for (c, i, j, k) in [(0, 0, 0, 0), (2, 3, 4, 5), (1, 2, 0, 3), (0, 1, 3, 2)]:
        assert flat[c + 3 * (i + 4 * (j + 5 * k))] == A[c, i, j, k]
    assert np.allclose(g.spacing, [1.0, 2.0, 0.5]) and np.allclose(g.origin, 0)
    with pytest.raises(TP.ArgumentError):
        TP.GridField(A[:2], x, y, z)
    with pytest.raises(NotImplementedError):
        TP.GridField(A, np.array([0, 1, 3, 4.0]), y, z)
    # order 3, non-uniform axes, 2D / 1D and spherical grids route to a mesh
    assert isinstance(TP.build_interpolator(A, x, y, z, 3), TP.MeshField)
    assert isinstance(TP.build_interpolator(A, np.array([0, 1, 3, 4.0]), y, z), TP.MeshField)
    assert isinstance(TP.build_interpolator(A, x, y, z), TP.GridField)
    m2 = TP.build_interpolator(A[..., 0], x, y)
    assert m2.shape == (4, 5, 1) and [len(a) for a in m2.axes] == [4, 5, 1]
    m1 = TP.build_interpolator(A[:, 0, 0, :], z, dir=3)
    assert m1.shape == (1, 1, 6)
    with pytest.raises(NotImplementedError):           # constant (order 0) interpolation
        TP.build_interpolator(A, x, y, z, 0)
    with pytest.raises(TP.ArgumentError):              # test_utility.jl:236-238
        TP.build_interpolator(A, x, y, z, 2)
    with pytest.raises(TP.ArgumentError):              # interpolation.jl:250-252
        TP.build_interpolator("RectilinearGrid", A, x, y, z, 3)
    with pytest.raises(TP.ArgumentError):              # interpolation.jl:275-277
        TP.build_interpolator("StructuredGrid", A, x, y, z)


def test_initial_states_prob_func():
    base = [0, 0, 0, 1.0, 0, 0]

    def pf(prob, ctx):
        return TP.remake(prob, u0=[0, 0, 0, ctx.sim_id * 1e4, 0, 0])

    prob = TP.TraceProblem(base, (0, 1), (1, 1, TP.ZeroField(), TP.ZeroField(), TP.ZeroField()),
                           prob_func=pf)
    xv = TP.initial_states(prob, 5)
    assert np.array_equal(xv[:, 3], [1e4, 2e4, 3e4, 4e4, 5e4])
    prob2 = TP.TraceProblem(np.ones((7, 6)), (0, 1), prob.p)
    assert TP.initial_states(prob2, 1).shape == (7, 6)


def test_solution_interpolation_and_indexing():
    t = np.array([0.0, 1.0, 3.0])
    u = np.arange(18.0).reshape(3, 6)
    s = TP.TraceSolution(t, u, "Success")
    np.testing.assert_allclose(s(2.0), 0.5 * (u[1] + u[2]))
    np.testing.assert_allclose(s(0.0), u[0])
    assert s[0, -1] == u[-1][0]
    with pytest.raises(ValueError):
        s(4.0)


def test_boundary_menu_host_predicate():
    b = TP.OutsideNaN() | TP.OutsideSphere(1.0)
    assert b([0.5, 0, 0, 0, 0, 0]) is False
    assert b([1.5, 0, 0, 0, 0, 0]) is True
    assert b([0.1, 0, 0, np.nan, 0, 0]) is True
    box = TP.OutsideBox((0, 0, 0), (1, 1, 1))
    assert box([0.5, 0.5, 0.5, 0, 0, 0]) is False and box([1.1, 0.5, 0.5, 0, 0, 0]) is True
    assert TP.InsideSphere(2.0)([1, 0, 0, 0, 0, 0]) is True


def test_solve_rejects_host_callables():
    prob = TP.TraceProblem([0] * 6, (0, 1), (1.0, 1.0, TP.ZeroField(), lambda x: x, None))
    with pytest.raises(TypeError):
        TP.solve(prob, TP.Boris(), dt=0.1)
    prob = TP.TraceProblem([0] * 6, (0, 1), (1.0, 1.0, TP.ZeroField(), TP.ZeroField(), None))
    with pytest.raises(TypeError):
        TP.solve(prob, TP.Boris(), dt=0.1, isoutside=lambda u, p, t: False)
    with pytest.raises(TypeError):
        TP.solve(prob, TP.AdaptiveBoris(), dt=0.1)


def test_record_width_and_mass_in_descriptor():
    # _prepare_saved_data (src/boris/boris.jl:405-440): 6 + 6 (E, B) + 4 (work)
    z = TP.ZeroField()
    prob = TP.TraceProblem([0.0] * 6, (0.0, 1.0), (2.5, 7.0, z, z, z))
    for sf, sw, w in ((False, False, 6), (True, False, 12), (False, True, 10), (True, True, 16)):
        d = TP.make_desc(prob, TP.Boris(), dt=0.1, save_fields=sf, save_work=sw)
        assert d.record_width == w
        assert d.mass == 7.0 and d.q2m == 2.5
    # the C struct keeps its size: the new fields took the reserved words
    assert C.sizeof(TP._lib.SolveDesc) == 224
    assert TP._lib.SolveDesc.particles_per_lane.offset == 216
    assert TP.make_desc(prob, TP.AdaptiveBoris(), particles_per_lane=4).particles_per_lane == 4


def test_solve_batch_argument_checks():
    z = TP.ZeroField()
    a = TP.TraceProblem(np.zeros((2, 6)), (0.0, 1.0), (1.0, 1.0, z, z, z))
    b = TP.TraceProblem(np.zeros((3, 6)), (0.0, 2.0), (1.0, 1.0, z, z, z))
    c = TP.TraceProblem(np.zeros((3, 6)), (0.0, 1.0), (2.0, 1.0, z, z, z))
    with pytest.raises(TP.ArgumentError, match="share"):
        TP.solve_batch([a, b], TP.Boris(), dt=0.1)
    with pytest.raises(TP.ArgumentError, match="share"):
        TP.solve_batch([a, c], TP.Boris(), dt=0.1)
    assert TP.solve_batch([], TP.Boris(), dt=0.1) == []


def test_batched_workload_generator():
    from testparticle_amd import workloads as W
    inst, tspan = W.batched_turbulence(n_inst=5, total=5000, sizes=(8, 12), seed=9)
    assert sum(x.shape[0] for _, _, x in inst) == pytest.approx(5000, abs=5)
    for A, ax, xv in inst:
        n = A.shape[1]
        assert A.shape == (3, n, n, n) and n in (8, 12)
        assert ax[0][-1] == pytest.approx(1e7) and np.all((xv[:, :3] >= 0) & (xv[:, :3] < 1e7))
    assert tspan[1] > 0


def test_workload_generators_are_seeded():
    from testparticle_amd import workloads as W
    a = W.uniform_gyration(N=16)
    b = W.uniform_gyration(N=16)
    assert np.array_equal(a.xv0, b.xv0)
    A = W.turbulent_field(16, seed=3, dtype=np.float64)
    assert A.shape == (3, 16, 16, 16)
    dB = A.copy()
    dB[2] -= 5e-9
    rms = np.sqrt(np.mean(np.sum(dB * dB, axis=0)))
    assert rms == pytest.approx(0.5 * 5e-9, rel=1e-10)
    # divergence-free: spectral divergence vanishes to rounding
    k = np.fft.fftfreq(16, 1 / 16)
    Fx, Fy, Fz = (np.fft.fftn(dB[c]) for c in range(3))
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    div = KX * Fx + KY * Fy + KZ * Fz
    assert np.max(np.abs(div)) < 1e-10 * np.max(np.abs(Fx))
    w = W.dipole_belt(N=64)
    r = np.linalg.norm(w.xv0[:, :3], axis=1) / TP.Re
    assert np.all((r >= 4) & (r <= 7)) and np.all(w.xv0[:, 2] == 0)


@pytest.mark.skipif(not any(os.access(os.path.join(d, "gcc"), os.X_OK)
                            for d in os.environ.get("PATH", "").split(os.pathsep)),
                    reason="no C compiler")
def test_struct_layouts_match_the_c_header(tmp_path):
    # the ctypes mirrors against the header, compiled by the C compiler
    import subprocess

    structs = {"tpamd_solve_desc": _lib.SolveDesc, "tpamd_stats": _lib.Stats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "tpamd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for name, _ in py._fields_:
            lines.append(f'printf("{cname} {name} %zu\\n", offsetof({cname}, {name}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {(c, n): int(v) for c, n, v in (l.split() for l in out if l)}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for name, _ in py._fields_:
            assert got[(cname, name)] == getattr(py, name).offset, (cname, name)

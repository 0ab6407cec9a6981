"""The trace() caller's host logic (bindings/python/src/testparticle/
__init__.py:173-243,245-439): result columns, time-dependence detection,
tabulation of callable fields, solver-name resolution and the errors a GPU
path must raise for fields it cannot evaluate."""
import numpy as np
import pytest

import testparticle_amd as TP
from testparticle_amd import wrapper as Wr


def test_trace_result_columns():
    st = np.arange(5 * 16, dtype=float).reshape(5, 16)
    r = TP.TraceResult(np.arange(5.0), st, True, True)
    np.testing.assert_array_equal(r.x, st[:, 0:3])
    np.testing.assert_array_equal(r.v, st[:, 3:6])
    np.testing.assert_array_equal(r.E, st[:, 6:9])
    np.testing.assert_array_equal(r.B, st[:, 9:12])
    np.testing.assert_array_equal(r.work, st[:, 12:16])
    plain = TP.TraceResult(np.arange(5.0), st[:, :6])
    for name in ("E", "B", "work"):
        with pytest.raises(AttributeError):
            getattr(plain, name)
    w = TP.TraceResult(np.arange(5.0), st[:, :10], save_work=True)
    np.testing.assert_array_equal(w.work, st[:, 6:10])


def test_time_dependence_rule():
    # __init__.py:186-200: two positional parameters or *args
    assert not Wr._time_dependent(lambda x: x)
    assert Wr._time_dependent(lambda x, t: x)

    def with_default(x, t=0.0):
        return x

    def var(*a):
        return a[0]

    assert Wr._time_dependent(with_default) and Wr._time_dependent(var)


def test_tabulation_and_field_routing():
    g = (np.linspace(0, 2, 3), np.linspace(0, 1, 2), np.linspace(-1, 1, 3))
    A = Wr._tabulate(lambda x: [x[0], 2 * x[1], x[2] ** 2], g)
    assert A.shape == (3, 3, 2, 3)
    assert A[0, 2, 0, 0] == 2.0 and A[1, 0, 1, 0] == 2.0 and A[2, 0, 0, 0] == 1.0
    f = Wr._as_device_field(lambda x: [0, 0, 1e-8], g, None, TP.ClampExtrap())
    assert isinstance(f, TP.GridField) and f.extrap == TP._lib.EXTRAP_CLAMP
    td = Wr._as_device_field(lambda x, t: [0, 0, t], g, [0.0, 1.0, 3.0], TP.ClampExtrap())
    assert isinstance(td, TP.LazyTimeInterpolator)
    snaps = [s for _, s in td.snapshots()]
    assert [s.data[0, 0, 0, 2] for s in snaps] == [0.0, 1.0, 3.0]
    assert isinstance(Wr._as_device_field([0, 0, 1.0], None, None, None), TP.UniformField)
    with pytest.raises(TypeError):                     # a callable needs a grid on the GPU
        Wr._as_device_field(lambda x: [0, 0, 1.0], None, None, None)
    with pytest.raises(TypeError):                     # f(x, t) needs snapshot times
        Wr._as_device_field(lambda x, t: [0, 0, 1.0], g, None, TP.ClampExtrap())


def test_solver_names_and_stubs():
    assert isinstance(Wr._as_alg(None), TP.Boris) and not Wr._as_alg(None).adaptive
    assert Wr._as_alg("MultistepBoris4").N == 4
    assert Wr._as_alg("AdaptiveBoris").adaptive
    with pytest.raises(TP.ArgumentError):
        Wr._as_alg("RK4")
    for f in (TP.trace_gc, TP.trace_fieldline):
        with pytest.raises(NotImplementedError):
            f()
    with pytest.raises(ValueError):
        TP.trace([0, 0], [1, 0, 0], (0, 1), 0.1, [0, 0, 1.0])

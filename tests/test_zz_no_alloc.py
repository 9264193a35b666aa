"""fl_run_steps_static: the multi-step run without any allocation (scratch in a fixed device
array), checked through the C ABI against fl_run_steps on the same inputs: bit for bit, on
two calls in a row, the second on a smaller grid, and with the source inside and outside
the update domain; FP32 and the FP16-recompute form (whose 8-row tiles run the
both-fields-per-lane kernel)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = (12.5, 12.5, 12.5)


def _case(fl, n3, absorb, seed):
    m = fl.model_gradient(n3, absorb, seed=seed, blob_smin=4.0, blob_smax=10.0, smooth_smin=6.0,
                          smooth_smax=16.0)
    return m, m[0].shape, fl.cfl_dt(m, 12.5, 0.4)


def _run(fl, m, shape, dt, isrc, st, rec, static, prec):
    L = fl.lib()
    d = fl.Driver(variant="fused", precision=prec, tiles=(8, 0))
    d.initialize(shape, D, dt, m)
    d.record_traces(rec, 3, st // 3)
    src = fl.sources(dt, st)
    before = L.fl_device_bytes()
    if static:
        rc = L.fl_run_steps_static(0, st, isrc, src)
        assert rc == 0, fl.last_error()
        assert L.fl_device_bytes() == before  # nothing allocated by the call
    else:
        d.run(st, isrc, graph_chunk=10, src=src)
    out = (d.field("pc"), d.field("qc"), d.field("pp"), d.traces())
    d.finalize()
    return out


@pytest.mark.parametrize("prec", ["fp32", "fp16r"])
def test_static_run_equals_run_steps_twice(fl, prec):
    m1, s1, dt1 = _case(fl, (44, 40, 36), 6, 21)
    m2, s2, dt2 = _case(fl, (20, 18, 16), 4, 22)
    assert np.prod(s2) < np.prod(s1)
    calls = [(m1, s1, dt1, fl.index(s1, 25, 20, 14), 60, 14),   # source inside the update domain
             (m2, s2, dt2, fl.index(s2, 14, 13, 10), 24, 9),    # smaller grid, fewer steps
             (m2, s2, dt2, fl.index(s2, 2, 13, 10), 18, 9)]     # source in the halo: never stored
    for m, s, dt, isrc, st, rec in calls:
        print("static run: %s shape %s steps %d source %d" % (prec, s, st, isrc), flush=True)
        a = _run(fl, m, s, dt, isrc, st, rec, static=False, prec=prec)
        b = _run(fl, m, s, dt, isrc, st, rec, static=True, prec=prec)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert float(np.abs(a[0]).max()) > 0


def test_static_run_limits(fl):
    m, s, dt = _case(fl, (20, 18, 16), 4, 23)
    L = fl.lib()
    d = fl.Driver(variant="fused")
    d.initialize(s, D, dt, m)
    src = np.zeros(65537, np.float32)
    assert L.fl_run_steps_static(0, 65537, fl.center_index(s), src) == -1  # FL_ERR_ARG
    assert "scratch limit" in fl.last_error()
    L.fl_clear_error()
    d.set_variant("twopass")
    assert L.fl_run_steps_static(0, 4, fl.center_index(s), src) == -1
    L.fl_clear_error()
    d.set_variant("fused")
    d.finalize()

"""The device FP64 accuracy reference (fl_ref64_*) against the FP64 oracle.

Both evaluate the reference's der1der1 time loop in double with exact rational weights, in
the same operation order and without FMA contraction: the results are bitwise equal.  This
pins the device reference, which measures the FP32 and 16-bit forms at sizes the oracle
cannot run in test time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = (12.5, 12.5, 12.5)


@pytest.mark.parametrize("n3,absorb,st", [((21, 17, 13), 4, 12), ((40, 36, 30), 8, 40)])
def test_ref64_bitwise_equals_fp64_oracle(fl, oracle, n3, absorb, st):
    m = fl.model_gradient(n3, absorb, seed=77, blob_smin=4.0, blob_smax=9.0, smooth_smin=5.0,
                          smooth_smax=12.0)
    shape = m[0].shape
    dt = fl.cfl_dt(m, 12.5, 0.4)
    rec = 4 + absorb + 2
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, rec)
    ref = oracle.run("der1der1", shape, 4, D, dt, st, isrc, oracle.precompute(m, "f64"), fp="f64",
                     rec_iz=rec, rec_every=3)
    r = fl.Ref64(shape, D, dt, m)
    try:
        r.run(st, isrc, rec_iz=rec, rec_every=3, max_traces=st // 3)
        got = {k: r.field(k) for k in ("pp", "pc", "qp", "qc")}
        tr = r.traces()
    finally:
        r.destroy()
    assert np.abs(ref["pc"]).max() > 0
    for k in ("pp", "pc", "qp", "qc"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(tr, ref["traces"])


def test_ref64_continues_across_runs(fl):
    m = fl.model_gradient((24, 20, 18), 4, seed=3)
    shape = m[0].shape
    dt = fl.cfl_dt(m, 12.5, 0.4)
    isrc = fl.center_index(shape)
    src = fl.sources(dt, 20)
    r = fl.Ref64(shape, D, dt, m)
    try:
        r.run(8, isrc, src=src[:8])
        r.run(12, isrc, it0=8, src=src[8:])
        a = r.field("pc")
    finally:
        r.destroy()
    r = fl.Ref64(shape, D, dt, m)
    try:
        r.run(20, isrc, src=src)
        b = r.field("pc")
    finally:
        r.destroy()
    assert np.array_equal(a, b)

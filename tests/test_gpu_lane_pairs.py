"""The both-fields-per-lane step (tti_step2: the 16-bit recompute forms on 8-row tiles)
against the one-field-per-lane step (tti_step): the same expressions in the same order, so pc,
qc and the receiver traces must be the same bits, for chunk depths with and without whole
groups of ten planes, stream-hint settings, grid spacings, call-by-call and captured runs,
and batched shots (one workgroup per item and persistent).  Other tile heights ignore the
setting."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = (12.5, 12.5, 12.5)


@pytest.fixture
def pairs_default(fl):
    yield
    fl.lib().fl_set_lane_pairs(1)  # the default


def _model(fl, n3=(70, 45, 37), absorb=8, seed=21):
    m = fl.model_gradient(n3, absorb, seed=seed, blob_smin=4.0, blob_smax=12.0, smooth_smin=6.0,
                          smooth_smax=20.0)
    return m, fl.cfl_dt(m, 12.5, 0.4)


def _single(fl, prec, tiles, hints, pair, d3, m, dt, chunk, steps=40):
    shape = m[0].shape
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, 14)
    d = fl.Driver(variant="fused", precision=prec, tiles=tiles)
    try:
        d.set_lane_pairs(pair)
        assert d.lane_pairs() == pair
        d.set_stream_hints(hints)
        d.initialize(shape, d3, dt, m)
        d.record_traces(12, 4, steps // 4)
        if chunk is None:
            d.step_loop(steps, isrc)
        else:
            d.run(steps, isrc, graph_chunk=chunk)
        return d.field("pc"), d.field("qc"), d.traces()
    finally:
        d.finalize()


@pytest.mark.parametrize("prec", ["fp16r", "q16r"])
@pytest.mark.parametrize("tiles,hints,chunk", [((8, 0), True, 10), ((8, 7), True, 0), ((8, 11), False, 10),
                                               ((8, 29), False, 0), ((8, 0), False, None), ((12, 0), True, 10)])
def test_pairs_equal_split_lanes(fl, pairs_default, prec, tiles, hints, chunk):
    m, dt = _model(fl)
    print("lane pairs: %s tiles %s hints %s chunk %s" % (prec, tiles, hints, chunk), flush=True)
    ref = _single(fl, prec, tiles, hints, False, D, m, dt, chunk)
    got = _single(fl, prec, tiles, hints, True, D, m, dt, chunk)
    for x, y in zip(got, ref):
        assert np.abs(y).max() > 0
        assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", ["fp16r", "q16r"])
def test_pairs_equal_split_lanes_anisotropic(fl, pairs_default, prec):
    m, dt = _model(fl, n3=(41, 66, 30), seed=5)
    d3 = (12.5, 10.0, 15.0)
    dt = dt * 0.8
    print("lane pairs anisotropic: %s" % prec, flush=True)
    ref = _single(fl, prec, (8, 9), True, False, d3, m, dt, 10)
    got = _single(fl, prec, (8, 9), True, True, d3, m, dt, 10)
    for x, y in zip(got, ref):
        assert np.abs(y).max() > 0
        assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", ["fp16r", "q16r"])
@pytest.mark.parametrize("persist", [0, 1])
def test_batch_pairs_equal_split_lanes(fl, pairs_default, prec, persist):
    from fletcher_petrobras_amd import shots as S
    absorb = 6
    parent = fl.model_gradient((96, 88, 40), absorb, seed=512, blob_smin=5.0, blob_smax=12.0,
                               smooth_smin=8.0, smooth_smax=24.0)
    interior = S.parent_interior(parent, absorb)
    shots = S.make_shots(interior, 3, absorb, 30, 60, src_depth=4 + absorb + 5, rec_depth=4 + absorb + 5,
                         seed=3)
    dt = min(fl.cfl_dt(s["model"], 12.5, 0.4) for s in shots)
    st, every = 30, 3
    out = []
    for pair in (False, True):
        print("batch lane pairs: %s persistent %d pairs %s" % (prec, persist, pair), flush=True)
        b = fl.Batch(shots, D, dt, precision=prec, tiles=(8, 0))
        try:
            b.set_lane_pairs(pair)
            b.set_persistent(persist)
            b.run(st, graph_chunk=10, rec_every=every, max_traces=st // every)
            out.append([(b.field(k, "pc"), b.field(k, "qc"), b.traces(k)) for k in range(len(shots))])
        finally:
            b.set_persistent(0)
            b.destroy()
    for k in range(len(shots)):
        for x, y in zip(out[1][k], out[0][k]):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y), k


def test_lane_pairs_knob(fl, pairs_default):
    L = fl.lib()
    assert L.fl_set_lane_pairs(2) == -1
    L.fl_clear_error()
    assert L.fl_get_lane_pairs() == 1  # the default
    assert L.fl_set_lane_pairs(0) == 0 and L.fl_get_lane_pairs() == 0
    assert L.fl_set_lane_pairs(1) == 0 and L.fl_get_lane_pairs() == 1

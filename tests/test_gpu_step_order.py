"""Odd steps take the tiles in reverse order and march each z chunk from its top plane down;
even steps go forwards and upwards (csrc/fl_kernels.h step_order, step_down).  One step from
the same levels must give the same bits at an even and at an odd step index: the z second
derivative is a sum of symmetric pairs, and the z first derivatives come out exactly negated
and are scaled back exactly.  The receiver plane each step writes is compared too.  Cases:
FP32 with isotropic and anisotropic spacing, the two-pass variant, FP16 recompute with both
fields per lane and with one, q16r and BF16, ragged and whole z chunks, per-step calls and
captured runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LEVELS = ("pc", "qc", "pp", "qp")
ISO = (12.5, 12.5, 12.5)
ANISO = (12.5, 10.0, 15.0)


def _driver(fl, variant, prec, tiles, d3):
    # interior 45 x 38 x 53, absorbing zone 6: 65 updated planes (chunks of 14: 14 x 4 + 9)
    m = fl.model_gradient((45, 38, 53), 6, seed=77, blob_smin=4.0, blob_smax=12.0, smooth_smin=6.0,
                          smooth_smax=20.0)
    dt = fl.cfl_dt(m, min(d3), 0.4)
    shape = m[0].shape
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, shape[0] // 2)
    d = fl.Driver(variant=variant, precision=prec, tiles=tiles)
    d.initialize(shape, d3, dt, m)
    d.step_loop(9, isrc)  # a wavefield that reaches every chunk
    return d, isrc, shape[0] // 2 + 2


def _one_step(d, state, it, graph, isrc):
    for k, v in state.items():
        d.set_field(k, v)
    if graph:
        d.run(1, isrc, it0=it - 1, graph_chunk=1, src=np.zeros(1, np.float32))
    else:
        d.propagate(it)
    return {k: d.field(k) for k in LEVELS}


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("variant,prec,tiles,d3", [
    ("fused", "fp32", (12, 14), ISO),
    ("fused", "fp32", (12, 0), ISO),
    ("fused", "fp32", (8, 14), ANISO),
    ("twopass", "fp32", (8, 11), ISO),
    ("fused", "fp16r", (8, 14), ISO),
    ("fused", "fp16r", (12, 14), ISO),
    ("fused", "q16r", (8, 9), ANISO),
    ("fused", "bf16", (12, 13), ANISO),
])
def test_odd_step_equals_even_step(fl, variant, prec, tiles, d3, graph):
    print("step order: %s %s tiles %s spacing %s graph %s" % (variant, prec, tiles, d3, graph), flush=True)
    d, isrc, rec = _driver(fl, variant, prec, tiles, d3)
    try:
        state = {k: d.field(k) for k in LEVELS}
        assert np.abs(state["pc"]).max() > 0
        d.record_traces(rec, 1, 12)  # step it writes slot it - 1
        even = _one_step(d, state, 10, graph, isrc)
        odd = _one_step(d, state, 11, graph, isrc)
        even2 = _one_step(d, state, 12, graph, isrc)
        tr = d.traces()
    finally:
        d.finalize()
    assert np.abs(tr[9]).max() > 0
    assert np.array_equal(tr[10].view(np.uint32), tr[9].view(np.uint32))
    assert np.array_equal(tr[11], tr[9])
    changed = False
    for k in LEVELS:
        assert np.array_equal(even[k], even2[k]), k
        assert np.array_equal(odd[k].view(np.uint32), even[k].view(np.uint32)), k
        changed |= not np.array_equal(even[k], state[k])
    assert changed

"""Model setup on the device: coefficient records computed on the GPU in FP64, and grids
initialised from a model streamed in z slabs (no whole host arrays).

* device records against the host records of every precision: the velocity-factor and
  recompute forms bit for bit (no trigonometry); the stored angular coefficients bit for bit
  except where the device's double sine or cosine rounds differently from the host library's,
  then within one unit in the last place of the stored format;
* a layered model streamed in slabs equals the whole model initialised the same way: same
  records and same fields after a run, bit for bit; and the run equals the oracle (FP32) at
  the 1e-5 bar.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

D = (12.5, 12.5, 12.5)
# record slots that hold the six angular coefficients (the FP32 record puts v2px in slot 4)
TRIG = {"fp32": (np.uint32, 10, (0, 1, 2, 3, 5, 6)), "bf16": (np.uint16, 10, range(6)),
        "fp16": (np.uint16, 10, range(6))}


def _records(fl, prec, shape, dt, m, device):
    d = fl.Driver(variant="fused", precision=prec)
    d.set_device_precompute(device)
    try:
        d.initialize(shape, D, dt, m)
        return d.records()
    finally:
        d.set_device_precompute(False)
        d.finalize()


@pytest.mark.parametrize("prec", ["fp32", "fp32r", "bf16", "fp16", "fp16r", "q16r"])
def test_device_records_match_host(fl, prec):
    m = fl.model_gradient((40, 36, 30), 6, seed=9, blob_smin=4.0, blob_smax=10.0, smooth_smin=6.0,
                          smooth_smax=16.0)
    shape = m[0].shape
    dt = fl.cfl_dt(m, 12.5, 0.4)
    host = _records(fl, prec, shape, dt, m, False)
    dev = _records(fl, prec, shape, dt, m, True)
    assert host.size == dev.size
    if prec not in TRIG:
        assert np.array_equal(host, dev)
        return
    dtype, nrec, angular_slots = TRIG[prec]
    h = host.view(dtype).reshape(-1, nrec)
    g = dev.view(dtype).reshape(-1, nrec)
    diff = h != g
    print("%s: %d of %d stored values differ" % (prec, int(diff.sum()), diff.size))
    # only angular coefficients may differ
    vel_slots = [s for s in range(nrec) if s not in angular_slots]
    assert not diff[:, vel_slots].any()
    assert diff.sum() <= 1e-4 * diff.size
    ia = h[diff].astype(np.int64)
    ib = g[diff].astype(np.int64)
    assert np.all(np.abs(ia - ib) <= 1)  # one unit in the last place (same sign, same exponent range)


@pytest.mark.parametrize("prec", ["fp32", "q16r"])
def test_streamed_layered_model_initialisation(fl, oracle, prec):
    n3, absorb = (32, 40, 40), 32
    kw = dict(smooth_smin=6.0, smooth_smax=20.0)
    whole = fl.model_layered(n3, absorb, seed=1024, **kw)
    shape = whole[0].shape
    s = fl.ModelStream("layered", n3, absorb, 1024, **kw)
    try:
        dt = s.cfl_dt(10.0, 0.4)
        assert dt == fl.cfl_dt(whole, 10.0, 0.4)
        d3 = (10.0,) * 3
        rec = 4 + absorb + 8
        isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, rec)
        st = 60
        out = []
        for streamed in (True, False):
            d = fl.Driver(variant="fused", precision=prec)
            d.set_device_precompute(True)
            try:
                if streamed:
                    d.initialize_streamed(s, d3, dt)
                else:
                    d.initialize(shape, d3, dt, whole)
                recs = d.records()
                d.record_traces(rec, 4, st // 4)
                d.run(st, isrc, graph_chunk=20)
                out.append((recs, d.field("pc"), d.field("qc"), d.traces()))
            finally:
                d.set_device_precompute(False)
                d.finalize()
        for x, y in zip(out[0], out[1]):
            assert np.array_equal(x, y)
        if prec == "fp32":
            r = oracle.run("der1der1", shape, 4, d3, dt, st, isrc, oracle.precompute(whole), rec_iz=rec, rec_every=4)
            for k, v in (("pc", out[0][1]), ("qc", out[0][2]), ("traces", out[0][3])):
                e = rel_l2(v, r[k])
                print("streamed layered %s: relL2 vs oracle %.3e" % (k, e))
                assert e <= 1e-5
    finally:
        s.close()

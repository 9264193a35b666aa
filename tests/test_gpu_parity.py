"""GPU parity of the HIP back-end (through the C ABI) against the reference and the oracle.

Bar: relative L2 <= 1e-5 on pc, qc and receiver traces against the reference C/OpenMP
path (golden fixtures written by the reference program; the oracle, bitwise equal to it,
elsewhere); bitwise equality where the arithmetic is the same by construction (tile
shapes, graph vs call-by-call runs, the ISO reduction p == q, batched vs single shot).
"""
import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

D = (12.5, 12.5, 12.5)
TOL = 1e-5


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


def ref_model(fl, form, nx, absorb):
    shape = fl.grid_shape(nx, nx, nx, absorb)
    m = fl.model_homogeneous(form, shape)
    fl.random_boundary(m, (nx, nx, nx), absorb)
    return shape, m


def hetero(fl, n3=(70, 45, 37), absorb=8, seed=11):
    m = fl.model_gradient(n3, absorb, seed=seed, blob_smin=4.0, blob_smax=12.0, smooth_smin=6.0,
                          smooth_smax=20.0)
    dt = fl.cfl_dt(m, 12.5, 0.4)
    return m, dt


@pytest.mark.parametrize("variant", ["fused", "direct", "twopass"])
@pytest.mark.parametrize("form", ["ISO", "VTI", "TTI"])
@pytest.mark.parametrize("steps", [10, 30])
def test_reference_fixture_parity(fl, drv, form, variant, steps):
    ref_variant = "original" if variant == "direct" else "der1der1"
    g = golden("ref_%s_%s_n8_a4_s%d.npz" % (form, ref_variant, steps))
    shape, m = ref_model(fl, form, 8, 4)
    d = drv(variant=variant)
    d.initialize(shape, D, 0.001, m)
    d.step_loop(steps, int(g["isource"]))
    pc = d.update_pointers()
    assert rel_l2(pc, g["pc"]) <= TOL
    assert rel_l2(d.field("qc"), g["qc"]) <= TOL
    assert rel_l2(d.field("pp"), g["pp"]) <= TOL


@pytest.mark.parametrize("form", ["ISO", "VTI", "TTI"])
@pytest.mark.parametrize("steps", [10, 30])
def test_fp32_recompute_fixture_parity(fl, drv, form, steps):
    """FP32 recompute (velocity factors of precomp.h bit for bit, the six angular
    coefficients rebuilt in FP32 from theta, phi per cell) against the reference fixtures."""
    g = golden("ref_%s_der1der1_n8_a4_s%d.npz" % (form, steps))
    shape, m = ref_model(fl, form, 8, 4)
    d = drv(variant="fused", precision="fp32r")
    d.initialize(shape, D, 0.001, m)
    d.step_loop(steps, int(g["isource"]))
    assert rel_l2(d.update_pointers(), g["pc"]) <= TOL
    assert rel_l2(d.field("qc"), g["qc"]) <= TOL
    assert rel_l2(d.field("pp"), g["pp"]) <= TOL


def test_reference_fixture_parity_absorb16(fl, drv):
    g = golden("ref_TTI_original_n24_a16_s10.npz")
    shape, m = ref_model(fl, "TTI", 24, 16)
    for variant in ("fused", "direct", "twopass"):
        d = drv(variant=variant)
        d.initialize(shape, D, 0.001, m)
        d.step_loop(10, int(g["isource"]))
        assert rel_l2(d.field("pc"), g["pc"]) <= TOL, variant
        assert rel_l2(d.field("qc"), g["qc"]) <= TOL, variant
        d.finalize()


@pytest.mark.parametrize("variant", ["fused", "direct", "twopass"])
def test_oracle_parity_heterogeneous_ragged(fl, oracle, drv, variant):
    """Ragged grid (no extent a multiple of the tile), heterogeneous TTI model, off-centre
    source, 160 steps, receiver traces every 4 steps."""
    m, dt = hetero(fl)
    shape = m[0].shape
    isrc = fl.index(shape, shape[2] // 3, shape[1] // 2, 4 + 8 + 5)
    rec = 4 + 8 + 5
    st = 160
    d = drv(variant=variant)
    d.initialize(shape, D, dt, m)
    d.record_traces(rec, 4, st // 4)
    d.run(st, isrc, graph_chunk=0)
    ov = "original" if variant == "direct" else "der1der1"
    r = oracle.run(ov, shape, 4, D, dt, st, isrc, oracle.precompute(m), rec_iz=rec, rec_every=4)
    assert rel_l2(d.field("pc"), r["pc"]) <= TOL
    assert rel_l2(d.field("qc"), r["qc"]) <= TOL
    assert rel_l2(d.traces(), r["traces"]) <= TOL
    # and the other reference form (the two differ only by rounding)
    r2 = oracle.run("der1der1" if ov == "original" else "original", shape, 4, D, dt, st, isrc,
                    oracle.precompute(m))
    assert rel_l2(d.field("pc"), r2["pc"]) <= TOL


def test_direct_bitwise_equals_original_oracle(fl, oracle, drv):
    """The direct variant evaluates the reference's DerCross, Der2, combination and leapfrog
    in their association order, staged through LDS: on a grid with partial tiles in x and y,
    several z chunks and anisotropic spacings, pc, qc and the traces equal the FP32 oracle of
    original/ (itself bitwise equal to the reference build) bit for bit."""
    d3 = (12.5, 10.0, 15.0)
    m = fl.model_gradient((150, 37, 41), 6, seed=5, blob_smin=4.0, blob_smax=10.0, smooth_smin=6.0,
                          smooth_smax=16.0)
    shape = m[0].shape
    dt = fl.cfl_dt(m, 10.0, 0.4)
    rec, st = 4 + 6 + 3, 40
    isrc = fl.index(shape, 100, shape[1] // 2, rec)
    d = drv(variant="direct")
    d.initialize(shape, d3, dt, m)
    d.record_traces(rec, 4, st // 4)
    d.run(st, isrc, graph_chunk=0)
    r = oracle.run("original", shape, 4, d3, dt, st, isrc, oracle.precompute(m), rec_iz=rec, rec_every=4)
    assert np.abs(r["pc"]).max() > 0
    for k in ("pc", "qc"):
        assert np.array_equal(d.field(k), r[k]), (k, rel_l2(d.field(k), r[k]))
    assert np.array_equal(d.traces(), r["traces"])


def test_fp32_recompute_oracle_parity_heterogeneous_ragged(fl, oracle, drv):
    m, dt = hetero(fl)
    shape = m[0].shape
    isrc = fl.index(shape, shape[2] // 3, shape[1] // 2, 4 + 8 + 5)
    rec, st = 4 + 8 + 5, 160
    d = drv(variant="fused", precision="fp32r")
    d.initialize(shape, D, dt, m)
    d.record_traces(rec, 4, st // 4)
    d.run(st, isrc, graph_chunk=20)
    r = oracle.run("der1der1", shape, 4, D, dt, st, isrc, oracle.precompute(m), rec_iz=rec, rec_every=4)
    e = [rel_l2(d.field("pc"), r["pc"]), rel_l2(d.field("qc"), r["qc"]), rel_l2(d.traces(), r["traces"])]
    print("fp32r vs FP32 oracle relL2 pc %.3e qc %.3e traces %.3e" % tuple(e))
    assert max(e) <= TOL


@pytest.mark.parametrize("prec", ["fp32", "fp16r"])
def test_tiles_from_x_bord(fl, oracle, drv, prec):
    """sx = 72: the update domain [4, 68) takes two 32-cell tile columns from x = 4 against
    three from x = 0, so the first tile column starts at x = 4.  FP32 against the oracle,
    and both forms bitwise equal to a run whose tile rows differ."""
    m, dt = hetero(fl, (48, 41, 30), 8)
    shape = m[0].shape
    assert shape[2] == 72
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, 4 + 8 + 3)
    st = 60
    outs = []
    for tiles in ((8, 0), (12, 9), (16, 11)):
        d = drv(variant="fused", precision=prec, tiles=tiles)
        d.initialize(shape, D, dt, m)
        d.run(st, isrc, graph_chunk=0)
        outs.append(d.field("pc"))
        d.finalize()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    if prec == "fp32":
        r = oracle.run("der1der1", shape, 4, D, dt, st, isrc, oracle.precompute(m))
        assert rel_l2(outs[0], r["pc"]) <= TOL


@pytest.mark.parametrize("variant,prec", [("fused", "fp32"), ("direct", "fp32"), ("twopass", "fp32"),
                                          ("fused", "fp32r")])
def test_anisotropic_spacing_oracle_parity(fl, oracle, drv, variant, prec):
    """dx != dy != dz: the kernels that scale every derivative by its own spacing (the
    isotropic grids take the ones that apply 1/h^2 once)."""
    d3 = (12.5, 10.0, 15.0)
    m = fl.model_gradient((44, 38, 30), 6, seed=21, blob_smin=4.0, blob_smax=10.0, smooth_smin=6.0,
                          smooth_smax=16.0)
    shape = m[0].shape
    dt = fl.cfl_dt(m, 10.0, 0.4)
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 3, 4 + 6 + 3)
    rec, st = 4 + 6 + 3, 80
    d = drv(variant=variant, precision=prec)
    d.initialize(shape, d3, dt, m)
    d.record_traces(rec, 4, st // 4)
    d.run(st, isrc, graph_chunk=0)
    ov = "original" if variant == "direct" else "der1der1"
    r = oracle.run(ov, shape, 4, d3, dt, st, isrc, oracle.precompute(m), rec_iz=rec, rec_every=4)
    e = [rel_l2(d.field("pc"), r["pc"]), rel_l2(d.field("qc"), r["qc"]), rel_l2(d.traces(), r["traces"])]
    print("%s/%s anisotropic: relL2 pc %.3e qc %.3e traces %.3e" % ((variant, prec) + tuple(e)))
    assert max(e) <= TOL


def test_fp64_oracle_error_reported(fl, oracle, drv):
    m, dt = hetero(fl, (48, 40, 36), 6)
    shape = m[0].shape
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, 4 + 6 + 4)
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    d.run(200, isrc, graph_chunk=0)
    r = oracle.run("der1der1", shape, 4, D, dt, 200, isrc, oracle.precompute(m, "f64"), fp="f64")
    err = rel_l2(d.field("pc"), r["pc"])
    print("fused FP32 vs FP64 oracle relL2(pc) = %.3e" % err)
    assert err <= TOL




@pytest.mark.parametrize("tiles", [(4, 16), (8, 0), (12, 0), (12, 7), (16, 24), (8, 5), (4, 200)])
def test_tile_shapes_are_bitwise_identical(fl, drv, tiles):
    m, dt = hetero(fl, (60, 33, 41), 6)
    shape = m[0].shape
    isrc = fl.index(shape, 20, 17, 15)
    out = []
    for t in [(8, 0), tiles]:
        d = drv(variant="fused", tiles=t)
        d.initialize(shape, D, dt, m)
        d.run(40, isrc, graph_chunk=0)
        out.append((d.field("pc"), d.field("qc")))
        d.finalize()
    for k in (0, 1):
        assert np.array_equal(out[0][k], out[1][k])


@pytest.mark.parametrize("variant,prec,tiles", [("fused", "fp32", (12, 0)), ("fused", "q16r", (8, 11)),
                                                ("twopass", "fp32", (16, 0))])
def test_strip_orders_are_bitwise_identical(fl, drv, variant, prec, tiles):
    """The workgroup order (strips of sw tile columns) moves tiles between workgroups and
    XCDs only: every width gives the same bits, the plain row-major order (a width at or
    above the tile columns) included, and the default picks a width within the columns."""
    m, dt = hetero(fl, (150, 61, 41), 6)
    shape = m[0].shape
    isrc = fl.index(shape, 70, 30, 15)
    out = []
    for sw in (0, 1, 2, 3, 5, 1000):
        d = drv(variant=variant, precision=prec, tiles=tiles)
        d.set_tile_strip(sw)
        d.initialize(shape, D, dt, m)
        got = d.tile_strip()
        assert got == min(sw or 8, 6)  # 6 tile columns of 32 from x = bord
        d.run(24, isrc, graph_chunk=8)
        out.append((d.field("pc"), d.field("qc")))
        d.finalize()
    for o in out[1:]:
        for k in (0, 1):
            assert np.array_equal(out[0][k], o[k])
    assert fl.lib().fl_set_tile_strip(-1) == -1  # FL_ERR_ARG
    fl.lib().fl_clear_error()


@pytest.mark.parametrize("variant,prec", [("fused", "fp32"), ("fused", "fp32r"), ("fused", "bf16"),
                                          ("fused", "fp16r"), ("fused", "q16r"), ("twopass", "fp32")])
def test_stream_hints_are_bitwise_identical(fl, drv, variant, prec):
    """The streaming cache hints change only the cache policy of the previous-level and
    coefficient accesses: the fields after 40 steps are the same bits with and without."""
    m, dt = hetero(fl, (60, 33, 41), 6)
    shape = m[0].shape
    isrc = fl.index(shape, 20, 17, 15)
    out = []
    for h in (False, True):
        d = drv(variant=variant, precision=prec, tiles=(12, 0))
        d.set_stream_hints(h)
        assert d.stream_hints() == h
        d.initialize(shape, D, dt, m)
        d.run(40, isrc, graph_chunk=10)
        out.append((d.field("pc"), d.field("qc")))
        d.finalize()
    for k in (0, 1):
        assert np.array_equal(out[0][k], out[1][k])
    assert fl.lib().fl_set_stream_hints(2) == -1  # FL_ERR_ARG
    fl.lib().fl_clear_error()


def test_twopass_tiles_bitwise(fl, drv):
    m, dt = hetero(fl, (40, 30, 28), 6)
    shape = m[0].shape
    isrc = fl.index(shape, 20, 17, 15)
    out = []
    for t in [(8, 0), (4, 9), (16, 64)]:
        d = drv(variant="twopass", tiles=t)
        d.initialize(shape, D, dt, m)
        d.run(30, isrc, graph_chunk=0)
        out.append(d.field("pc"))
        d.finalize()
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])


def test_twopass_first_step_derivative_volumes(fl, oracle, drv):
    """Pass-1 volumes of the very first step (freshly allocated, zero-filled on the same
    stream before pass 1 runs) equal the oracle's first derivatives of the staged pc/qc."""
    m, dt = hetero(fl, (30, 26, 22), 5)
    shape = m[0].shape
    isrc = fl.index(shape, 17, 14, 12)
    d = drv(variant="twopass")
    d.initialize(shape, D, dt, m)
    d.run(3, isrc, graph_chunk=0)          # some non-trivial pc, qc
    pc, qc = d.field("pc"), d.field("qc")
    d.propagate(4)
    vols = [d.derivative_volume(k) for k in range(4)]
    sz, sy, sx = shape
    L = [0.8, -0.2, 0.0380952380952381, -0.0035714285714285713]
    for k, (f, ax) in enumerate([(pc, 2), (pc, 1), (qc, 2), (qc, 1)]):
        e = np.zeros(shape)
        for j in range(1, 5):
            e += L[j - 1] * (np.roll(f, -j, axis=ax).astype(np.float64) - np.roll(f, j, axis=ax))
        e /= 12.5
        sl = (slice(None), slice(None), slice(4, sx - 4)) if ax == 2 else \
            (slice(None), slice(4, sy - 4), slice(4, sx))
        assert np.abs(e[sl]).max() > 0
        assert np.allclose(vols[k][sl], e[sl], rtol=1e-5, atol=1e-6 * np.abs(e[sl]).max()), k


def test_graph_run_equals_call_by_call(fl, drv):
    m, dt = hetero(fl, (50, 40, 30), 6)
    shape = m[0].shape
    isrc = fl.index(shape, 25, 20, 12)
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    d.record_traces(14, 3, 30)
    d.step_loop(77, isrc)
    a = (d.field("pc"), d.field("qc"), d.field("pp"), d.traces())
    d.finalize()
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    d.record_traces(14, 3, 30)
    d.run(40, isrc, graph_chunk=10)            # 4 graph replays
    d.run(37, isrc, it0=40, graph_chunk=8)     # 4 replays + 5 direct steps
    b = (d.field("pc"), d.field("qc"), d.field("pp"), d.traces())
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_iso_p_equals_q_and_halo_zero(fl, drv):
    shape, m = ref_model(fl, "ISO", 40, 8)
    for variant in ("fused", "direct", "twopass"):
        d = drv(variant=variant)
        d.initialize(shape, D, 0.001, m)
        d.step_loop(60, fl.center_index(shape))
        pc, qc, pp = d.field("pc"), d.field("qc"), d.field("pp")
        assert np.array_equal(pc, qc), variant
        assert np.abs(pc).max() > 0
        for f in (pc, pp):
            assert not f[:4].any() and not f[-4:].any() and not f[:, :4].any()
            assert not f[:, -4:].any() and not f[:, :, :4].any() and not f[:, :, -4:].any()
        d.finalize()


def test_autotune_keeps_state_and_result(fl, drv, tmp_path):
    m, dt = hetero(fl, (64, 40, 48), 6)
    shape = m[0].shape
    isrc = fl.index(shape, 30, 20, 14)
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    d.run(20, isrc, graph_chunk=0)
    before = d.field("pc")
    cache = str(tmp_path / "tiles.txt")
    ty, zt = d.autotune(2, cache)
    hints = d.stream_hints()
    assert ty in (4, 8, 12, 16) and zt > 0
    assert np.array_equal(d.field("pc"), before)
    d.run(20, isrc, it0=20, graph_chunk=0)
    after_tuned = d.field("pc")
    d.finalize()
    d = drv(variant="fused", tiles=(8, 0))
    d.set_stream_hints(not hints)
    d.initialize(shape, D, dt, m)
    d.run(40, isrc, graph_chunk=0)
    assert np.array_equal(d.field("pc"), after_tuned)
    with open(cache) as f:
        assert f.read().count("|") >= 4  # cache line written
    d.finalize()
    d = drv(variant="fused")
    d.initialize(shape, D, dt, m)
    assert d.autotune(2, cache) == (ty, zt)  # served from the cache
    assert d.stream_hints() == hints


def test_initialize_with_host_fields_and_update_pointers(fl, drv):
    shape, m = ref_model(fl, "TTI", 16, 4)
    rng = np.random.default_rng(0)
    f = [np.zeros(shape, np.float32) for _ in range(4)]
    inner = (slice(4, -4),) * 3
    for a in f:
        a[inner] = rng.standard_normal(a[inner].shape).astype(np.float32) * 1e-3
    d = drv(variant="fused")
    d.initialize(shape, D, 0.001, m, fields=f)
    for k, name in enumerate(("pp", "pc", "qp", "qc")):
        assert np.array_equal(d.field(name), f[k])
    d.propagate(1)
    assert np.array_equal(d.update_pointers(), d.field("pc"))


def test_errors_are_reported_not_fatal(fl, drv):
    shape, m = ref_model(fl, "TTI", 8, 4)
    d = drv(variant="fused")
    d.initialize(shape, D, 0.001, m)
    with pytest.raises(fl.FletcherError):
        d.insert_source(0, int(np.prod(shape)) + 5, 1.0)
    with pytest.raises(fl.FletcherError):
        fl.lib().DRIVER_Propagate(shape[2] + 1, shape[1], shape[0], 4, 12.5, 12.5, 12.5, 0.001, 1,
                                  None, None, None, None)
        d._status()
    d.step_loop(2, fl.center_index(shape))  # still usable


@pytest.mark.parametrize("variant", ["fused", "direct", "twopass"])
@pytest.mark.parametrize("n3", [(1, 1, 1), (3, 2, 5), (33, 1, 2)])
def test_smallest_grids(fl, oracle, drv, variant, n3):
    """Interiors of one to a few cells with a one-cell absorbing zone (update domain of 3 to 35
    cells a side, single partial tiles and chunks): the same fields as the oracle."""
    shape = fl.grid_shape(n3[0], n3[1], n3[2], 1)
    m = fl.model_homogeneous(2, shape)
    fl.random_boundary(m, n3, 1)
    isrc = fl.center_index(shape)
    st = 6
    d = drv(variant=variant)
    d.initialize(shape, D, 0.001, m)
    d.run(st, isrc, graph_chunk=0)
    ov = "original" if variant == "direct" else "der1der1"
    r = oracle.run(ov, shape, 4, D, 0.001, st, isrc, oracle.precompute(m))
    assert np.abs(r["pc"]).max() > 0
    for k in ("pc", "qc"):
        got = d.field(k)
        if variant == "direct":
            assert np.array_equal(got, r[k]), k
        else:
            assert rel_l2(got, r[k]) <= TOL, (k, rel_l2(got, r[k]))

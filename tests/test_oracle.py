"""Pins the CPU oracle (oracle/fl_oracle.c) before it is trusted as the checker.

* golden fixtures written by the reference C/OpenMP program itself (tests/golden, made by
  tests/golden/make_fixtures.py from oracle/_ref): bitwise equality of pc, qc, pp, qp;
* the live reference build when oracle/_ref is present: bitwise on more shapes;
* mathematics: polynomial exactness and error constants of Der1/Der2 (which fix the
  weights of derivatives.h:13-41 uniquely), DerCross = Der1 o Der1 and exactness on
  x^a y^b, the closed-form Ricker, the ISO reduction p == q, point symmetry of the
  operator, the untouched halo.
"""
import os

import numpy as np
import pytest

from conftest import golden, rel_l2

FORMS = ("ISO", "VTI", "TTI")
D = (12.5, 12.5, 12.5)


def ref_model(fl, form, nx, absorb):
    shape = fl.grid_shape(nx, nx, nx, absorb)
    m = fl.model_homogeneous(form, shape)
    fl.random_boundary(m, (nx, nx, nx), absorb)
    return shape, m


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("variant", ["original", "der1der1"])
@pytest.mark.parametrize("steps", [10, 30])
def test_oracle_matches_reference_fixture(oracle, fl, form, variant, steps):
    g = golden("ref_%s_%s_n8_a4_s%d.npz" % (form, variant, steps))
    shape, m = ref_model(fl, form, 8, 4)
    r = oracle.run(variant, shape, 4, D, 0.001, steps, int(g["isource"]), oracle.precompute(m))
    for k in ("pc", "qc", "pp", "qp"):
        assert np.array_equal(r[k], g[k]), (k, rel_l2(r[k], g[k]))


def test_oracle_matches_reference_fixture_absorb16(oracle, fl):
    g = golden("ref_TTI_original_n24_a16_s10.npz")
    shape, m = ref_model(fl, "TTI", 24, 16)
    r = oracle.run("original", shape, 4, D, 0.001, 10, int(g["isource"]), oracle.precompute(m))
    assert np.array_equal(r["pc"], g["pc"]) and np.array_equal(r["qc"], g["qc"])


@pytest.mark.parametrize("variant", ["der1der1hm", "der1der1lm"])
@pytest.mark.parametrize("fp", ["f32", "f64"])
def test_loop_orders_equal_der1der1(oracle, fl, variant, fp):
    """der1der1hm (first derivatives formed plane by plane in z order) and der1der1lm (a
    9-plane ring) restate der1der1's arithmetic in other loop orders: same bits, FP32 and
    FP64, on a ragged heterogeneous grid and on the reference TTI model."""
    m = fl.model_gradient((15, 11, 9), 5, seed=11, nblobs=4, blob_smin=2.0, blob_smax=5.0,
                          smooth_smin=3.0, smooth_smax=8.0)
    shape = m[0].shape
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, 4 + 5 + 2)
    c = oracle.precompute(m, fp)
    a = oracle.run("der1der1", shape, 4, D, 0.0008, 24, isrc, c, fp=fp)
    b = oracle.run(variant, shape, 4, D, 0.0008, 24, isrc, c, fp=fp)
    for k in ("pc", "qc", "pp", "qp"):
        assert np.array_equal(a[k], b[k]), k
    assert float(np.abs(a["pc"]).max()) > 0
    g = golden("ref_TTI_der1der1_n8_a4_s30.npz")
    shape, m = ref_model(fl, "TTI", 8, 4)
    r = oracle.run(variant, shape, 4, D, 0.001, 30, int(g["isource"]), oracle.precompute(m))
    assert np.array_equal(r["pc"], g["pc"]) and np.array_equal(r["qc"], g["qc"])


def test_lowmem_reference_equals_der1der1():
    """der1der1lm's 9-plane ring (the z-queue precedent) gives der1der1 bit for bit."""
    a = golden("ref_TTI_der1der1lm_n8_a4_s30.npz")
    b = golden("ref_TTI_der1der1_n8_a4_s30.npz")
    assert np.array_equal(a["pc"], b["pc"]) and np.array_equal(a["qc"], b["qc"])


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                     "libref_original.so")),
                    reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("variant", ["original", "der1der1", "der1der1hm", "der1der1lm"])
def test_oracle_matches_live_reference_ragged(oracle, fl, variant):
    """Non-cubic grid, heterogeneous model: the reference back-end driven on the same arrays."""
    n3, absorb = (13, 9, 11), 5
    m = fl.model_gradient(n3, absorb, seed=7, nblobs=4, blob_smin=2.0, blob_smax=5.0,
                          smooth_smin=3.0, smooth_smax=8.0)
    shape = m[0].shape
    R = oracle.Ref(variant)
    R.setup(shape, 4, D, 0.0008, m)
    isrc = fl.index(shape, shape[2] // 2, shape[1] // 2, 4 + absorb + 3)
    R.steps(0, 25, isrc)
    coefs = oracle.precompute(m)
    for k, c in enumerate(R.fetch_coefs()):
        assert np.array_equal(c, coefs[k]), k
    r = oracle.run(variant, shape, 4, D, 0.0008, 25, isrc, coefs)
    assert np.array_equal(r["pc"], R.fetch("pc"))
    assert np.array_equal(r["qc"], R.fetch("qc"))


# ------------------------------------------------------------------ derivative closed forms

def _line(fn, h, n=9):
    x = (np.arange(n) - n // 2) * h
    return np.ascontiguousarray(fn(x), np.float64)


def test_der1_polynomial_exactness_and_error_constant(oracle):
    L = oracle.lib()
    h = 0.7
    for k in range(0, 9):  # exact through degree 8
        p = _line(lambda x: x ** k, h)
        exact = k * 0.0 ** (k - 1) if k >= 1 else 0.0
        assert abs(L.oracle_der1_f64(p, 4, 1, 1.0 / h) - exact) < 1e-9
    # x^9: D1 = f'(0) + error with (4!)^2/9! * 9! h^8 = 576 h^8 (sign from the central form)
    p = _line(lambda x: x ** 9, h)
    assert np.isclose(L.oracle_der1_f64(p, 4, 1, 1.0 / h), -576.0 * h ** 8, rtol=1e-12)
    # derivative of a linear function, stride != 1
    p = np.ascontiguousarray(np.repeat(3.0 * np.arange(9) * h, 5), np.float64)
    assert np.isclose(L.oracle_der1_f64(p, 20, 5, 1.0 / h), 3.0, rtol=1e-13)


def test_der2_polynomial_exactness_and_error_constant(oracle):
    L = oracle.lib()
    h = 1.3
    for k in range(0, 10):  # exact through degree 9
        p = _line(lambda x: x ** k, h)
        exact = 2.0 if k == 2 else 0.0
        assert abs(L.oracle_der2_f64(p, 4, 1, 1.0 / h ** 2) - exact) < 1e-8, k
    # x^10 at 0 (f''=0): the 8th-order central error -2 (4!)^2/10! h^8 f^(10) = -1152 h^8
    p = _line(lambda x: x ** 10, h)
    err = L.oracle_der2_f64(p, 4, 1, 1.0 / h ** 2)
    assert np.isclose(err, -2.0 * 576.0 * h ** 8, rtol=1e-10)


def test_der2_plane_wave_symbol_at_nyquist(oracle):
    """(-1)^i sampled wave: symbol (K0 + 2 sum K_m cos(m pi)) = -6.5016 (1/h^2)."""
    L = oracle.lib()
    p = np.ascontiguousarray((-1.0) ** np.arange(9), np.float64)
    assert np.isclose(L.oracle_der2_f64(p, 4, 1, 1.0), -(205 / 72 + 3.2 + 0.4 + 16 / 315 + 1 / 280),
                      rtol=1e-14)


def _grid2(fn, h1, h2):
    a = (np.arange(9) - 4) * h1
    b = (np.arange(9) - 4) * h2
    X, Y = np.meshgrid(a, b, indexing="xy")  # Y rows (stride 9), X fastest
    return np.ascontiguousarray(fn(X, Y), np.float64)


def test_dercross_exact_on_monomials(oracle):
    L = oracle.lib()
    h1, h2 = 0.9, 1.1
    for a in range(0, 9):
        for b in range(0, 9):
            p = _grid2(lambda X, Y: X ** a * Y ** b, h1, h2)
            v = L.oracle_dercross_f64(p, 40, 1, 9, 1.0 / (h1 * h2))
            exact = 1.0 if (a == 1 and b == 1) else 0.0
            assert abs(v - exact) < 1e-7, (a, b, v)


def test_dercross_equals_der1_of_der1(oracle):
    """DerCross = Der1_s21(Der1_s11(p)) in FP64 for random data; catches a wrong L_ij
    (the reference labels L13 'L1*L2' but its value is L1*L3)."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    p = np.ascontiguousarray(rng.standard_normal(81), np.float64)
    # first derivative along x (stride 1) on the 9x9 patch, then along y (stride 9)
    dx = np.zeros((9, 9))
    P = p.reshape(9, 9)
    for r in range(9):
        row = np.ascontiguousarray(P[r], np.float64)
        dx[r, 4] = L.oracle_der1_f64(row, 4, 1, 1.0)
    col = np.ascontiguousarray(dx[:, 4], np.float64)
    nested = L.oracle_der1_f64(col, 4, 1, 1.0)
    direct = L.oracle_dercross_f64(p, 40, 1, 9, 1.0)
    assert np.isclose(direct, nested, rtol=1e-12, atol=1e-12)
    # FP32 instantiation uses the reference literals: agree to FP32 rounding
    p32 = np.ascontiguousarray(p, np.float32)
    assert np.isclose(L.oracle_dercross_f32(p32, 40, 1, 9, 1.0), direct, rtol=1e-5, atol=1e-5)


# ------------------------------------------------------------------ source, symmetries

def test_ricker_closed_form(oracle):
    """Ricker wavelet of peak frequency FCUT/3 delayed by 2 sqrt(pi)/FCUT."""
    fcut = 40.0
    f = fcut / 3.0
    t0 = 2.0 * np.sqrt(np.pi) / fcut
    for it in range(0, 400, 7):
        t = it * 0.001
        a = (np.pi * f * (t - t0)) ** 2
        exact = (1 - 2 * a) * np.exp(-a)
        assert abs(oracle.source(0.001, it) - exact) < 2e-6


def test_iso_reduction_p_equals_q(oracle, fl):
    """ISO: ch1dzz=1, others 0, v2px=v2pz=v2pn, v2sz=0 and the source feeds p and q equally,
    so p and q stay bitwise equal."""
    shape, m = ref_model(fl, "ISO", 12, 4)
    c = oracle.precompute(m)
    assert np.all(c[2] == 1.0) and all(np.all(c[k] == 0) for k in (0, 1, 3, 4, 5, 8))
    r = oracle.run("original", shape, 4, D, 0.001, 40, fl.center_index(shape), c)
    assert np.array_equal(r["pc"], r["qc"]) and np.abs(r["pc"]).max() > 0


@pytest.mark.parametrize("variant", ["original", "der1der1"])
def test_point_symmetry_and_halo(oracle, fl, variant):
    """Homogeneous TTI without the random layer, centred source: the even operator keeps
    p(x) = p(-x) about the source (FP64, before the wave reaches the edges); the 4-cell
    halo stays exactly zero."""
    nx, absorb = 17, 4  # odd extent: the update domain is symmetric about the centre cell
    shape = fl.grid_shape(nx, nx, nx, absorb)
    m = fl.model_homogeneous("TTI", shape)
    c = oracle.precompute(m, "f64")
    isrc = fl.center_index(shape)
    r = oracle.run(variant, shape, 4, D, 0.001, 20, isrc, c, fp="f64")
    pc = r["pc"]
    sz, sy, sx = shape
    cz, cy, cx = sz // 2, sy // 2, sx // 2
    w = 8
    box = pc[cz - w:cz + w + 1, cy - w:cy + w + 1, cx - w:cx + w + 1]
    assert np.allclose(box, box[::-1, ::-1, ::-1], rtol=1e-10, atol=1e-14 * np.abs(box).max())
    for f in (r["pc"], r["qc"], r["pp"], r["qp"]):
        assert not f[:4].any() and not f[-4:].any()
        assert not f[:, :4].any() and not f[:, -4:].any()
        assert not f[:, :, :4].any() and not f[:, :, -4:].any()


def test_fp64_oracle_close_to_fp32(oracle, fl):
    shape, m = ref_model(fl, "TTI", 12, 4)
    isrc = fl.center_index(shape)
    a = oracle.run("original", shape, 4, D, 0.001, 30, isrc, oracle.precompute(m))
    b = oracle.run("original", shape, 4, D, 0.001, 30, isrc, oracle.precompute(m, "f64"), fp="f64")
    assert rel_l2(a["pc"], b["pc"]) < 1e-5


def test_traces_are_plane_samples(oracle, fl):
    shape, m = ref_model(fl, "TTI", 10, 4)
    isrc = fl.center_index(shape)
    c = oracle.precompute(m)
    r = oracle.run("der1der1", shape, 4, D, 0.001, 12, isrc, c, rec_iz=shape[0] // 2, rec_every=4)
    assert r["traces"].shape == (3, shape[1] - 8, shape[2] - 8)
    last = r["pc"][shape[0] // 2, 4:-4, 4:-4]
    assert np.array_equal(r["traces"][2], last)

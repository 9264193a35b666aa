"""The reference's own front-end on the device: the unmodified original/ main.c, model.c,
boundary.c, source.c, utils.c (oracle/_ref/libref_dropin.so, built by `make -C oracle ref`)
linked against libfletcher_amd.so in place of its OpenMP back-end, run with the reference's
default invocation (TTI 88^3, absorb 16, dt 0.001, 100 steps).  Its pc after the last step
(copied back by the front-end's own DRIVER_Update_pointers call, model.c:99-101) must match
the reference OpenMP build of the same front-end.  Variants are chosen through FL_VARIANT,
as a front-end that only calls driver.h would.
"""
import os

import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ARGS = ["TTI", 88, 88, 88, 16, 12.5, 12.5, 12.5, 0.001, 0.1]


@pytest.fixture(scope="module")
def cpu_ref(oracle):
    if not (oracle.ref_available("original") and oracle.ref_available("dropin")):
        pytest.skip("reference builds absent (oracle/_ref)")
    return oracle.Ref("original").main(ARGS)


@pytest.mark.parametrize("variant", ["fused", "twopass", "direct"])
def test_reference_front_end_drives_the_device(oracle, cpu_ref, variant, monkeypatch):
    monkeypatch.setenv("FL_VARIANT", variant)
    got = oracle.Ref("dropin").main(ARGS)
    assert got["shape"] == cpu_ref["shape"] == (128, 128, 128)
    assert abs(got["pc"]).max() > 0
    err = rel_l2(got["pc"], cpu_ref["pc"])
    print("%s: relL2(pc) vs reference OpenMP = %.3e" % (variant, err))
    assert err <= 1e-5


@pytest.mark.parametrize("prec", ["fp32", "fp32r"])
def test_reference_front_end_stream_hints(oracle, cpu_ref, prec, monkeypatch):
    """FL_STREAM_HINTS=1 then 0: the same pc bits both ways, within 1e-5 of the reference
    OpenMP build."""
    monkeypatch.setenv("FL_VARIANT", "fused")
    monkeypatch.setenv("FL_PRECISION", prec)
    got = []
    for h in ("1", "0"):
        monkeypatch.setenv("FL_STREAM_HINTS", h)
        got.append(oracle.Ref("dropin").main(ARGS)["pc"])
    assert (got[0] == got[1]).all()
    assert rel_l2(got[0], cpu_ref["pc"]) <= 1e-5


def test_reference_front_end_fp32_recompute(oracle, cpu_ref, monkeypatch):
    """FL_PRECISION=fp32r: 24-byte records, angular coefficients rebuilt in FP32 on the
    device; the same 1e-5 bar against the reference OpenMP build."""
    monkeypatch.setenv("FL_VARIANT", "fused")
    monkeypatch.setenv("FL_PRECISION", "fp32r")
    got = oracle.Ref("dropin").main(ARGS)
    err = rel_l2(got["pc"], cpu_ref["pc"])
    print("fp32r: relL2(pc) vs reference OpenMP = %.3e" % err)
    assert err <= 1e-5

"""The arithmetic behind marching a z chunk downwards (odd steps, csrc/fl_step.h step_tile):
the z queues then hold the column in reverse order.  Evaluated here on the host with the
kernels' own sums (fl_kernels.h fd1 / fd2: the reference's float weights, one rounding per
multiply or fused multiply-add, libm fmaf), over random FP32 columns with repeated values,
mixed signs and subnormals:
  - the second derivative (sums of symmetric pairs) is bit for bit the same;
  - the first derivative comes out exactly negated, and a negated scale (or a multiply by -1)
    gives the upward value back, bit for bit apart from the sign of an exact zero.
A mistake in the pairing of the reversed queue (an off-by-one slot) fails both."""
import ctypes

import numpy as np

L = [np.float32(v) for v in ("0.8", "-0.2", "0.0380952380952381", "-0.0035714285714285713")]
K = [np.float32(v) for v in ("-2.84722222222222222222", "1.6", "-0.2", "0.02539682539682539682",
                             "-0.00178571428571428571")]

_m = ctypes.CDLL("libm.so.6")
_m.fmaf.restype = ctypes.c_float
_m.fmaf.argtypes = [ctypes.c_float] * 3


def fmaf(a, b, c):
    return np.float32(_m.fmaf(float(a), float(b), float(c)))


def fd1(q):  # der1_q: pairs (5,3) (6,2) (7,1) (8,0), unscaled
    t = np.float32(L[0] * np.float32(q[5] - q[3]))
    for k in (1, 2, 3):
        t = fmaf(L[k], np.float32(q[5 + k] - q[3 - k]), t)
    return t


def fd2(q):  # der2_q: centre 4, pairs (5,3) .. (8,0), unscaled
    t = np.float32(K[0] * q[4])
    for k in (1, 2, 3, 4):
        t = fmaf(K[k], np.float32(q[4 + k] + q[4 - k]), t)
    return t


def _columns(n, seed=1234):
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-44, 6, size=(n, 9))
    q = (rng.standard_normal((n, 9)) * mag).astype(np.float32)
    # repeated values (exact zero differences and sums of opposite values)
    rep = rng.random((n, 9)) < 0.2
    q[rep] = q[:, :1].repeat(9, axis=1)[rep]
    q[rng.random((n, 9)) < 0.1] = 0.0
    neg = rng.random(n) < 0.1  # antisymmetric about the centre: the pair sums vanish
    q[neg, 5:] = -q[neg, 3::-1]
    sym = (rng.random(n) < 0.1) & ~neg  # symmetric: the first derivative is an exact zero
    q[sym, 5:] = q[sym, 3::-1]
    return q


def _bits(x):
    return np.float32(x).view(np.uint32)


def test_reversed_column_derivatives():
    q = _columns(6000)
    dz = np.float32(1.0 / 12.5)
    zero_sign_only = 0
    for col in q:
        rev = col[::-1].copy()
        up1, dn1 = fd1(col), fd1(rev)
        assert _bits(fd2(rev)) == _bits(fd2(col))
        assert _bits(dn1) == _bits(-up1) or (up1 == 0 and dn1 == 0)
        for a, b in ((np.float32(up1 * dz), np.float32(dn1 * -dz)), (up1, np.float32(dn1 * np.float32(-1.0)))):
            if _bits(a) != _bits(b):
                assert a == 0 and b == 0  # only the sign of an exact zero may differ
                zero_sign_only += 1
    assert 0 < zero_sign_only < len(q) // 2


def test_off_by_one_reversal_is_caught():
    # planted mistake: the reversed queue read one slot off (plane z+5 .. z-3)
    q = _columns(200, seed=7)
    bad = 0
    for col in q:
        ext = np.concatenate([col, col[:1] * np.float32(1.5)])
        rev = ext[::-1][:9].copy()
        if _bits(fd1(rev)) != _bits(-fd1(col)) or _bits(fd2(rev)) != _bits(fd2(col)):
            bad += 1
    assert bad > len(q) // 2

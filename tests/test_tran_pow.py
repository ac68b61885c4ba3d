"""The adaptive transient controller's power (ddpow.hpp, run on the host
through rs_host_pow_dd; no GPU needed).

The controller raises 1/e_n to 0.35, 1/e_prev to 0.2 and 1/err to 0.5
(analysis.rs:173, 530-536); the reference computes these with libm's pow.
The double-double power is pinned by
- libm (Python's math.pow is the C library's pow) on 300,000 pairs: every
  disagreement must be one the device flags as near a rounding midpoint, so
  the device hands exactly those to the host's libm;
- exact values where the power is representable (4^0.5, 2^-1 ...) and
  high-precision decimal arithmetic on a sample of the disagreements, where the
  double-double result must be the correctly rounded one.
"""
import ctypes as C
import math
from decimal import Decimal, getcontext

import numpy as np

from rustspice_amd import _lib


def pow_dd(x, y, zone=0.02):
    L = _lib.lib()
    L.rs_host_pow_dd.restype = C.c_int32
    L.rs_host_pow_dd.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p]
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    out = np.zeros(len(x))
    near = np.zeros(len(x), np.uint8)
    L.rs_host_pow_dd(x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x), zone,
                     out.ctypes.data_as(C.c_void_p), near.ctypes.data_as(C.c_void_p))
    return out, near.astype(bool)


def test_exact_powers():
    x = np.array([4.0, 0.25, 1.0, 9.0, 1024.0, 2.0 ** -20, 1e10])
    y = np.array([0.5, 0.5, 0.35, 0.5, 0.2, 0.5, 0.0])
    out, _ = pow_dd(x, y)
    assert list(out) == [2.0, 0.5, 1.0, 3.0, 4.0, 2.0 ** -10, 1.0]


def test_matches_libm_except_flagged_midpoints():
    rng = np.random.RandomState(11)
    x = np.exp(rng.uniform(-23.0, 23.0, 100000))
    xs = np.concatenate([x, x, x])
    ys = np.concatenate([np.full(len(x), 0.7 / 2.0), np.full(len(x), 0.4 / 2.0), np.full(len(x), 0.5)])
    out, near = pow_dd(xs, ys)
    libm = np.array([math.pow(a, b) for a, b in zip(xs, ys)])
    differ = out != libm
    assert not (differ & ~near).any(), np.flatnonzero(differ & ~near)[:5]
    assert 0.01 < near.mean() < 0.08  # about 2 * 0.02 of the calls
    # where they differ, the double-double power is the correctly rounded one
    getcontext().prec = 60
    for i in np.flatnonzero(differ)[:20]:
        exact = (Decimal(float(xs[i])).ln() * Decimal(float(ys[i]))).exp()
        assert abs(Decimal(float(out[i])) - exact) <= abs(Decimal(float(libm[i])) - exact)

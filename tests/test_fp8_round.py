"""FP8 E4M3 rounding of the 8-bit pool (qdfn_fp8_round, the host rounding
behind QDFN_PREC_MIXED8's stored entries; no reference counterpart).

Pinned against torch's float8_e4m3fn (the OCP E4M3 format: bias 7, no
infinities, 448 the largest finite value): every one of the 254 finite codes
round-trips exactly; the midpoints between neighbouring values round to the
neighbour with the even code; random float32-exact inputs across the normal
and subnormal range agree with torch's conversion.  Values past 448 saturate
(torch gives NaN there, so those are checked against 448 directly)."""
import ctypes as C

import numpy as np
import pytest
import torch

from qdfn._lib import lib

PD = C.POINTER(C.c_double)


def _round(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    assert lib().qdfn_fp8_round(x.ctypes.data_as(PD), y.ctypes.data_as(PD), x.size) == 0
    return y


def _torch(x):
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(torch.float8_e4m3fn).to(torch.float64).numpy()


def _codes():
    v = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    return v[np.isfinite(v)]


def test_codes_round_trip():
    v = _codes()
    assert v.size == 254
    np.testing.assert_array_equal(_round(v), v)
    assert v.max() == 448.0 and np.min(v[v > 0]) == 2.0 ** -9


def test_midpoints_round_to_even():
    v = np.unique(_codes()[_codes() >= 0])
    mid = (v[:-1] + v[1:]) / 2
    got = _round(mid)
    np.testing.assert_array_equal(got, _torch(mid))
    # the even code: the neighbour whose mantissa's last bit is zero
    lo_code = torch.from_numpy(v[:-1].astype(np.float32)).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    expect = np.where(lo_code % 2 == 0, v[:-1], v[1:])
    np.testing.assert_array_equal(got, expect)
    np.testing.assert_array_equal(_round(-mid), -expect)


@pytest.mark.parametrize("lo,hi", [(2.0 ** -12, 2.0 ** -6), (2.0 ** -6, 1.0), (1.0, 448.0)])
def test_random_values_agree_with_torch(lo, hi):
    rng = np.random.default_rng(5)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), 20000)).astype(np.float32).astype(np.float64)
    x *= rng.choice([-1.0, 1.0], x.size)
    np.testing.assert_array_equal(_round(x), _torch(x))


def test_saturation_and_zero():
    np.testing.assert_array_equal(_round([450.0, 463.9, 465.0, 1e6, -1e6]), [448.0, 448.0, 448.0, 448.0, -448.0])
    y = _round([0.0, -0.0, 2.0 ** -11, 3 * 2.0 ** -11])
    assert y[0] == 0.0 and y[1] == 0.0 and np.signbit(y[1])
    assert y[2] == 0.0 and y[3] == 2.0 ** -9  # 0.25 quantum rounds down to 0, 0.75 up to one quantum

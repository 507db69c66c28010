"""The float32 kernels divide through fh::Divisor: a Newton-refined hardware
reciprocal shared by the dividends of one divisor and one remainder
correction.  This checks on the device that it is IEEE division bit for bit:
every reciprocal significand at nine exponents from 2^-100 to 2^99, and
quotients over random pairs in its exact domain (a = +-0, or 2^-100 <= |a|,
|a / b| <= 2^100, with 2^-100 <= b <= 2^100).  Dividends below 2^-100 lie
outside it (the remainder can underflow): there at most a small fraction of
quotients may differ, by one ulp.  Zero-sign and one-ulp mistakes are what
this catches; a plausible slip (the remainder's sign, a missing Newton step)
makes millions of mismatches."""

import ctypes

import pytest

from floodhip import _lib

pytestmark = pytest.mark.gpu


def test_division_is_ieee_division_in_its_domain():
    counts = (ctypes.c_int64 * 6)()
    assert _lib.load().fh_division_selftest(counts) == _lib.FH_OK
    recip_bad, recip_n, quot_bad, quot_n, tiny_bad, tiny_n = list(counts)
    print(f"\nreciprocals {recip_bad}/{recip_n}, quotients {quot_bad}/{quot_n}, "
          f"dividends below 2^-100 {tiny_bad}/{tiny_n}", flush=True)
    assert recip_n == 9 * 2**23 and recip_bad == 0
    assert quot_n > 40_000_000 and quot_bad == 0
    assert tiny_n > 5_000_000 and tiny_bad < 0.05 * tiny_n

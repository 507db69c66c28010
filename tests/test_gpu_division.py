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


def test_friction_power_is_correctly_rounded():
    """x^p, p = (float)(4/3), as the float32 friction terms compute it, against the
    correctly rounded value over every float in [2^-40, 2^20): misrounded in
    about 1e-5 of the values (measured 4,284 of 503 M), each by one ulp, where
    the device library's powf is one ulp off in about a quarter of them."""
    counts = (ctypes.c_int64 * 5)()
    assert _lib.load().fh_pow_selftest(counts) == _lib.FH_OK
    bad, n, lib_bad, ulp, lib_ulp = list(counts)
    print(f"\nfric_pow {bad}/{n} misrounded (max {ulp} ulp); powf {lib_bad}/{n} (max {lib_ulp} ulp)", flush=True)
    assert n == 60 * 2**23
    assert ulp <= 1 and bad <= 2e-5 * n
    assert 1000 * bad < lib_bad

"""Arithmetic the reference form takes from hardware instructions whose
rounding no document fixes, checked on the device: the correctly rounded
cube root of its XYB, for every positive normal float, and the division of
its SSIM' map, over a spread of significand pairs.

The kernels seed the root with the hardware base-2 logarithm and exponential
(v_log_f32, v_exp_f32), whose rounding no document fixes, and finish it with
one FP64 step and the exact midpoint test (vqm_amd/csrc/ss2_cbrt.h,
cbrt_rn_near_hw). ``vqm_ss2_check_cbrt`` runs, on the device, that root and
``cbrt_rn`` (IEEE steps only, which test_cbrt.py checks against exact integer
arithmetic for every float in [2^-126, 2^24) on the CPU) over a range of
inputs and counts the inputs where they differ.
"""
from __future__ import annotations

import ctypes

import pytest

from vqm_amd import _lib as L

pytestmark = pytest.mark.gpu

FIRST_NORMAL, END_NORMAL = 0x00800000, 0x7F800000


def check(first: int, end: int) -> L.CbrtCheckC:
    r = L.CbrtCheckC()
    rc = L.lib().vqm_ss2_check_cbrt(first, end, ctypes.byref(r))
    assert rc == 0, L.lib().vqm_last_error()
    return r


def test_hardware_seeded_root_equals_the_ieee_root_for_every_normal_float():
    r = check(FIRST_NORMAL, END_NORMAL)
    assert r.mismatches == 0, f"first differing input bits {r.first_mismatch:#010x}"
    assert r.first_mismatch == 0xFFFFFFFF
    # One FP64 step keeps the root inside the midpoint test's 2^-44 band
    # with room to spare while the seed's |1 - x r^3| stays below 2^-16
    # (dropped term 14/81 h^3 < 2^-50.5); measured: 1.08e-5.
    assert 0.0 < r.max_seed_error < 2.0**-16
    # Inputs within 2^9 units of a midpoint: about one in 2^19.
    assert 0 < r.exact_steps < (END_NORMAL - FIRST_NORMAL) >> 16


def test_check_counts_a_partial_range():
    """A range the check splits across blocks unevenly: its counters add up
    to the same result (no input skipped, none counted twice)."""
    one = check(0x3F800000, 0x40000001)  # [1, 2]
    assert one.mismatches == 0
    halves = [check(0x3F800000, 0x3FC00123), check(0x3FC00123, 0x40000001)]
    assert sum(h.exact_steps for h in halves) == one.exact_steps
    assert max(h.max_seed_error for h in halves) == one.max_seed_error


def check_div(first: int, end: int) -> L.DivCheckC:
    r = L.DivCheckC()
    rc = L.lib().vqm_ss2_check_div(first, end, ctypes.byref(r))
    assert rc == 0, L.lib().vqm_last_error()
    return r


def test_map_division_equals_the_ieee_quotient():
    """div_rn (hardware reciprocal, one Newton step, one correction) against
    the IEEE quotient for 8 runs of 8192 b significands spread over [1, 2),
    each against all 2^23 a significands (2^39 pairs; all 2^46 take ~30 s);
    rcp_rn(b) against 1 / b; and the hardware reciprocal's invariance under
    scaling by 2^e for those b. Fused multiply-adds scale exactly, so
    together these cover the map's operand domain."""
    total_bad = total_rcp = 0
    for k in range(8):
        first = k * (1 << 20) + (k * 7919) % (1 << 19)
        r = check_div(first, first + 8192)
        total_bad += r.mismatches
        total_rcp += r.reciprocal_mismatches
        assert r.mismatches == 0, f"first differing b significand {r.first_mismatch:#08x}"
    assert total_bad == 0 and total_rcp == 0

"""GPU parity of fixed_weight_masks_u32 / _u64 against the oracle (which
tests/test_comb.py pins to the reference's known answers, brute force and the
live reference), and properties at sizes the oracle would take long on."""
import math

import numpy as np
import pytest

from bmmpy import _bmmpy

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("bits", [32, 64])
def test_fixed_weight_masks_match_oracle(lib, ora, bits):
    cases = [(5, 3), (6, 2), (1, 1), (7, 7), (20, 1), (20, 10), (24, 12), (bits, 1), (bits, 2),
             (bits, 3), (bits, bits - 1), (bits, bits - 2), (bits - 1, 4)]
    if bits == 32:
        cases += [(32, 5), (32, 27), (32, 32)]
    for n, k in cases:
        _say("fixed weight", bits, n, k)
        got = lib.fixed_weight_masks(n, k, bits)
        exp = ora.fixed_weight_masks(n, k, bits)
        assert got.size == math.comb(n, k), (n, k)
        assert np.array_equal(got.astype(np.uint64), exp), (n, k)


def test_fixed_weight_masks_all_ones_64():
    # n = k = 64: the single mask with every bit set.
    assert _bmmpy.fixed_weight_masks_u64(64, 64) == [(1 << 64) - 1]


def test_fixed_weight_masks_large_properties(lib):
    # C(30, 15) = 155,117,520 masks: increasing, weight 15, below 2^30, and
    # the sum of all masks is C(29, 14) * (2^30 - 1) (each bit is set in
    # C(29, 14) of them).
    _say("fixed weight 30 15")
    got = lib.fixed_weight_masks(30, 15)
    assert got.size == math.comb(30, 15)
    assert got[0] == (1 << 15) - 1 and got[-1] == ((1 << 15) - 1) << 15
    assert np.all(got[1:] > got[:-1])
    v = got.astype(np.uint32)
    v = v - ((v >> 1) & 0x55555555)
    v = (v & 0x33333333) + ((v >> 2) & 0x33333333)
    v = (v + (v >> 4)) & 0x0F0F0F0F
    assert np.all(((v * 0x01010101) >> 24) == 15)
    assert int(got.astype(np.uint64).sum()) == math.comb(29, 14) * ((1 << 30) - 1)


def test_fixed_weight_masks_binding_lists():
    assert _bmmpy.fixed_weight_masks_u32(5, 3) == [7, 11, 13, 14, 19, 21, 22, 25, 26, 28]
    assert _bmmpy.fixed_weight_masks_u64(64, 1) == [1 << i for i in range(64)]

"""fixed_weight_masks_u32 / _u64 (src/math/comb.cpp:13-82) without a GPU:
the oracle against the reference's known answers (tests/cpp/fwht_tests.cpp:
13-56), brute force and the live reference build; the entry points' answers
that need no device (k = 0, k > n, too-wide n) through the C-ABI and the
module binding."""
import math

import numpy as np
import pytest

import bmmpy as bmm
from bmmpy import _bmmpy


def test_oracle_known_answers(ora):
    assert ora.fixed_weight_masks(5, 3).tolist() == [
        0x07, 0x0B, 0x0D, 0x0E, 0x13, 0x15, 0x16, 0x19, 0x1A, 0x1C]
    assert ora.fixed_weight_masks(6, 2, 64).tolist() == [
        0x03, 0x05, 0x06, 0x09, 0x0A, 0x0C, 0x11, 0x12, 0x14, 0x18,
        0x21, 0x22, 0x24, 0x28, 0x30]


def test_oracle_brute_force_and_counts(ora):
    for n in range(0, 15):
        every = np.arange(1 << n, dtype=np.uint64)
        pop = np.array([bin(int(v)).count("1") for v in every])
        for k in range(0, n + 2):
            want = [0] if k == 0 else every[pop == k].tolist()
            assert ora.fixed_weight_masks(n, k).tolist() == want, (n, k)
    for w in (32, 64):
        for n, k in ((w, 1), (w, 2), (w, w - 1), (w, w), (w - 1, 3)):
            got = ora.fixed_weight_masks(n, k, w)
            assert got.size == math.comb(n, k)
            assert np.all(np.diff(got.astype(object)) > 0) if got.size > 1 else True


def test_oracle_matches_live_reference(ora, ref):
    for w in (32, 64):
        for n in (0, 1, 5, 17, w - 2, w - 1, w):
            for k in (0, 1, 2, 3, n - 2, n - 1, n, n + 1):
                if k < 0 or (w, n, k) == (64, 64, 64):
                    continue
                assert np.array_equal(ora.fixed_weight_masks(n, k, w),
                                      ref.fixed_weight_masks(n, k, w)), (w, n, k)
    # n = k = 64: the reference computes uint64_t(1) << 64, which C++ leaves
    # undefined (x86-64 masks the count and it returns [0]); the one 64-bit
    # mask with 64 set bits is all ones.
    assert ora.fixed_weight_masks(64, 64, 64).tolist() == [(1 << 64) - 1]
    for w, n in ((32, 33), (64, 65)):
        with pytest.raises(Exception, match=f"n exceeds {w}-bit mask width"):
            ref.fixed_weight_masks(n, 1, w)


def test_entry_points_without_device(lib):
    # Answers that need no transform: checked before any device query.
    assert lib.fixed_weight_masks(40, 0).tolist() == [0]
    assert lib.fixed_weight_masks(3, 5).tolist() == []
    assert lib.fixed_weight_masks(70, 0, 64).tolist() == [0]
    with pytest.raises(ValueError, match="fixed_weight_masks_u32: n exceeds 32-bit mask width"):
        lib.fixed_weight_masks(33, 1)
    with pytest.raises(ValueError, match="fixed_weight_masks_u64: n exceeds 64-bit mask width"):
        lib.fixed_weight_masks(65, 2, 64)
    assert _bmmpy.fixed_weight_masks_u32(40, 0) == [0]
    assert _bmmpy.fixed_weight_masks_u64(2, 3) == []
    with pytest.raises(ValueError, match="n exceeds 32-bit mask width"):
        _bmmpy.fixed_weight_masks_u32(33, 1)
    with pytest.raises(TypeError):
        _bmmpy.fixed_weight_masks_u32(-1, 1)
    assert bmm._bmmpy is _bmmpy

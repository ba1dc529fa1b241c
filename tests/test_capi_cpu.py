"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, validates arguments before touching a device (same
exception classes as the reference), and its seeded generators agree with the
oracle's bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bmmpy_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bmm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lib):
    so = lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    # and the Python binding types every one of them
    assert not [n for n in names if n not in lib.SIGNATURES]


def test_version_and_device_info_without_gpu(lib):
    assert lib.load().bmm_hip_version().decode()
    info = lib.device_info()
    assert info.device_count >= 0


def test_validation_precedes_device(lib):
    with pytest.raises(ValueError):
        lib.fwht_i32([1, 2, 3])
    with pytest.raises(ValueError):
        lib.gray_coset_min(np.zeros((4, 40), np.uint64), 3, [0])     # width > 32 words
    with pytest.raises(IndexError):
        lib.gray_coset_min(np.zeros((30, 2), np.uint64), 20, [31])
    with pytest.raises(ValueError):
        lib.coset_min(np.zeros((40, 1), np.uint64), 10)              # > 31 rows
    with pytest.raises(ValueError):
        lib.coset_min(np.zeros((1, 1), np.uint64), 10)               # < 2 rows


def test_no_device_is_runtime_error(lib):
    if lib.device_info().device_count > 0:
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError):
        lib.fwht_i32([1, 2, 3, 4])
    with pytest.raises(RuntimeError):
        lib.coset_min(np.zeros((4, 1), np.uint64), 10)


def test_generators_match_oracle(lib, ora):
    assert np.array_equal(lib.gen_bernoulli(0x5EED0002, 32, 1024, 0.3),
                          ora.bernoulli(0x5EED0002, 32, 1024, 0.3))
    S, A, M = lib.gen_planted(0x5EED0003, 48, 4096, 0.02)
    S2, A2, M2 = ora.planted(0x5EED0003, 48, 4096, 0.02)
    assert np.array_equal(S, S2) and np.array_equal(A, A2) and np.array_equal(M, M2)
    r, c, d, w, o = lib.gen_batch(0x5EED0004, 200)
    r2, c2, d2, w2, o2 = ora.batch(0x5EED0004, 200)
    assert np.array_equal(r, r2) and np.array_equal(c, c2) and np.array_equal(w, w2)
    assert np.array_equal(d, d2) and np.array_equal(o, o2)
    assert r.min() >= 8 and r.max() <= 28 and (c % 64 == 0).all() and c.max() <= 2048

"""GPU checks of the BitMatrix surface that runs on the device (power through
the GF(2) row-combination kernel) and of batched coset minima read from
BMMPYB1 files: the same result as from the in-memory batch, which the batch
parity tests pin to the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_power_matches_repeated_oracle_products(ora):
    import bmmpy as bmm
    for n, e in ((1, 0), (5, 1), (16, 7), (70, 13), (128, 100)):
        _say("power", n, e)
        a = ora.bernoulli(500 + n, n, n, 0.2)
        exp = np.zeros_like(a)
        for i in range(n):
            exp[i, i // 64] = np.uint64(1 << (i % 64))
        for _ in range(e):
            exp = ora.mul(exp, n, a, n)
        got = bmm.BitMatrix.from_words(a, n).power(e)
        assert np.array_equal(np.asarray(got.to_words(), np.uint64).reshape(n, -1), exp)


def test_batch_from_bmmpyb1_files(lib, tmp_path):
    from bmmpy_amd import batch
    _say("batch from BMMPYB1 files")
    rows, cols, dens, words, offs = lib.gen_batch(0x5EED0004, 64)
    paths = batch.save_batch(tmp_path, words, offs, rows, cols)
    w2, o2, r2, c2 = batch.load_batch(paths)
    got_w, got_i = lib.coset_min_batch(w2, o2, r2, c2)
    exp_w, exp_i = lib.coset_min_batch(words, offs, rows, cols)
    assert np.array_equal(got_w, exp_w) and np.array_equal(got_i, exp_i)

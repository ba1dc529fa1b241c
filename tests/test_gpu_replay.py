"""bmm_hip_replay (RowWindow::row_xor replay, src/core/bit_matrix_rows.cpp:
12-20) on the reference's own output: 400 ops over a 24 x 300 matrix,
including source == target no-ops, written by the reference build
(tests/golden/make_golden.py); plus the oracle on wider and edge cases."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_arrays.npz")


def test_replay_reference_golden(lib):
    g = np.load(GOLDEN)
    ops = g["replay_ops"]
    assert (ops[:, 0] == ops[:, 1]).any()          # the fixture holds no-op pairs
    rows, w = lib.replay(g["replay_in"], ops)
    assert np.array_equal(rows, g["replay_out"])
    assert w == int(g["replay_weight"][0])


@pytest.mark.parametrize("n_rows,cols,n_ops", [(256, 8192, 512), (7, 64, 0), (3, 130, 50),
                                                (64, 4096, 3000)])
def test_replay_vs_oracle(lib, ora, n_rows, cols, n_ops):
    m = lib.gen_bernoulli(900 + n_ops, n_rows, cols, 0.3)
    rng = np.random.default_rng(n_ops)
    ops = rng.integers(0, n_rows, size=(n_ops, 2)).astype(np.uint16)
    rows, w = lib.replay(m, ops)
    er, ew = ora.replay(m, cols, ops)
    assert np.array_equal(rows, er) and w == ew


def test_replay_rejects_out_of_range_rows(lib):
    m = lib.gen_bernoulli(1, 4, 128, 0.5)
    with pytest.raises(IndexError):
        lib.replay(m, np.array([[0, 4]], np.uint16))

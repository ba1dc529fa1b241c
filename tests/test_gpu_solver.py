"""GPU parity of ExactBasisSolver (the next row of the hot path: a minimum-
total-weight basis from the (weight, Gray index) candidate stream) against
the oracle, which is pinned to the reference solver (tests/test_oracle.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("t,cols,p,k", [(1, 10, 0.5, 5), (3, 64, 0.4, 7), (8, 130, 0.3, 255),
                                        (12, 200, 0.2, 4095), (16, 96, 0.5, 3000),
                                        (20, 1024, 0.1, 20000), (24, 256, 0.3, 500)])
def test_weight_order_matches_oracle(lib, ora, t, cols, p, k):
    _say("weight order", t, cols, k)
    m = ora.bernoulli(7000 + t, t, cols, p)
    gm, gw = lib.weight_order(m, cols, k)
    om, ow = ora.weight_order(m, cols, k)
    assert np.array_equal(gm, om) and np.array_equal(gw, ow)


def _matrix(bmm, m, cols):
    rows = []
    for r in m:
        bits = "".join("1" if (int(r[c // 64]) >> (c % 64)) & 1 else "0" for c in range(cols))
        rows.append(bits)
    return bmm.matrix_from_rows(rows)


@pytest.mark.parametrize("t,cols,p,mode", [(3, 3, 0.0, "kat"), (6, 70, 0.3, "rand"),
                                           (12, 150, 0.3, "deficient"), (16, 300, 0.1, "rand"),
                                           (20, 128, 0.4, "rand"), (24, 200, 0.2, "rand")])
def test_exact_basis_matches_oracle(lib, ora, t, cols, p, mode):
    import bmmpy as bmm
    _say("exact basis", t, cols, mode)
    if mode == "kat":   # the reference's own example (tests/python/test_matrix_api.py:45-65)
        mat = bmm.matrix_from_rows(["111", "110", "001"])
        import oracle
        m, cols = oracle.matrix_from_rows(["111", "110", "001"])
    else:
        m = ora.bernoulli(7100 + t, t, cols, p)
        if mode == "deficient":
            m[3] = m[7]
            m[5] = 0
            m[9] = m[1] ^ m[2] ^ m[4]
        mat = _matrix(bmm, m, cols)
    res = bmm.solver.ExactBasisSolver().solve(mat.row_window(list(range(t))))
    om, ow, ot, orank = ora.exact_basis(m, cols)
    assert res.input_rows == t and res.cols == cols
    assert res.rank == orank and res.total_weight == ot
    assert res.enumerated_states == (1 << t) - 1
    assert list(res.basis_masks) == om.tolist() and list(res.basis_weights) == ow.tolist()
    assert res.transform_matrix.shape == (orank, t)
    assert res.basis_matrix.shape == (orank, cols)
    assert res.basis_matrix.rank() == orank
    assert res.basis_matrix.weight() == ot
    for q, mask in enumerate(om):
        for r in range(t):
            assert res.transform_matrix[q, r] == bool((int(mask) >> r) & 1)
    if mode == "kat":
        assert res.rank == 2 and res.total_weight == 3 and sorted(res.basis_weights) == [1, 2]


def _gray_inverse(g):
    s = 0
    while g:
        s ^= g
        g >>= 1
    return s


@pytest.mark.parametrize("t", [29, 30])
def test_weight_order_29_30_rows_properties(lib, ora, t):
    """Windows past the oracle's enumeration range (2^29 and 2^30 states; the
    2^30 transform takes the global-pointer passes): every returned weight is
    the weight of its row combination, the stream is ordered by (weight, Gray
    index), and no sampled mask lighter than the last returned weight is
    missing from it."""
    _say("weight order", t)
    cols, k = 128, 3000
    m = ora.bernoulli(7300 + t, t, cols, 0.3)
    gm, gw = lib.weight_order(m, cols, k)
    assert gm.size == k and len(set(gm.tolist())) == k and 0 not in set(gm.tolist())

    def combo_weights(masks):
        acc = np.zeros((masks.size, m.shape[1]), np.uint64)
        for r in range(t):
            sel = ((masks >> r) & 1).astype(bool)
            acc[sel] ^= m[r]
        return np.bitwise_count(acc).sum(axis=1)

    assert np.array_equal(combo_weights(gm.astype(np.int64)), gw.astype(np.int64))
    keys = [(int(w), _gray_inverse(int(x))) for x, w in zip(gm, gw)]
    assert keys == sorted(keys)
    rng = np.random.default_rng(t)
    sample = rng.integers(1, 1 << t, 20000, dtype=np.int64)
    sw = combo_weights(sample)
    lighter = set(sample[sw < int(gw[-1])].tolist())
    assert lighter <= set(gm.tolist())

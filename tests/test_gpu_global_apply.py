"""GPU parity of the full-transform path: calc_scores_and_order (spectrum ->
stable score order), the GF(2) row combination behind BitMatrix.mul, and
GlobalGreedyApplier, against the reference's outputs (tests/golden, from the
reference build) and the oracle, which tests/test_oracle.py pins to the
reference KATs and the live reference."""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gglobal():
    return np.load(os.path.join(GOLD, "golden_global.npz"), allow_pickle=False)


# ------------------------------------------------- calc_scores_and_order ---
def test_scores_order_kats(lib):
    # fwht_tests.cpp:72-107.
    _say("scores order KATs")
    s, order, cnt, off = lib.calc_scores_and_order(np.array([0, 0, 0, 4, 4], np.int32), 4)
    assert s.tolist() == [0, 2, 2, 0, 0] and order.tolist() == [3, 4, 1, 2]
    assert cnt.tolist() == [2, 0, 2, 0, 0] and off.tolist() == [2, 2, 4, 4, 4]
    s, order, cnt, off = lib.calc_scores_and_order(np.array([0, 4, 2, 0, -2], np.int16), 4)
    assert s.dtype == np.int16 and s.tolist() == [0, 0, 1, 2, 3]
    assert order.tolist() == [1, 2, 3, 4] and cnt.tolist() == [1, 1, 1, 1, 0]


def test_scores_order_golden(lib, gglobal):
    for q in range(int(gglobal["scores_cases"][0])):
        h, n = gglobal[f"scores{q}_h"], int(gglobal[f"scores{q}_n"][0])
        _say("scores order golden", q, h.dtype, h.size, n)
        got = lib.calc_scores_and_order(h, n)
        for g, key in zip(got, ("s", "order", "cnt", "off")):
            assert np.array_equal(g, gglobal[f"scores{q}_{key}"]), (q, key)


# n around the one-pass limit (255/256), the LDS-histogram limit (16383 /
# 16384), the two/three/four-pass digit counts and uneven digit splits
# (300: 5 + 4 bits; 1023: 5 + 5), with partial last tiles.  Two-pass sorts
# pack the mask and the score bits above the first digit into one 32-bit
# word between the passes: 2^28 masks with n = 1023 widen the first digit to
# 6 bits so that 28 + 4 bits fill the word (its top bit set for scores of
# 512 and above); (1 << 22) + 5 masks with n = 65535 pack 23 + 8 bits, and
# (1 << 25) + 3 masks cannot (26 + 8 bits): separate keys and values.
# Scores of at most 10 bits (n <= 1023) take the one-pass order over tiles
# of 2^15 masks: one tile (m = 2, 3, 5000), partial last tiles, the largest
# key range (n = 1023, also in int16), and the first sizes past it (n = 1024,
# 2046) on the radix passes.
@pytest.mark.parametrize("dt,m,n", [
    (np.int32, 1, 5), (np.int32, 2, 0), (np.int32, 4097, 0), (np.int32, 65539, 255),
    (np.int32, 3, 7), (np.int32, 3 * 32768 + 17, 1023), (np.int16, 40000, 1023),
    (np.int32, 3 * 32768 + 17, 2046), (np.int16, 40000, 1500), (np.int32, 32768, 1024),
    (np.int32, 70001, 300), (np.int32, (1 << 20) + 3, 1023), (np.int16, 5000, 1),
    (np.int32, 65536, 256), (np.int32, 1 << 20, 16383), (np.int32, 1 << 20, 16384),
    (np.int32, (1 << 22) + 5, 65535), (np.int32, 1 << 20, 65536),
    (np.int32, 300000, (1 << 24) + 5), (np.int16, 1 << 20, 4000), (np.int16, 12345, 32767),
    (np.int16, 1 << 16, 60000), (np.int32, 1 << 28, 1023), (np.int32, (1 << 25) + 3, 65535)])
def test_scores_order_matches_oracle(lib, ora, dt, m, n):
    _say("scores order", dt.__name__, m, n)
    rng = np.random.default_rng(m ^ n)
    if dt == np.int16:
        lo = max(-32768, n - 2 * 32767)   # keeps every score within int16
        h = rng.integers(lo, 32768, m, dtype=np.int64).astype(dt)
    else:
        h = rng.integers(-2 * n - 7, 2 * n + 8, m, dtype=np.int64).astype(dt)
    got = lib.calc_scores_and_order(h, n)
    exp = ora.scores_order(h, n)
    for g, e, key in zip(got, exp, ("s", "order", "cnt", "off")):
        assert np.array_equal(g, e), key


def test_scores_order_full_window_spectrum(lib, ora):
    """A real window: h = FWHT of the column-pattern histogram of 20 rows
    (2^20 masks), n = active columns; the order is the masks by weight."""
    _say("scores order window spectrum")
    m = ora.bernoulli(9101, 20, 4096, 0.3)
    hist, n = ora.histogram(m, 4096)
    h = lib.fwht_i32(hist)
    got = lib.calc_scores_and_order(h, n)
    exp = ora.scores_order(h, n)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    # Scores are the row-combination weights: spot-check a few masks.
    for u in (1, 2, 3, 1 << 19, (1 << 20) - 1):
        acc = np.zeros(m.shape[1], np.uint64)
        for r in range(20):
            if (u >> r) & 1:
                acc ^= m[r]
        assert int(got[0][u]) == sum(bin(int(v)).count("1") for v in acc)


def test_scores_order_errors(lib):
    _say("scores order errors")
    with pytest.raises(ValueError, match="non-negative"):
        lib.calc_scores_and_order(np.zeros(4, np.int32), -1)
    with pytest.raises(OverflowError, match="int16"):
        lib.calc_scores_and_order(np.full(8, -30000, np.int16), 40000)
    s, order, cnt, off = lib.calc_scores_and_order(np.zeros(0, np.int32), 3)
    assert s.size == 0 and order.size == 0 and not cnt.any() and not off.any()


# --------------------------------------------------------- row combination --
def test_row_combine_matches_oracle_mul(lib, ora):
    rng = np.random.default_rng(31)
    for n_rows, cols, n_out in ((1, 64, 1), (70, 1000, 37), (130, 5000, 300), (64, 64, 64)):
        _say("row combine", n_rows, cols, n_out)
        rows = ora.bernoulli(3100 + n_rows, n_rows, cols, 0.4)
        a = ora.bernoulli(3200 + n_rows, n_out, n_rows, 0.3)
        got, wt = lib.row_combine(rows, a)
        exp = ora.mul(a, n_rows, rows, cols)
        assert np.array_equal(got, exp)
        assert wt == sum(bin(int(v)).count("1") for v in exp.reshape(-1))
    with pytest.raises(IndexError, match="beyond n_rows"):
        lib.row_combine(np.zeros((3, 1), np.uint64), np.array([[0b1000]], np.uint64))


def test_bitmatrix_mul_on_gpu(ora):
    import bmmpy as bmm
    _say("BitMatrix.mul")
    i3 = bmm.BitMatrix.identity(3)
    assert (i3 @ i3).to_rows() == i3.to_rows()
    a = ora.bernoulli(41, 50, 90, 0.3)
    b = ora.bernoulli(42, 90, 700, 0.5)
    got = bmm.BitMatrix.from_words(a, 90).mul(bmm.BitMatrix.from_words(b, 700))
    exp = ora.mul(a, 90, b, 700)
    assert np.array_equal(np.array(got.to_words(), np.uint64).reshape(50, -1)[:, :b.shape[1]],
                          exp)
    with pytest.raises(bmm.MatrixError):
        bmm.BitMatrix(2, 3).mul(bmm.BitMatrix(2, 3))


# ----------------------------------------------------- GlobalGreedyApplier --
def _candidates(bmm, masks, ws):
    return [bmm.Candidate([int(v) for v in mk], int(w)) for mk, w in zip(masks, ws)]


def test_global_greedy_reference_kat():
    # tests/python/test_matrix_api.py:66-105.
    import bmmpy as bmm
    _say("global greedy KAT")
    m = bmm.matrix_from_rows(["111", "110", "001"])
    rank = m.rank()
    cands = [bmm.Candidate([0b011], 1), bmm.Candidate([0b101], 2), bmm.Candidate([0b110], 3)]
    ap = bmm.GlobalGreedyApplier(require_improvement=True)
    assert ap.name() == "global_greedy"
    assert repr(ap) == "GlobalGreedyApplier(require_improvement=True)"
    res = ap.apply(m.row_window([0, 1, 2]), cands)
    assert res.applied_count == 1 and res.weight_improvement > 0 and m.rank() == rank
    eye = bmm.matrix_from_rows(["10", "01"])
    res = bmm.GlobalGreedyApplier().apply(eye.row_window([0, 1]), [])
    assert res.applied_count == 0 and eye.to_rows() == ["10", "01"]


def test_global_greedy_golden(gglobal):
    import bmmpy as bmm
    for c in range(int(gglobal["gg_cases"][0])):
        t, cols, req, ap, imp = (int(v) for v in gglobal[f"gg{c}_meta"])
        _say("global greedy golden", c, t, cols, req)
        m = bmm.BitMatrix.from_words(gglobal[f"gg{c}_in"], cols)
        res = bmm.GlobalGreedyApplier(require_improvement=bool(req)).apply(
            m.row_window(list(range(t))), _candidates(bmm, gglobal[f"gg{c}_masks"],
                                                      gglobal[f"gg{c}_w"]))
        assert (res.applied_count, res.weight_improvement) == (ap, imp), c
        got = np.array(m.to_words(), np.uint64).reshape(t, -1)
        exp = gglobal[f"gg{c}_out"]
        assert np.array_equal(got[:, :exp.shape[1]], exp), c


def test_global_greedy_window_of_larger_matrix(ora):
    """A window over unsorted rows of a bigger matrix: only those rows change,
    as the oracle applies the transform to the extracted rows."""
    import bmmpy as bmm
    _say("global greedy sub-window")
    S, A, M = ora.planted(77, 30, 640, 0.05)
    big = np.concatenate([ora.bernoulli(78, 10, 640, 0.5), A])
    rows = [int(r) for r in np.random.default_rng(5).permutation(np.arange(10, 40))]
    local = big[rows]
    rng = np.random.default_rng(6)
    masks = np.zeros((60, 1), np.uint64)
    for i, j in zip(*np.nonzero(rng.random((60, 30)) < 0.1)):
        masks[i, 0] |= np.uint64(1 << int(j))
    ws = rng.integers(0, 200, 60).astype(np.uint32)
    exp_rows, ap, imp = ora.global_greedy(local, 640, masks, ws, False)
    m = bmm.BitMatrix.from_words(big, 640)
    res = bmm.GlobalGreedyApplier(require_improvement=False).apply(
        m.row_window(rows), _candidates(bmm, masks, ws))
    assert (res.applied_count, res.weight_improvement) == (ap, imp)
    got = np.array(m.to_words(), np.uint64).reshape(40, -1)[:, :big.shape[1]]
    assert np.array_equal(got[rows], exp_rows) and np.array_equal(got[:10], big[:10])


def test_global_greedy_rejects_out_of_range_mask():
    import bmmpy as bmm
    _say("global greedy mask range")
    m = bmm.matrix_from_rows(["10", "01"])
    with pytest.raises(ValueError, match="outside the supported row range"):
        bmm.GlobalGreedyApplier().apply(m.row_window([0, 1]), [bmm.Candidate([0b100], 0)])
    assert m.to_rows() == ["10", "01"]


def test_oracle_is_pinned_here_too(ora, gglobal):
    """The oracle used above reproduces the fixtures (fails loudly if the
    checker itself drifts)."""
    t, cols, req, ap, imp = (int(v) for v in gglobal["gg12_meta"])
    out, a, i = ora.global_greedy(gglobal["gg12_in"], cols, gglobal["gg12_masks"],
                                  gglobal["gg12_w"], bool(req))
    assert (a, i) == (ap, imp) and np.array_equal(out, gglobal["gg12_out"])
    assert isinstance(oracle.OracleError("x"), RuntimeError)

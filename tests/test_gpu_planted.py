"""GPU parity of the wide-window coset minimum (planted meet-in-the-middle
workload: 48 x 4096 Bernoulli(0.02) mixed by a random invertible 48 x 48
GF(2) matrix), against the oracle on reduced windows and against closed-form
properties on wider ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED_PLANTED_48X4096 = 0x5EED0003


def _say(*a):
    print(*a, flush=True)


def _oracle_window(ora, A, cols, target, window):
    sub = np.concatenate([A[list(window)], A[target:target + 1]])
    w, idx = ora.coset_min(sub, cols)
    return int(w[-1]), int(idx[-1])


@pytest.mark.parametrize("width", [1, 5, 13, 14, 18, 21])
def test_window_coset_vs_oracle(lib, ora, width):
    _say("window coset", width)
    S, A, M = lib.gen_planted(SEED_PLANTED_48X4096, 48, 4096, 0.02)
    targets = [0, 17, 47]
    windows = [[r for r in range(48) if r != t][:width] for t in targets]
    w, m = lib.window_coset_min(A, 4096, targets, windows)
    for i, t in enumerate(targets):
        assert (int(w[i]), int(m[i])) == _oracle_window(ora, A, 4096, t, windows[i]), t


def test_window_coset_unordered_and_narrow(lib, ora):
    _say("window coset misc")
    m = lib.gen_bernoulli(77, 30, 300, 0.3)
    targets = [3, 29]
    windows = [[20, 1, 7, 0, 12, 25, 9, 14, 2, 28, 22, 5], [0, 3, 6, 9, 12, 15, 18, 21, 24, 27, 1, 2]]
    w, mk = lib.window_coset_min(m, 300, targets, windows)
    for i, t in enumerate(targets):
        assert (int(w[i]), int(mk[i])) == _oracle_window(ora, m, 300, t, windows[i])


def test_window_coset_duplicate_rows_reach_zero(lib):
    m = lib.gen_bernoulli(5, 20, 500, 0.3)
    m[7] = m[3] ^ m[11] ^ m[15]
    w, mk = lib.window_coset_min(m, 500, [7], [[r for r in range(20) if r != 7]])
    assert w[0] == 0
    win = [r for r in range(20) if r != 7]
    assert int(mk[0]) == (1 << win.index(3)) | (1 << win.index(11)) | (1 << win.index(15))


def test_planted_40_wide_properties(lib):
    """Full-width windows (40 other rows) on two targets: the minimum is at most
    the target's own weight (empty subset) and the returned mask reproduces
    the returned weight."""
    _say("planted 40 wide")
    S, A, M = lib.gen_planted(SEED_PLANTED_48X4096, 48, 4096, 0.02)
    targets = [0, 47]
    windows = [[r for r in range(48) if r != t][:40] for t in targets]
    w, mk = lib.window_coset_min(A, 4096, targets, windows)
    for i, t in enumerate(targets):
        acc = A[t].copy()
        for j, r in enumerate(windows[i]):
            if (int(mk[i]) >> j) & 1:
                acc ^= A[r]
        assert int(w[i]) == sum(bin(int(v)).count("1") for v in acc)
        assert int(w[i]) <= sum(bin(int(v)).count("1") for v in A[t])


def test_window_queues_mixed_in_one_call(lib, ora):
    """One call whose targets land in different kernels: a dense 16-row window
    over 40000 columns (more than 32767 active columns: int32 blocks) and a
    sparse one (packed int16 blocks), plus a 10-row window (narrow blocks)."""
    _say("window mixed queues")
    dense = lib.gen_bernoulli(0x5EED0200, 17, 40000, 0.3)
    sparse = lib.gen_bernoulli(0x5EED0201, 17, 40000, 0.001)
    A = np.concatenate([dense, sparse])
    targets = [0, 17, 5]
    windows = [list(range(1, 17)), list(range(18, 34)),
               [1, 2, 3, 4, 6, 7, 8, 9, 10, 11] + [None] * 6]
    # The C ABI takes one width per call: run the 10-row window separately.
    w, m = lib.window_coset_min(A, 40000, targets[:2], windows[:2])
    for i, t in enumerate(targets[:2]):
        assert (int(w[i]), int(m[i])) == _oracle_window(ora, A, 40000, t, windows[i]), t
    w3, m3 = lib.window_coset_min(A, 40000, [5], [windows[2][:10]])
    assert (int(w3[0]), int(m3[0])) == _oracle_window(ora, A, 40000, 5, windows[2][:10])


@pytest.mark.parametrize("width", [15, 16])
def test_window_coset_memory_lists_vs_oracle(lib, ora, width):
    """6144 active columns over 16-17 row windows: each half of the packed
    pattern list holds more than 2560 entries, so the blocks read their
    entries from memory instead of registers."""
    _say("window coset memory lists", width)
    m = lib.gen_bernoulli(4242 + width, 24, 6144, 0.5)
    targets = [0, 23]
    windows = [[r for r in range(24) if r != t][:width] for t in targets]
    w, mk = lib.window_coset_min(m, 6144, targets, windows)
    for i, t in enumerate(targets):
        assert (int(w[i]), int(mk[i])) == _oracle_window(ora, m, 6144, t, windows[i]), t


def test_window_coset_47_rows_top_bit(lib):
    """A 47-row window (33 high bits per pattern: the top one travels in the
    entry's spare address bit): the target is planted as the XOR of ten
    window rows, including the last one (mask bit 46), plus three flipped
    bits, so the unique minimum is 3 at exactly that mask.  2^47 subsets
    (about 45 s)."""
    _say("window coset 47 rows")
    rng = np.random.default_rng(47)
    m = lib.gen_bernoulli(4747, 48, 4096, 0.5)
    window = list(range(1, 48))
    pick = [2, 5, 11, 17, 23, 30, 36, 41, 44, 46]   # window positions
    t = np.zeros(m.shape[1], np.uint64)
    for k in pick:
        t ^= m[window[k]]
    for c in rng.choice(4096, 3, replace=False):
        t[c // 64] ^= np.uint64(1 << (int(c) % 64))
    m[0] = t
    w, mk = lib.window_coset_min(m, 4096, [0], [window])
    assert int(w[0]) == 3
    assert int(mk[0]) == sum(1 << k for k in pick)


def _kat40(lib, target):
    S, A, M = lib.gen_planted(SEED_PLANTED_48X4096, 48, 4096, 0.02)
    window = [r for r in range(48) if r != target][:40]
    rng = np.random.default_rng(4000 + target)
    pick = sorted({0, 39} | set(int(x) for x in rng.choice(np.arange(1, 39), 15, replace=False)))
    t = np.zeros(A.shape[1], np.uint64)
    for k in pick:
        t ^= A[window[k]]
    for c in rng.choice(4096, 3, replace=False):
        t[c // 64] ^= np.uint64(1 << (int(c) % 64))
    m = A.copy()
    m[target] = t
    w, mk = lib.window_coset_min(m, 4096, [target], [window])
    return int(w[0]), int(mk[0]), sum(1 << k for k in pick), len(pick)


@pytest.mark.parametrize("target", [0, 47])
def test_planted_40_wide_known_answer(lib, target):
    """Known answer at the workload's full width on the planted 48 x 4096
    instance: the target row is replaced by the XOR of 17 of its 40 window
    rows (the lowest-index other rows, window positions 0 and 39 included)
    plus three flipped bits.  The rows of the mixed matrix are independent,
    so any other subset adds a nonzero word of the window's row space, whose
    weight is far above 6; the unique minimum is 3 at exactly that mask
    (2^40 subsets, about 0.3 s)."""
    _say("planted 40 wide known answer", target)
    w, mk, expect, n = _kat40(lib, target)
    assert n == 17
    assert w == 3 and mk == expect

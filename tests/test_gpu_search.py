"""GPU parity of the exact top-k searches (FWHT, MITM split, brute force)
against the reference's golden outputs and the oracle."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "golden_searches.json")) as f:
        return json.load(f)


def rows_of(js):
    return np.array([[int(v) for v in r] for r in js], dtype=np.uint64)


def res_of(js):
    return [(int(w), int(m)) for w, m in js]


def gpu_search(lib, ora, kind, m, cols, k, chunk=0):
    if kind == "fwht":
        return lib.fwht_topk(m, cols, k)
    if kind == "bruteforce":
        return lib.bruteforce_topk(m, k, chunk)
    t = m.shape[0]
    high = t // 2 if t <= 26 else min(t // 2, 13)
    low = t - high
    total, q, r, mult = ora.split_window(m, cols, low)
    if q.size == 0:
        return []
    return lib.mitm_topk(q, r, mult, low, high, total, k)


def test_golden_searches(lib, ora, gold):
    for s in gold["searches"]:
        _say("golden", s["name"])
        m = rows_of(s["rows"])
        got = gpu_search(lib, ora, s["kind"], m, s["cols"], s["k"], s["chunk_bits"])
        assert got == res_of(s["result"]), s["name"]


@pytest.mark.parametrize("t,cols,p,k", [(1, 7, 0.5, 5), (5, 64, 0.3, 31), (13, 300, 0.2, 64),
                                        (14, 1000, 0.4, 128), (17, 257, 0.1, 500),
                                        (21, 96, 0.5, 1000), (22, 2048, 0.05, 4096)])
def test_fwht_topk_vs_oracle(lib, ora, t, cols, p, k):
    _say("fwht_topk", t, cols, k)
    m = ora.bernoulli(t * 31 + cols, t, cols, p)
    assert lib.fwht_topk(m, cols, k) == ora.fwht_topk(m, cols, k)


@pytest.mark.parametrize("t,cols,k", [(16, 512, 8), (16, 1024, 128), (16, 4096, 8),
                                      (20, 640, 64), (12, 100, 300)])
def test_bruteforce_topk_vs_oracle(lib, ora, t, cols, k):
    _say("bruteforce_topk", t, cols, k)
    m = ora.bernoulli(t * 7 + cols, t, cols, 0.3)
    assert lib.bruteforce_topk(m, k) == ora.bruteforce_topk(m, cols, k)


def test_bruteforce_counts_tail_bits(lib, ora):
    m = ora.bernoulli(4, 10, 70, 0.3)
    m[2, 1] |= np.uint64(1 << 50)
    assert lib.bruteforce_topk(m, 20) == ora.bruteforce_topk(m, 70, 20)


@pytest.mark.parametrize("k", [8, 128])
def test_mitm_equivalence_matrix(lib, ora, k):
    """GPU MITM == CPU MITM on the 28x96 fixture (search_cuda_tests.cpp:72-111)."""
    import oracle
    _say("mitm equiv", k)
    e = oracle.equivalence_matrix(28, 96)
    assert gpu_search(lib, ora, "mitm", e, 96, k) == ora.mitm_topk(e, 96, k)


def test_mitm_any_split_same_result(lib, ora):
    _say("mitm splits")
    m = ora.bernoulli(71, 18, 333, 0.3)
    exp = ora.fwht_topk(m, 333, 50)
    for low in (1, 5, 9, 17):
        total, q, r, mult = ora.split_window(m, 333, low)
        assert lib.mitm_topk(q, r, mult, low, 18 - low, total, 50) == exp


def test_zero_window(lib, ora):
    z = np.zeros((5, 1), np.uint64)
    assert lib.fwht_topk(z, 10, 4) == [(0, 1), (0, 2), (0, 3), (0, 4)]
    assert lib.bruteforce_topk(z, 4) == [(0, 1), (0, 2), (0, 3), (0, 4)]


def test_mitm_planted_reduced_windows(lib, ora, gold):
    S, A, M = lib.gen_planted(0x5EED0003, 48, 4096, 0.02)
    for case in gold["c3"]:
        _say("c3 reduced", case["target"])
        others, t = case["others"], case["target"]
        win = np.concatenate([A[others], A[t:t + 1]])
        assert gpu_search(lib, ora, "mitm", win, 4096, 64) == res_of(case["result"])


@pytest.mark.parametrize("kind", ["fwht", "mitm", "bruteforce"])
@pytest.mark.parametrize("t,cols,k", [(14, 300, 5000), (16, 1000, 65535), (20, 700, 123457)])
def test_large_k_full_ranking(lib, ora, kind, t, cols, k):
    """max_candidates above the on-chip top-k capacity (the reference takes
    any k): the whole (weight, mask) ranking from the HBM transform and the
    stable score order, against the oracle; k = 2^16 - 1 is every mask."""
    _say("large k", kind, t, cols, k)
    m = lib.gen_bernoulli(31 * t + k % 97, t, cols, 0.3)
    if kind == "bruteforce" and cols % 64:
        m[:, -1] |= np.uint64(1) << np.uint64(63)          # tail bits count here
    got = gpu_search(lib, ora, kind, m, cols, k)
    exp = {"fwht": ora.fwht_topk, "mitm": ora.mitm_topk,
           "bruteforce": ora.bruteforce_topk}[kind](m, cols, k)
    assert len(got) == min(k, 2 ** t - 1)
    assert got == exp


def test_large_k_every_mask_sorted_by_brute_force(lib):
    """All 2^12 - 1 masks of a 12-row window: equal to sorting (weight, mask)
    of directly computed XOR popcounts."""
    m = lib.gen_bernoulli(4242, 12, 200, 0.4)
    got = lib.fwht_topk(m, 200, 10 ** 6)
    exp = []
    for mask in range(1, 1 << 12):
        acc = np.zeros(m.shape[1], np.uint64)
        for r in range(12):
            if (mask >> r) & 1:
                acc ^= m[r]
        exp.append((int(np.unpackbits(acc.view(np.uint8)).sum()), mask))
    assert got == sorted(exp)


def test_large_k_limits_and_searcher_api(lib):
    import bmmpy as bmm
    m = lib.gen_bernoulli(5, 31, 128, 0.3)
    with pytest.raises(ValueError, match="at most 30 rows"):
        lib.fwht_topk(m, 128, 5000)
    rows = ["".join("1" if (int(r[c // 64]) >> (c % 64)) & 1 else "0" for c in range(128))
            for r in m[:18]]
    win = bmm.matrix_from_rows(rows).row_window(list(range(18)))
    a = bmm.FwhtSearch(max_rows=18, max_candidates=20000).search(win)
    b = bmm.MitmFwhtSearch(max_candidates=20000).search(win)
    c = bmm.BruteforceSearch(max_candidates=20000).search(win)
    assert len(a) == 20000
    assert [(x.weight, x.mask_u64()) for x in a] == [(x.weight, x.mask_u64()) for x in b]
    assert [(x.weight, x.mask_u64()) for x in a] == [(x.weight, x.mask_u64()) for x in c]
    assert [(x.weight, x.mask_u64()) for x in a[:4096]] == \
        [(x.weight, x.mask_u64()) for x in bmm.FwhtSearch(max_rows=18,
                                                            max_candidates=4096).search(win)]


def _sampled_bound_window(lib, case):
    """Windows of 22-28 rows, where the packed search first bounds the keys
    from a sample of blocks (each workgroup sees only a few blocks)."""
    if case == "random24":
        return lib.gen_bernoulli(0x5EED0021, 24, 200, 0.3), 200
    if case == "dup_rows23":
        # Ten distinct rows and thirteen copies of them: most masks tie.
        base = lib.gen_bernoulli(0x5EED0022, 10, 160, 0.4)
        return np.concatenate([base, base[np.arange(13) % 10]]), 160
    if case == "few_cols22":
        # Five active columns: 2^17 - 1 masks of weight 0, the list fills
        # with them and every crowded block is rescanned exactly.
        m = lib.gen_bernoulli(0x5EED0023, 22, 5, 0.5)
        return m, 5
    if case == "sparse28":
        return lib.gen_bernoulli(0x5EED0024, 28, 700, 0.05), 700
    raise ValueError(case)


@pytest.mark.parametrize("case,k", [("random24", 1), ("random24", 1000), ("dup_rows23", 300),
                                    ("few_cols22", 128), ("few_cols22", 4096),
                                    ("sparse28", 64), ("sparse28", 513)])
def test_sampled_bound_windows(lib, ora, case, k):
    """The sampled key bound never drops a key of the exact top-k: FWHT and
    MITM searches equal the oracle, ties at the bound included."""
    _say("sampled bound", case, k)
    m, cols = _sampled_bound_window(lib, case)
    exp = ora.fwht_topk(m, cols, k)
    assert lib.fwht_topk(m, cols, k) == exp
    assert gpu_search(lib, ora, "mitm", m, cols, k) == exp

"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker:
against the reference's golden outputs (tests/golden, written from the
reference's own compiled CPU code), against the reference library directly
where it is built here, and against closed forms / brute force."""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def arrays():
    return np.load(os.path.join(GOLD, "golden_arrays.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "golden_searches.json")) as f:
        return json.load(f)


def rows_of(js):
    return np.array([[int(v) for v in r] for r in js], dtype=np.uint64)


def res_of(js):
    return [(int(w), int(m)) for w, m in js]


# ---------------------------------------------------------------- RNG ----
def test_xorshift_recurrence(ora):
    M = (1 << 64) - 1
    x = 0x5EED0002
    exp = []
    for _ in range(8):
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        exp.append(x)
    assert ora.xs_stream(0x5EED0002, 8).tolist() == exp
    # seed 0 -> default state (xorshift64.hpp:8,15-16)
    assert ora.xs_stream(0, 1)[0] == ora.xs_stream(0x9E3779B97F4A7C15, 1)[0]


# --------------------------------------------------------------- FWHT ----
def test_fwht_reference_kats(ora):
    assert ora.fwht_i32([1, 2, 3, 4]).tolist() == [10, -2, -4, 0]   # fwht_tests.cpp:58-62
    assert ora.fwht_i16([32767, 1]).tolist() == [-32768, 32766]     # fwht_tests.cpp:64-70


def test_fwht_golden(ora, arrays):
    for nb in (1, 4, 10, 14):
        assert np.array_equal(ora.fwht_i32(arrays[f"fwht_i32_{nb}_in"]),
                              arrays[f"fwht_i32_{nb}_out"])
        assert np.array_equal(ora.fwht_i16(arrays[f"fwht_i16_{nb}_in"]),
                              arrays[f"fwht_i16_{nb}_out"])


def test_fwht_closed_forms(ora):
    rng = np.random.default_rng(3)
    n = 1 << 9
    x = rng.integers(-50, 50, n).astype(np.int32)
    h = ora.fwht_i32(x)
    # brute-force Walsh sum H[m] = sum_p x[p] (-1)^popcount(p & m)
    idx = np.arange(n)
    par = np.array([[bin(int(p & m)).count("1") & 1 for p in idx] for m in idx[:64]])
    assert np.array_equal(h[:64], ((1 - 2 * par) * x[None, :]).sum(1))
    assert np.array_equal(ora.fwht_i32(h), n * x)                    # involution
    assert int((h.astype(np.int64) ** 2).sum()) == n * int((x.astype(np.int64) ** 2).sum())
    d = np.zeros(n, np.int32)
    d[37] = 1                                                         # delta -> sign row
    assert np.array_equal(ora.fwht_i32(d), 1 - 2 * par_row(37, n))
    assert h[0] == x.sum()


def par_row(p, n):
    return np.array([bin(p & m).count("1") & 1 for m in range(n)], dtype=np.int32)


def test_fwht_rejects_non_power_of_two(ora):
    with pytest.raises(oracle.OracleError):
        ora.fwht_i32([1, 2, 3])


# ----------------------------------------------------------- histogram ---
def test_weight_identity_every_mask(ora):
    """(total - H[m]) / 2 equals popcount(XOR of rows in m) for every m."""
    m = ora.bernoulli(11, 10, 150, 0.35)
    b, total = ora.histogram(m, 150)
    h = ora.fwht_i32(b)
    assert b.sum() == total
    for mask in range(1 << 10):
        acc = np.zeros(m.shape[1], np.uint64)
        for r in range(10):
            if mask >> r & 1:
                acc ^= m[r]
        direct = sum(bin(int(v)).count("1") for v in acc)
        assert (total - int(h[mask])) // 2 == direct


# -------------------------------------------------------------- searches --
def _run(ora, s):
    m = rows_of(s["rows"])
    if s["kind"] == "fwht":
        return ora.fwht_topk(m, s["cols"], s["k"])
    if s["kind"] == "mitm":
        return ora.mitm_topk(m, s["cols"], s["k"])
    return ora.bruteforce_topk(m, s["cols"], s["k"], s["chunk_bits"])


def test_searches_match_reference_golden(ora, gold):
    for s in gold["searches"]:
        assert _run(ora, s) == res_of(s["result"]), s["name"]


def test_search_kats():
    """search_cpu_tests.cpp:18-52: 2x5 window -> best mask 0x3 weight 0, 3 results."""
    o = oracle.oracle()
    m = np.array([[0b10101], [0b10101]], dtype=np.uint64)
    res = o.fwht_topk(m, 5, 64)
    assert len(res) == 3 and res[0] == (0, 3)
    assert o.fwht_topk(m, 5, 1) == [(0, 3)]
    assert o.bruteforce_topk(m, 5, 1) == [(0, 3)]


def test_topk_equals_sorted_brute_force(ora):
    for seed in range(4):
        t, cols = 8 + seed, 90 + 17 * seed
        m = ora.bernoulli(seed + 50, t, cols, 0.3)
        allw = []
        for mask in range(1, 1 << t):
            acc = np.zeros(m.shape[1], np.uint64)
            for r in range(t):
                if mask >> r & 1:
                    acc ^= m[r]
            allw.append((sum(bin(int(v)).count("1") for v in acc), mask))
        allw.sort()
        for k in (1, 7, 60):
            exp = allw[:k]
            assert ora.fwht_topk(m, cols, k) == exp
            assert ora.mitm_topk(m, cols, k) == exp
            assert ora.bruteforce_topk(m, cols, k) == exp


def test_bruteforce_counts_tail_bits_fwht_does_not(ora, gold):
    by = {s["name"]: s for s in gold["searches"]}
    b = res_of(by["tail_12x70_bruteforce_k16"]["result"])
    f = res_of(by["tail_12x70_fwht_k16"]["result"])
    assert b != f   # the planted bit beyond cols changes brute-force weights only


def test_split_window_golden(ora, gold):
    for s in gold["splits"]:
        tot, q, r, mult = ora.split_window(rows_of(s["rows"]), s["cols"], s["low_bits"])
        assert tot == s["total"], s["name"]
        assert q.tolist() == s["q"] and r.tolist() == s["r"] and mult.tolist() == s["mult"]


# ------------------------------------------------------------- configs ----
def test_coset_c1_golden(ora, arrays):
    m = arrays["c1_rows"]
    w, idx = ora.coset_min(m, 128)
    assert np.array_equal(w, arrays["c1_weight"])
    for i in range(20):
        mk = int(arrays["c1_mask"][i])
        below = mk & ((1 << i) - 1)
        above = mk >> (i + 1)
        assert int(idx[i]) == below | (above << i)


def test_coset_matches_brute_force_small(ora):
    m = ora.bernoulli(9, 9, 77, 0.45)
    w, idx = ora.coset_min(m, 77)
    for t in range(9):
        others = [r for r in range(9) if r != t]
        best = None
        for s in range(1 << 8):
            acc = m[t].copy()
            for j, r in enumerate(others):
                if s >> j & 1:
                    acc ^= m[r]
            wt = sum(bin(int(v)).count("1") for v in acc)
            if best is None or wt < best[0]:
                best = (wt, s)
        assert (int(w[t]), int(idx[t])) == best


def test_gray_c2_golden(ora, arrays):
    m = arrays["c2_rows"]
    for i, t in enumerate(arrays["c2_targets"]):
        w, s, low = ora.gray_min(m, 24, int(t))
        assert w == arrays["c2_weight"][i]
        assert s == arrays["c2_first"][i]
        assert low == arrays["c2_mask"][i]


def test_gray_first_step_small_brute_force(ora):
    m = ora.bernoulli(19, 12, 64, 0.5)
    for t in (0, 4, 11):
        blk = [r for r in range(10) if r != t]
        seq = []
        for s in range(1 << len(blk)):
            g = s ^ (s >> 1)
            acc = m[t].copy()
            for j, r in enumerate(blk):
                if g >> j & 1:
                    acc ^= m[r]
            seq.append((sum(bin(int(v)).count("1") for v in acc), s, g))
        wmin = min(x[0] for x in seq)
        first = min(x[1] for x in seq if x[0] == wmin)
        low = min(x[2] for x in seq if x[0] == wmin)
        assert ora.gray_min(m, 10, t) == (wmin, first, low)


def test_c3_reduced_windows_golden(ora, gold):
    S, A, M = ora.planted(0x5EED0003, 48, 4096, 0.02)
    assert ora.rank(M, 48) == 48
    for case in gold["c3"]:
        others, t = case["others"], case["target"]
        win = np.concatenate([A[others], A[t:t + 1]])
        assert ora.mitm_topk(win, 4096, 64) == res_of(case["result"])


def test_c4_small_matrices_golden(ora, gold):
    rows, cols, dens, words, offs = ora.batch(0x5EED0004, 64)
    for case in gold["c4"]:
        mi = case["matrix"]
        W = cols[mi] // 64
        mat = words[offs[mi]: offs[mi] + rows[mi] * W].reshape(rows[mi], W)
        w, idx = ora.coset_min(mat, int(cols[mi]))
        assert w.tolist() == case["weight"], mi
        for i, mk in enumerate(case["mask"]):
            mk = int(mk)
            assert int(idx[i]) == (mk & ((1 << i) - 1)) | ((mk >> (i + 1)) << i)


def test_replay_golden(ora, arrays):
    out, wt = ora.replay(arrays["replay_in"], 300, arrays["replay_ops"])
    assert np.array_equal(out, arrays["replay_out"])
    assert wt == int(arrays["replay_weight"][0])


def test_replay_identity_and_noop(ora):
    a = ora.bernoulli(3, 2, 200, 0.5)
    pop = lambda r: sum(bin(int(v)).count("1") for v in r)   # noqa: E731
    out, _ = ora.replay(a, 200, np.array([[0, 1]], np.uint16))
    assert pop(out[1]) == pop(a[0]) + pop(a[1]) - 2 * pop(a[0] & a[1])
    same, _ = ora.replay(a, 200, np.array([[1, 1]], np.uint16))
    assert np.array_equal(same, a)


def test_greedy_apply_golden(ora, gold):
    g = gold["greedy"]
    rows = rows_of(g["rows"])
    cands = res_of(g["candidates"])
    for case in g["cases"]:
        out, ap, imp = ora.greedy_apply(rows, g["cols"], cands, case["min_gain"],
                                        case["stochastic"], case["seed"])
        assert (ap, imp) == (case["applied"], case["improvement"])
        assert np.array_equal(out, rows_of(case["rows"]))


# ------------------------------------------------- live reference (here) --
def test_oracle_matches_live_reference_random(ora):
    r = oracle.reference()
    if r is None:
        pytest.skip("reference CPU library not built here")
    for seed in range(6):
        t = 6 + 3 * seed
        cols = 50 + 61 * seed
        m = ora.bernoulli(900 + seed, t, cols, 0.2 + 0.05 * seed)
        for k in (1, 33):
            assert ora.fwht_topk(m, cols, k) == r.fwht_search(m, cols, k)
            assert ora.mitm_topk(m, cols, k) == r.mitm_search(m, cols, k)
            if t <= 18:
                assert ora.bruteforce_topk(m, cols, k, 0) == r.bruteforce_search(m, cols, k, 0)
                assert ora.bruteforce_topk(m, cols, k, 3) == r.bruteforce_search(m, cols, k, 3)


def test_planted_generator_properties(ora):
    S, A, M = ora.planted(0x5EED0005, 64, 512, 0.05)
    assert ora.rank(M, 64) == 64
    assert ora.rank(A, 512) == ora.rank(S, 512)
    # 256 x 256 transforms need nonlinear bit extraction (XorShift64 is
    # GF(2)-linear); the planted 256 x 8192 instance must be full rank.
    S, A, M = ora.planted(0x5EED0005, 256, 8192, 0.01)
    assert ora.rank(M, 256) == 256
    assert ora.rank(A, 8192) == ora.rank(S, 8192)


def test_linear_bits_cap_rank():
    """Why the transform bits are mixed: p = 0.5 Bernoulli bits read off
    XorShift64 leave a 100 x 100 matrix far from full rank."""
    o = oracle.oracle()
    assert o.rank(o.bernoulli(5, 100, 100, 0.5), 100) <= 66


# ------------------------------------------------------- exact basis -----
def _ginv(m):
    g = 0
    while m:
        g ^= m
        m >>= 1
    return g


def test_exact_basis_kat(ora):
    """tests/python/test_matrix_api.py:45-65 of the reference: rows 111, 110,
    001 -> rank 2, total weight 3, basis weights {1, 2}."""
    m, cols = oracle.matrix_from_rows(["111", "110", "001"])
    masks, ws, total, rank = ora.exact_basis(m, cols)
    assert rank == 2 and total == 3 and sorted(ws.tolist()) == [1, 2]
    r = oracle.reference(scalar=True)
    if r is not None:
        rm, rw, rt, rr, rb = r.exact_basis(m, cols)
        assert (rr, rt) == (2, 3)
        assert rm.tolist() == masks.tolist() and rw.tolist() == ws.tolist()


def test_weight_order_is_weight_then_gray_index(ora):
    """The candidate stream against a direct sort of every nonzero mask by
    (weight, inverse Gray index), weights by explicit XOR + popcount."""
    for seed, (t, cols, p) in enumerate([(5, 70, 0.4), (9, 130, 0.2), (12, 64, 0.5)]):
        m = ora.bernoulli(4000 + seed, t, cols, p)
        keys = []
        for mask in range(1, 1 << t):
            acc = np.zeros(m.shape[1], np.uint64)
            for r in range(t):
                if (mask >> r) & 1:
                    acc ^= m[r]
            keys.append((sum(bin(int(v)).count("1") for v in acc), _ginv(mask), mask))
        keys.sort()
        k = min(len(keys), 300)
        masks, ws = ora.weight_order(m, cols, k)
        assert masks.tolist() == [x[2] for x in keys[:k]]
        assert ws.tolist() == [x[0] for x in keys[:k]]


def test_exact_basis_matches_live_reference(ora):
    """Random, rank-deficient (duplicate / zero / dependent rows) and tail-bit
    windows against the reference ExactBasisSolver compiled here (its scalar
    build: the AVX2 build of this solver crashes on rows of >= 4 words)."""
    r = oracle.reference(scalar=True)
    if r is None:
        pytest.skip("reference CPU library not built here")
    cases = []
    for seed, (t, cols, p) in enumerate([(1, 10, 0.5), (4, 65, 0.3), (10, 200, 0.1),
                                         (14, 96, 0.5), (18, 300, 0.05)]):
        cases.append((ora.bernoulli(5000 + seed, t, cols, p), cols))
    m = ora.bernoulli(5100, 12, 150, 0.3)
    m[3] = m[7]
    m[5] = 0
    m[9] = m[1] ^ m[2] ^ m[4]
    cases.append((m, 150))
    for m, cols in cases:
        masks, ws, total, rank = ora.exact_basis(m, cols)
        rm, rw, rt, rr, rb = r.exact_basis(m, cols)
        assert rank == rr and total == rt
        assert masks.tolist() == rm.tolist() and ws.tolist() == rw.tolist()
        # basis rows = transform x window rows
        for q, mask in enumerate(masks):
            acc = np.zeros(m.shape[1], np.uint64)
            for row in range(m.shape[0]):
                if (int(mask) >> row) & 1:
                    acc ^= m[row]
            assert np.array_equal(acc, rb[q])


# ------------------------------------------ scores / global greedy apply ---
@pytest.fixture(scope="module")
def gglobal():
    return np.load(os.path.join(GOLD, "golden_global.npz"), allow_pickle=False)


def test_scores_order_kats(ora):
    # fwht_tests.cpp:72-107 (off: the end of each bucket after the scatter).
    s, order, cnt, off = ora.scores_order(np.array([0, 0, 0, 4, 4], np.int32), 4)
    assert s.tolist() == [0, 2, 2, 0, 0] and order.tolist() == [3, 4, 1, 2]
    assert cnt.tolist() == [2, 0, 2, 0, 0] and off.tolist() == [2, 2, 4, 4, 4]
    s, order, cnt, off = ora.scores_order(np.array([0, 4, 2, 0, -2], np.int16), 4)
    assert s.dtype == np.int16 and s.tolist() == [0, 0, 1, 2, 3]
    assert order.tolist() == [1, 2, 3, 4] and cnt.tolist() == [1, 1, 1, 1, 0]


def test_scores_order_golden(ora, gglobal):
    for q in range(int(gglobal["scores_cases"][0])):
        h, n = gglobal[f"scores{q}_h"], int(gglobal[f"scores{q}_n"][0])
        s, order, cnt, off = ora.scores_order(h, n)
        for got, key in ((s, "s"), (order, "order"), (cnt, "cnt"), (off, "off")):
            assert np.array_equal(got, gglobal[f"scores{q}_{key}"]), (q, key)


@pytest.mark.parametrize("dt", [np.int32, np.int16])
def test_scores_order_is_a_stable_sort(ora, dt):
    """Closed forms: score = clamp(floor((n - h) / 2), 0, n); order = numpy's
    stable argsort of the scores of masks 1..m-1; cnt = bincount; off =
    cumulative cnt.  Covers m = 0 and 1 and n = 0."""
    rng = np.random.default_rng(21)
    for m, n in ((0, 3), (1, 5), (2, 0), (37, 0), (1000, 7), (4096, 300), (3000, 2000)):
        lim = min(3 * n + 5, 32767) if dt == np.int16 else 3 * n + 5
        h = rng.integers(-lim, lim + 1, m, dtype=np.int64).astype(dt)
        s, order, cnt, off = ora.scores_order(h, n)
        exp = np.clip(np.floor_divide(n - h.astype(np.int64), 2), 0, n)
        if m:
            exp[0] = 0
        assert np.array_equal(s.astype(np.int64), exp)
        assert np.array_equal(order, np.argsort(exp[1:], kind="stable") + 1)
        assert np.array_equal(cnt, np.bincount(exp[1:], minlength=n + 1))
        assert np.array_equal(off, np.cumsum(cnt) if m else np.zeros(n + 1, np.int32))


def test_scores_order_matches_live_reference(ora, ref):
    rng = np.random.default_rng(22)
    for dt, m, n in ((np.int32, 5000, 999), (np.int32, 1 << 15, 70000), (np.int16, 4096, 1200),
                     (np.int16, 777, 0)):
        lim = min(2 * n + 9, 32767)
        h = rng.integers(-lim, lim + 1, m, dtype=np.int64).astype(dt)
        for a, b in zip(ora.scores_order(h, n), ref.scores_order(h, n)):
            assert np.array_equal(a, b)


def test_scores_order_rejects_negative_n(ora, ref):
    with pytest.raises(oracle.OracleError):
        ora.scores_order(np.zeros(4, np.int32), -1)
    with pytest.raises(oracle.RefError, match="non-negative"):
        ref.scores_order(np.zeros(4, np.int32), -1)


def test_global_greedy_kat(ora):
    # tests/python/test_matrix_api.py:66-105.
    m, cols = oracle.matrix_from_rows(["111", "110", "001"])
    out, ap, imp = ora.global_greedy(m, cols, [[0b011], [0b101], [0b110]], [1, 2, 3])
    assert ap == 1 and imp > 0
    assert ora.rank(out, cols) == ora.rank(m, cols)
    eye, cols = oracle.matrix_from_rows(["10", "01"])
    out, ap, imp = ora.global_greedy(eye, cols, np.zeros((0, 1), np.uint64), [])
    assert (ap, imp) == (0, 0) and np.array_equal(out, eye)


def test_global_greedy_golden(ora, gglobal):
    for c in range(int(gglobal["gg_cases"][0])):
        t, cols, req, ap, imp = (int(v) for v in gglobal[f"gg{c}_meta"])
        out, a, i = ora.global_greedy(gglobal[f"gg{c}_in"], cols, gglobal[f"gg{c}_masks"],
                                      gglobal[f"gg{c}_w"], bool(req))
        assert (a, i) == (ap, imp), c
        assert np.array_equal(out, gglobal[f"gg{c}_out"]), c


def test_global_greedy_matches_live_reference(ora, ref):
    rng = np.random.default_rng(23)
    for t, cols in ((1, 64), (5, 70), (33, 130), (65, 64), (128, 200)):
        mat = ora.bernoulli(90 + t, t, cols, 0.3)
        mw = (t + 63) // 64
        masks = np.zeros((2 * t, mw), np.uint64)
        for i, j in zip(*np.nonzero(rng.random((2 * t, t)) < 2.5 / t)):
            masks[i, j // 64] |= np.uint64(1 << (j % 64))
        ws = rng.integers(0, cols // 3 + 1, 2 * t).astype(np.uint32)
        for req in (True, False):
            a = ora.global_greedy(mat, cols, masks, ws, req)
            b = ref.global_greedy(mat, cols, masks, ws, req)
            assert a[1:] == b[1:] and np.array_equal(a[0], b[0]), (t, req)


def test_global_greedy_rejects_out_of_range_masks(ora, ref):
    m, cols = oracle.matrix_from_rows(["10", "01"])
    for masks in ([[0b100]], [[0b1, 0b1]]):
        with pytest.raises(oracle.OracleError):
            ora.global_greedy(m, cols, masks, [0])
        with pytest.raises(oracle.RefError, match="outside the supported row range"):
            ref.global_greedy(m, cols, masks, [0])


def test_mul_and_rank_match_live_reference(ora, ref):
    """or_mul / or_rank (the checkers of the GPU row combination) against the
    reference BitMatrix::mul and rank, plus identity and associativity."""
    for ar, ac, bc, p in ((1, 1, 1, 0.5), (7, 70, 130, 0.3), (64, 64, 64, 0.5), (33, 200, 65, 0.1)):
        a = ora.bernoulli(700 + ar, ar, ac, p)
        b = ora.bernoulli(800 + bc, ac, bc, p)
        got = ora.mul(a, ac, b, bc)
        exp, rank = ref.mul_rank(a, ac, b, bc)
        assert np.array_equal(got, exp)
        assert ora.rank(a, ac) == rank
    n = 40
    x = ora.bernoulli(1, n, n, 0.5)
    y = ora.bernoulli(2, n, n, 0.5)
    z = ora.bernoulli(3, n, 90, 0.5)
    eye = np.zeros((n, 1), np.uint64)
    for i in range(n):
        eye[i, 0] = np.uint64(1 << i)
    assert np.array_equal(ora.mul(eye, n, z, 90), z)
    assert np.array_equal(ora.mul(ora.mul(x, n, y, n), n, z, 90),
                          ora.mul(x, n, ora.mul(y, n, z, 90), 90))

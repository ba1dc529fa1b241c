"""Writes the golden fixtures in this directory from the reference's own CPU
code (oracle/_ref/libbmmref.so, compiled from the reference sources by
`make -C oracle ref`).  Run here, where the reference tree exists:

    python tests/golden/make_golden.py

Every fixture is data: inputs (packed uint64 rows) and the reference outputs.
Inputs come from this project's seeded generators (XorShift64) or from the
reference test fixtures restated in oracle/__init__.py.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

SEED_COSET_20X128, SEED_GRAY_32X1024, SEED_PLANTED_48X4096, SEED_BATCH_4096 = 0x5EED0001, 0x5EED0002, 0x5EED0003, 0x5EED0004


def ginv(g: int) -> int:
    s = 0
    while g:
        s ^= g
        g >>= 1
    return s


def first_containing(ws, masks, bit):
    sel = (masks >> np.uint64(bit)) & np.uint64(1)
    i = int(np.argmax(sel == 1))
    return int(ws[i]), int(masks[i])


def main():
    ref = oracle.reference()
    ora = oracle.oracle()
    assert ref is not None, "build oracle/_ref first (make -C oracle ref)"
    out = {}

    # --- FWHT (src/math/fwht.cpp) -------------------------------------------
    rng = np.random.default_rng(1)
    for nb in (1, 4, 10, 14):
        x = rng.integers(-2**31, 2**31, 1 << nb, dtype=np.int64).astype(np.int32)
        out[f"fwht_i32_{nb}_in"] = x
        out[f"fwht_i32_{nb}_out"] = ref.fwht_i32(x)
        y = rng.integers(-2**15, 2**15, 1 << nb, dtype=np.int64).astype(np.int16)
        out[f"fwht_i16_{nb}_in"] = y
        out[f"fwht_i16_{nb}_out"] = ref.fwht_i16(y)

    searches = []  # (name, rows, cols, kind, k, extra) -> list of (w, mask)

    def add_search(name, m, cols, kind, k, chunk=0):
        if kind == "fwht":
            res = ref.fwht_search(m, cols, k, max_rows=64)
        elif kind == "mitm":
            res = ref.mitm_search(m, cols, k)
        else:
            res = ref.bruteforce_search(m, cols, k, chunk)
        searches.append({"name": name, "kind": kind, "k": k, "chunk_bits": chunk, "cols": cols,
                         "rows": [[int(v) for v in r] for r in m],
                         "result": [[w, str(mk)] for w, mk in res]})

    # Reference KAT windows (tests/cpp/search_cpu_tests.cpp).
    m25 = np.zeros((2, 1), np.uint64)
    for c in (0, 2, 4):
        m25[0, 0] |= np.uint64(1 << c)
        m25[1, 0] |= np.uint64(1 << c)
    for kind in ("fwht", "mitm", "bruteforce"):
        add_search(f"kat_2x5_{kind}_k64", m25, 5, kind, 64)
        add_search(f"kat_2x5_{kind}_k1", m25, 5, kind, 1)
    m612, c612 = oracle.matrix_from_rows(["110010101001", "101100101100", "011010011001",
                                          "111000010111", "000111100011", "101011110000"])
    m35, c35 = oracle.matrix_from_rows(["10101", "11100", "00111"])
    for kind in ("fwht", "mitm", "bruteforce"):
        add_search(f"kat_6x12_{kind}_k8", m612, c612, kind, 8)
        add_search(f"kat_3x5_{kind}_k4", m35, c35, kind, 4)
    # GPU equivalence fixtures (tests/cpp/search_cuda_tests.cpp:72-197).
    eq = oracle.equivalence_matrix(28, 96)
    add_search("equiv_28x96_mitm_k8", eq, 96, "mitm", 8)
    add_search("equiv_28x96_mitm_k128", eq, 96, "mitm", 128)
    add_search("equiv_28x96_sub18_fwht_k128", eq[:18], 96, "fwht", 128)
    for cols, k in ((512, 8), (1024, 128), (4096, 8)):
        e = oracle.equivalence_matrix(16, cols)
        add_search(f"equiv_16x{cols}_bruteforce_k{k}", e, cols, "bruteforce", k)
    # Tail bits beyond cols: brute force counts them, FWHT masks them.
    tail = oracle.equivalence_matrix(12, 70)
    tail[3, 1] |= np.uint64(1 << 40)   # column 104 > cols
    add_search("tail_12x70_bruteforce_k16", tail, 70, "bruteforce", 16)
    add_search("tail_12x70_fwht_k16", tail, 70, "fwht", 16)
    add_search("tail_12x70_mitm_k16", tail, 70, "mitm", 16)
    # Random windows of assorted widths for all three searchers.
    for i, (t, cols, p, k) in enumerate([(9, 40, 0.5, 20), (14, 200, 0.3, 64),
                                         (20, 130, 0.4, 100), (24, 333, 0.25, 50),
                                         (30, 64, 0.5, 40)]):
        m = ora.bernoulli(1000 + i, t, cols, p)
        if t <= 24:
            add_search(f"rand_{t}x{cols}_fwht_k{k}", m, cols, "fwht", k)
        add_search(f"rand_{t}x{cols}_mitm_k{k}", m, cols, "mitm", k)
        if t <= 24:
            add_search(f"rand_{t}x{cols}_bruteforce_k{k}", m, cols, "bruteforce", k, 0)
    # All-zero window: MITM returns nothing, FWHT/brute return weight-0 masks.
    z = np.zeros((5, 1), np.uint64)
    for kind in ("fwht", "mitm", "bruteforce"):
        add_search(f"zero_5x10_{kind}_k4", z, 10, kind, 4)

    # --- compact split window (src/search/split_window_prep.cpp) ------------
    splits = []
    m, c = oracle.matrix_from_rows(["11001", "00100", "01101", "01001"])
    for nm, mm, cc, low in (("kat_4x5_low2", m, c, 2), ("equiv_28x96_low14", eq, 96, 14),
                            ("rand_20x500_low7", ora.bernoulli(77, 20, 500, 0.1), 500, 7)):
        tot, q, r, mult = ref.compact_split(mm, cc, low)
        splits.append({"name": nm, "cols": cc, "low_bits": low,
                       "rows": [[int(v) for v in rr] for rr in mm], "total": tot,
                       "q": [int(v) for v in q], "r": [int(v) for v in r],
                       "mult": [int(v) for v in mult]})

    # --- 20x128 coset minimum from the full reference ranking ---------------
    c1 = ora.bernoulli(SEED_COSET_20X128, 20, 128, 0.5)
    ws, masks = ref.fwht_search_arrays(c1, 128, (1 << 20) - 1)
    c1_res = [first_containing(ws, masks, i) for i in range(20)]
    out["c1_rows"] = c1
    out["c1_weight"] = np.array([w for w, _ in c1_res], np.uint32)
    out["c1_mask"] = np.array([mk for _, mk in c1_res], np.uint64)

    # --- Gray sweep minima for four targets via the reference ranking -------
    c2 = ora.bernoulli(SEED_GRAY_32X1024, 32, 1024, 0.3)
    c2_t, c2_w, c2_s, c2_m = [], [], [], []
    for target in (0, 5, 24, 31):
        blk = [r for r in range(24) if r != target]
        win = np.concatenate([c2[blk], c2[target:target + 1]])
        b = len(blk)
        ws, masks = ref.fwht_search_arrays(win, 1024, (1 << (b + 1)) - 1)
        sel = ((masks >> np.uint64(b)) & np.uint64(1)) == 1
        wmin = int(ws[sel].min())
        mins = masks[sel & (ws == wmin)] & np.uint64((1 << b) - 1)
        c2_t.append(target)
        c2_w.append(wmin)
        c2_m.append(int(mins.min()))
        c2_s.append(min(ginv(int(x)) for x in mins))
        print("c2 target", target, wmin, flush=True)
    out["c2_rows"] = c2
    out["c2_targets"] = np.array(c2_t, np.int32)
    out["c2_weight"] = np.array(c2_w, np.uint32)
    out["c2_first"] = np.array(c2_s, np.uint64)
    out["c2_mask"] = np.array(c2_m, np.uint64)

    # --- planted instance, reduced windows through reference MITM ----------
    S, A, M = ora.planted(SEED_PLANTED_48X4096, 48, 4096, 0.02)
    out["c3_sparse_weights"] = np.array([bin(int(x)).count("1") for r in S for x in r],
                                        np.uint32).reshape(48, -1).sum(1).astype(np.uint32)
    out["c3_transform"] = M
    c3_cases = []
    for target, width in ((0, 20), (17, 22), (47, 24)):
        others = [r for r in range(48) if r != target][:width - 1]
        win = np.concatenate([A[others], A[target:target + 1]])
        res = ref.mitm_search(win, 4096, 64)
        c3_cases.append({"target": target, "others": others,
                         "result": [[w, str(mk)] for w, mk in res]})

    # --- batched uneven workload, small matrices via reference ranking ------
    rows, cols, dens, words, offs = ora.batch(SEED_BATCH_4096, 64)
    c4 = []
    for mi in range(64):
        if rows[mi] > 16:
            continue
        W = cols[mi] // 64
        mat = words[offs[mi]: offs[mi] + rows[mi] * W].reshape(rows[mi], W)
        ws, masks = ref.fwht_search_arrays(mat, int(cols[mi]), (1 << int(rows[mi])) - 1)
        res = [first_containing(ws, masks, i) for i in range(int(rows[mi]))]
        c4.append({"matrix": mi, "weight": [w for w, _ in res],
                   "mask": [str(mk) for _, mk in res]})

    # --- replay / greedy apply (src/core/bit_matrix_rows.cpp, apply/) --------
    rng = np.random.default_rng(5)
    rep = ora.bernoulli(31, 24, 300, 0.2)
    ops = rng.integers(0, 24, size=(400, 2), dtype=np.int64).astype(np.uint16)
    rep_out, rep_w = ref.replay(rep, 300, ops)
    out["replay_in"], out["replay_ops"], out["replay_out"] = rep, ops, rep_out
    out["replay_weight"] = np.array([rep_w], np.int64)
    ga_in = ora.bernoulli(32, 12, 256, 0.3)
    cands = ref.fwht_search(ga_in, 256, 40)
    ga_cases = []
    for mg, st, sd in ((0, False, 0), (1, False, 0), (3, True, 7), (0, True, 0)):
        g_out, ap, imp = ref.greedy_apply(ga_in, 256, cands, mg, st, sd)
        ga_cases.append({"min_gain": mg, "stochastic": st, "seed": sd, "applied": ap,
                         "improvement": imp, "rows": [[str(int(v)) for v in r] for r in g_out]})

    np.savez_compressed(os.path.join(HERE, "golden_arrays.npz"), **out)
    with open(os.path.join(HERE, "golden_searches.json"), "w") as f:
        json.dump({"searches": searches, "splits": splits, "c3": c3_cases, "c4": c4,
                   "greedy": {"rows": [[str(int(v)) for v in r] for r in ga_in],
                              "cols": 256, "candidates": [[w, str(mk)] for w, mk in cands],
                              "cases": ga_cases}}, f)
    # --- SASelector (src/search/sa_selector*.cpp; the scalar reference build) ---
    refs = oracle.reference(scalar=True)
    sa_cases = []
    doc, doc_cols = oracle.matrix_from_rows(["111111000000"] * 4 + ["000000000000"] * 2)
    specs = [(doc, doc_cols, 4, 256, 4, 7), (doc, doc_cols, 0, 256, 4, 7),
             (doc, doc_cols, 6, 100, 2, 3),
             (ora.bernoulli(41, 40, 300, 0.2), 300, 8, 2000, 3, 11),
             (ora.bernoulli(42, 100, 512, 0.1), 512, 16, 5000, 2, 0),
             (ora.bernoulli(43, 70, 130, 0.35), 130, 20, 3000, 2, 5)]
    for mat, cols, k, iters, restarts, seed in specs:
        for policy in ("pairwise_synergy", "higher_order_synergy"):
            res = refs.sa_select(mat, cols, k, iterations=iters, restarts=restarts, seed=seed,
                                 policy=policy)
            sa_cases.append({"rows": [[str(int(v)) for v in r] for r in mat], "cols": cols,
                             "window_size": k, "iterations": iters, "restarts": restarts,
                             "seed": seed, "policy": policy, "expect": res})
    with open(os.path.join(HERE, "golden_sa.json"), "w") as f:
        json.dump({"cases": sa_cases}, f)
    global_fixtures(ref, ora)
    io_fixtures(ref, ora)
    print("wrote fixtures", flush=True)


def combo_weights(mat, masks):
    """True weights of the row combinations selected by multi-word masks."""
    out = []
    for mk in masks:
        acc = np.zeros(mat.shape[1], np.uint64)
        for r in range(mat.shape[0]):
            if (int(mk[r // 64]) >> (r % 64)) & 1:
                acc ^= mat[r]
        out.append(sum(bin(int(v)).count("1") for v in acc))
    return np.array(out, np.uint32)


def gf2_inverse(M, n):
    """Inverse of an invertible n x n GF(2) matrix (rows of mask words)."""
    rows = [sum(int(M[i, q]) << (64 * q) for q in range(M.shape[1])) for i in range(n)]
    inv = [1 << i for i in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if (rows[r] >> c) & 1)
        rows[c], rows[p], inv[c], inv[p] = rows[p], rows[c], inv[p], inv[c]
        for r in range(n):
            if r != c and (rows[r] >> c) & 1:
                rows[r] ^= rows[c]
                inv[r] ^= inv[c]
    mw = (n + 63) // 64
    return np.array([[(v >> (64 * q)) & (2**64 - 1) for q in range(mw)] for v in inv], np.uint64)


def global_fixtures(ref, ora):
    """calc_scores_and_order (src/math/fwht.cpp:59-126) and GlobalGreedyApplier
    (src/apply/global_greedy_applier.cpp) -> golden_global.npz."""
    out = {}
    rng = np.random.default_rng(11)
    score_specs = [("i32", 12, 37), ("i32", 13, 300), ("i32", 12, 0), ("i32", 14, 70000),
                   ("i16", 12, 37), ("i16", 13, 5000), ("i16", 11, 30000)]
    for q, (kind, nb, n) in enumerate(score_specs):
        dt, lo, hi = (np.int32, -3 * n - 5, 3 * n + 5) if kind == "i32" else \
            (np.int16, max(-32768, -n - 5), min(32767, n + 5))
        h = rng.integers(lo, hi + 1, 1 << nb, dtype=np.int64).astype(dt)
        s, order, cnt, off = ref.scores_order(h, n)
        out[f"scores{q}_h"], out[f"scores{q}_n"] = h, np.array([n], np.int32)
        out[f"scores{q}_s"], out[f"scores{q}_order"] = s, order
        out[f"scores{q}_cnt"], out[f"scores{q}_off"] = cnt, off
    # Spectrum of a real window: h = FWHT of the active-column histogram.
    win = ora.bernoulli(51, 12, 640, 0.3)
    hist = np.zeros(1 << 12, np.int32)
    for c in range(640):
        p = sum(((int(win[r, c // 64]) >> (c % 64)) & 1) << r for r in range(12))
        hist[p] += 1 if p else 0
    n_act = int(hist.sum())
    h = ref.fwht_i32(hist)
    s, order, cnt, off = ref.scores_order(h, n_act)
    q = len(score_specs)
    out[f"scores{q}_h"], out[f"scores{q}_n"] = h, np.array([n_act], np.int32)
    out[f"scores{q}_s"], out[f"scores{q}_order"] = s, order
    out[f"scores{q}_cnt"], out[f"scores{q}_off"] = cnt, off
    out["scores_cases"] = np.array([q + 1], np.int32)

    cases = 0
    for t, cols, n_cand, seed in ((3, 64, 3, 61), (12, 256, 40, 62), (16, 300, 200, 63),
                                  (40, 512, 120, 64), (70, 192, 150, 65), (24, 128, 0, 66)):
        for req in (True, False):
            mat = ora.bernoulli(seed, t, cols, 0.35)
            mw = (t + 63) // 64
            if t <= 16 and n_cand:
                cands = ref.fwht_search(mat, cols, n_cand)
                masks = np.array([[mk] for _, mk in cands], np.uint64).reshape(-1, 1)
                ws = np.array([w for w, _ in cands], np.uint32)
            else:
                masks = rng.integers(0, 2**63, (n_cand, mw), dtype=np.int64).astype(np.uint64)
                if t % 64:
                    masks[:, -1] &= np.uint64((1 << (t % 64)) - 1)
                sparse = rng.random((n_cand, t)) < 3.0 / t   # a few rows per mask
                masks = np.zeros((n_cand, mw), np.uint64)
                for i, j in zip(*np.nonzero(sparse)):
                    masks[i, j // 64] |= np.uint64(1 << (j % 64))
                ws = combo_weights(mat, masks)
                # A few understated weights: the applier must still refuse a
                # transform that does not lower the true weight.
                ws[::7] = ws[::7] // 2
            if n_cand == 0:
                masks = np.zeros((0, mw), np.uint64)
                ws = np.zeros(0, np.uint32)
            res, ap, imp = ref.global_greedy(mat, cols, masks, ws, req)
            out[f"gg{cases}_in"], out[f"gg{cases}_masks"], out[f"gg{cases}_w"] = mat, masks, ws
            out[f"gg{cases}_meta"] = np.array([t, cols, int(req), ap, imp], np.int64)
            out[f"gg{cases}_out"] = res
            cases += 1
    # Planted windows A = M S: the rows of M^-1 recover the sparse S, mixed
    # with random sparse masks; weights are the true combination weights.
    for t, cols, seed in ((40, 512, 71), (70, 192, 72), (64, 256, 73)):
        S, A, M = ora.planted(seed, t, cols, 0.05)
        inv = gf2_inverse(M, t)
        mw = inv.shape[1]
        extra = np.zeros((3 * t, mw), np.uint64)
        for i, j in zip(*np.nonzero(rng.random((3 * t, t)) < 2.0 / t)):
            extra[i, j // 64] |= np.uint64(1 << (j % 64))
        masks = np.concatenate([extra[:t], inv, extra[t:]])
        ws = combo_weights(A, masks)
        for req in (True, False):
            res, ap, imp = ref.global_greedy(A, cols, masks, ws, req)
            out[f"gg{cases}_in"], out[f"gg{cases}_masks"], out[f"gg{cases}_w"] = A, masks, ws
            out[f"gg{cases}_meta"] = np.array([t, cols, int(req), ap, imp], np.int64)
            out[f"gg{cases}_out"] = res
            cases += 1
    out["gg_cases"] = np.array([cases], np.int32)
    np.savez_compressed(os.path.join(HERE, "golden_global.npz"), **out)


def io_fixtures(ref, ora):
    """BitMatrix text / BMMPYB1 binary writers, readers and the canonical hash
    (src/core/bit_matrix_io.cpp, bit_matrix_hash.cpp) -> golden_io.npz."""
    out = {}
    specs = [(np.zeros((0, 0), np.uint64), 0), (np.zeros((3, 0), np.uint64), 0),
             (np.zeros((0, 1), np.uint64), 5),
             # 1 x 65 with stray bits past column 64 (bit_matrix_hash_tests.cpp:11-31).
             (np.array([[0x1, 0x1 | (1 << 17)]], np.uint64), 65),
             (np.array([[0x1, 0x1]], np.uint64), 65),
             (np.array([[0x3, 0x1]], np.uint64), 65),
             (ora.bernoulli(101, 5, 130, 0.5), 130), (ora.bernoulli(102, 64, 1000, 0.3), 1000),
             (ora.bernoulli(103, 7, 64, 0.5), 64)]
    for q, (m, cols) in enumerate(specs):
        h, b, t = ref.matrix_io(m, cols)
        out[f"io{q}_in"], out[f"io{q}_cols"] = m, np.array([cols], np.int64)
        out[f"io{q}_hash"] = np.array([h])
        out[f"io{q}_bin"] = np.frombuffer(b, np.uint8)
        out[f"io{q}_text"] = np.array([t])
    out["io_cases"] = np.array([len(specs)], np.int32)
    # Malformed inputs and the reference's messages.
    good = ref.matrix_io(ora.bernoulli(104, 2, 70, 0.5), 70)
    bad_bin = [b"", b"BMMPY", b"BMMPYB2\x00" + good[1][8:], good[1][:20], good[1][:-3],
               good[1] + b"\x00\x00"]
    bad_txt = [b"", b"2", b"2 3\n101\n01", b"1 3\n1x1\n", b"1 3\n101\n1", b"1 3 101 \n\t ",
               b"2 2\n1 0\n0\r\n1\n"]
    msgs = []
    for kind, cases in ((0, bad_bin), (1, bad_txt)):
        for data in cases:
            try:
                words, r, c = ref.matrix_load(data, text=bool(kind))
                msgs.append((kind, data, "", words, r, c))
            except oracle.RefError as e:
                msgs.append((kind, data, str(e), np.zeros((0, 1), np.uint64), 0, 0))
    for q, (kind, data, msg, words, r, c) in enumerate(msgs):
        out[f"bad{q}_kind"] = np.array([kind], np.int32)
        out[f"bad{q}_data"] = np.frombuffer(data, np.uint8)
        out[f"bad{q}_msg"] = np.array([msg])
        out[f"bad{q}_words"], out[f"bad{q}_shape"] = words, np.array([r, c], np.int64)
    out["bad_cases"] = np.array([len(msgs)], np.int32)
    np.savez_compressed(os.path.join(HERE, "golden_io.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["global"]:
        global_fixtures(oracle.reference(), oracle.oracle())
    elif sys.argv[1:] == ["io"]:
        io_fixtures(oracle.reference(), oracle.oracle())
    else:
        main()

"""Pairwise local improvement of GA rows (test infrastructure: a CPU
restatement used only as a checker by tests/).

The reference rule (src/ga/genetic_algorithm_operators.cpp:131-192) walks the
rows in order, up to 10 iterations until none improves.  For row i of weight
w_i > 0 it takes the other row j minimising w_i + w_j - 2 |r_i & r_j| (the
first such j, rows with w_j >= 2 w_i skipped) and replaces r_i by r_i ^ r_j
when that is lighter; later rows see the updated earlier ones.

The GPU applies the same rule in rounds so that one round's moves are
independent: in round (b, phase) the rows whose index bit b equals phase are
targets and the others sources, every target picks its best source among
the unchanged sources, and all picks apply at once.  On two rows the rounds
(0, 0), (0, 1) are exactly one iteration of the reference's loop.
"""
import numpy as np


def _ints(rows_words):
    return [int.from_bytes(np.ascontiguousarray(r).tobytes(), "little") for r in rows_words]


def _words(rows, n_words):
    return np.array([[(r >> (64 * k)) & ((1 << 64) - 1) for k in range(n_words)] for r in rows],
                    dtype=np.uint64).reshape(len(rows), n_words)


def _pc(x):
    return bin(x).count("1")


def li_round(rows, b, phase):
    """One round on a list of row ints: returns the new list."""
    n = len(rows)
    w = [_pc(r) for r in rows]
    moves = {}
    for i in range(n):
        if (i >> b) & 1 != phase or w[i] == 0:
            continue
        best = None
        for j in range(n):
            if (j >> b) & 1 == phase or w[j] >= 2 * w[i]:
                continue
            cw = w[i] + w[j] - 2 * _pc(rows[i] & rows[j])
            if cw < w[i] and (best is None or (cw, j) < best):
                best = (cw, j)
        if best:
            moves[i] = best[1]
    out = list(rows)
    for i, j in moves.items():
        out[i] = rows[i] ^ rows[j]
    return out


def li_rounds(rows_words, sweeps=1):
    """`sweeps` passes of the rounds (b, phase) for b < bits(n - 1), phase
    0 then 1, on uint64 rows (n, W): the GPU schedule."""
    rows = _ints(rows_words)
    n = len(rows)
    bits = max(1, (n - 1).bit_length())
    for _ in range(sweeps):
        for b in range(bits):
            for phase in (0, 1):
                rows = li_round(rows, b, phase)
    return _words(rows, rows_words.shape[1])


def li_sequential(rows_words, iters=10):
    """The reference's sequential loop on uint64 rows (n, W)."""
    rows = _ints(rows_words)
    n = len(rows)
    w = [_pc(r) for r in rows]
    if n < 2:
        return _words(rows, rows_words.shape[1])
    for _ in range(iters):
        improved = False
        for i in range(n):
            if w[i] == 0:
                continue
            best_j, best_w = n, w[i]
            for j in range(n):
                if j == i or w[j] >= 2 * w[i]:
                    continue
                cw = w[i] + w[j] - 2 * _pc(rows[i] & rows[j])
                if cw < best_w:
                    best_w, best_j = cw, j
            if best_j == n:
                continue
            rows[i] ^= rows[best_j]
            w[i] = best_w
            improved = True
        if not improved:
            break
    return _words(rows, rows_words.shape[1])

"""Pins the pairwise local improvement restatement (oracle/ga_li.py) that the
GPU GA is checked against: on two rows the round schedule is the reference's
sequential rule (genetic_algorithm_operators.cpp:131-192); every round equals
a brute force over the sources by direct popcounts of r_i ^ r_j (not the
overlap identity the rule uses); rounds never add weight and keep the row
space."""
import numpy as np

from oracle.ga_li import _ints, li_round, li_rounds, li_sequential


def _rows(rng, n, words, p):
    bits = rng.random((n, 64 * words)) < p
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(n, words)


def _rank(ints):
    rows, r = list(ints), 0
    for bit in range(max((x.bit_length() for x in rows), default=0)):
        piv = next((k for k in range(r, len(rows)) if (rows[k] >> bit) & 1), None)
        if piv is None:
            continue
        rows[r], rows[piv] = rows[piv], rows[r]
        for k in range(len(rows)):
            if k != r and (rows[k] >> bit) & 1:
                rows[k] ^= rows[r]
        r += 1
    return r


def test_two_rows_reduce_to_the_reference_rule():
    rng = np.random.default_rng(11)
    for case in range(300):
        A = _rows(rng, 2, 1 + case % 3, [0.1, 0.3, 0.6][case % 3])
        if case % 7 == 0:
            A[1] = A[0] ^ _rows(rng, 1, A.shape[1], 0.05)[0]   # near-duplicates improve
        for it in (1, 2, 10):
            assert np.array_equal(li_rounds(A, it), li_sequential(A, it)), (case, it)


def test_hand_case():
    # r0 = 0b1111, r1 = 0b0111: r0 -> r0 ^ r1 = 0b1000 (weight 1 < 4); then
    # r1 (weight 3) with r0' (weight 1): 0b1111 has weight 4, no move.
    A = np.array([[0b1111], [0b0111]], np.uint64)
    exp = np.array([[0b1000], [0b0111]], np.uint64)
    assert np.array_equal(li_sequential(A), exp)
    assert np.array_equal(li_rounds(A, 1), exp)


def test_round_equals_brute_force():
    rng = np.random.default_rng(12)
    for case in range(200):
        n = 2 + case % 7
        A = _rows(rng, n, 1 + case % 2, [0.15, 0.4][case % 2])
        if case % 3 == 0:   # planted near-duplicate pairs
            for i in range(0, n - 1, 2):
                A[i + 1] = A[i] ^ _rows(rng, 1, A.shape[1], 0.05)[0]
        rows = _ints(A)
        for b in range(max(1, (n - 1).bit_length())):
            for phase in (0, 1):
                got = li_round(rows, b, phase)
                src = [j for j in range(n) if (j >> b) & 1 != phase]
                for i in range(n):
                    wi = bin(rows[i]).count("1")
                    if (i >> b) & 1 != phase or wi == 0:
                        assert got[i] == rows[i]
                        continue
                    cands = sorted((bin(rows[i] ^ rows[j]).count("1"), j) for j in src)
                    best = cands[0] if cands else None
                    exp = rows[i] ^ rows[best[1]] if best and best[0] < wi else rows[i]
                    assert got[i] == exp, (case, b, phase, i)
                assert sum(bin(x).count("1") for x in got) <= sum(bin(x).count("1") for x in rows)
                assert _rank(got) == _rank(rows)
                rows = got

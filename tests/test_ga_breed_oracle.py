"""Pins of oracle/ga_breed.py (the restated GPU breeding step) against the
reference GA's rules: ranked elitism (genetic_algorithm_lifecycle.cpp:
127-134), tournaments (genetic_algorithm_operators.cpp:39-56), the
mutation count and rate adaptation (:100-129, genetic_algorithm_core.cpp:
33-41), the catastrophe (:15-37), restarts (lifecycle.cpp:118-142) and random
individuals (genetic_algorithm_population.cpp:20-36)."""
import numpy as np

from oracle.ga_breed import F, _Rng, breed, random_sequence


def _params(**kw):
    p = {"rows": 16, "pop": 32, "max_ops": 64, "tournament": 3, "elite": 2, "mutations": 0,
         "p_append": 0.6, "p_replace": 0.25, "p_delete": 0.15, "learned": 0, "n_cand": 0,
         "p_learned": 0.5, "num_parents": 2, "num_offspring": 1, "mutation_rate": 0.0,
         "catastrophe_threshold": 0, "catastrophe_survival": 0.2, "restart_after": 0,
         "restart_p": 0.2}
    p.update(kw)
    return p


def _state(best=10 ** 6, rate=0.3, noimp=0, cat=0):
    return {"best_weight": best, "best_len": 0, "no_improvement": noimp, "catastrophe": cat,
            "mutation_rate": rate}


def _pop(p, seed=1):
    rng = np.random.default_rng(seed)
    seqs = [[int(rng.integers(0, p["rows"])) | (int(rng.integers(0, p["rows"])) << 8)
             for _ in range(int(rng.integers(0, 12)))] for _ in range(p["pop"])]
    fit = rng.integers(100, 140, p["pop"]).astype(np.uint32)
    rs = [int(x) | 1 for x in rng.integers(1, 2 ** 62, p["pop"])]
    return seqs, fit, rs


def test_elites_are_the_ranked_best_and_best_is_recorded():
    p = _params(elite=5)
    seqs, fit, rs = _pop(p)
    kids, _, st, best, mode = breed(p, seqs, fit, rs, _state())
    order = np.lexsort((np.arange(p["pop"]), fit))     # by fitness, then index
    assert mode == 0
    for k in range(5):
        assert kids[k] == seqs[order[k]]
    assert best == order[0] and st["best_weight"] == int(fit[order[0]])
    assert st["no_improvement"] == 0 and st["catastrophe"] == 0
    # no improvement: counters grow, best kept
    _, _, st2, best2, _ = breed(p, seqs, fit, rs, _state(best=int(fit.min()), noimp=3, cat=3))
    assert best2 is None and st2["no_improvement"] == 4 and st2["catastrophe"] == 4


def test_large_tournaments_pick_the_lightest():
    """With 64 draws from 8 individuals every tournament sees the lightest
    (missing it has probability (7/8)^64 < 3e-4 per tournament; the draws are
    fixed by the seeds), so every crossover child is a prefix and a suffix of
    the best individual."""
    p = _params(pop=8, elite=0, tournament=64)
    seqs, fit, rs = _pop(p, seed=3)
    fit = np.array([50, 40, 60, 30, 70, 45, 55, 65], np.uint32)
    kids, *_ = breed(p, seqs, fit, rs, _state())
    b = seqs[3]
    for k in kids:
        assert any(k == b[:i] + b[j:] for i in range(len(b) + 1) for j in range(len(b) + 1))


def test_mutation_count_and_row_xor_form():
    """Reference mutation: max(1, floor(rows * rate)) draws of (i, j), i == j
    skipped, each appended as the op r_i ^= r_j (src j, dst i)."""
    p = _params(pop=64, elite=0, num_parents=1, mutation_rate=0.5, rows=16, max_ops=64)
    seqs = [[] for _ in range(64)]
    fit = np.full(64, 7, np.uint32)
    rs = [(0x1234567 * (i + 1)) | 1 for i in range(64)]
    kids, _, st, *_ = breed(p, seqs, fit, rs, _state(best=7, rate=0.5))
    count = int(F(16) * F(0.475))      # rate adapted once: max(0.3, 0.5 * 0.95)
    assert count == 7
    lens = [len(k) for k in kids]
    assert max(lens) <= count and min(lens) >= 1
    assert abs(np.mean(lens) - count * 15 / 16) < 1.0
    for k in kids:
        for op in k:
            assert (op & 0xff) != (op >> 8) and (op & 0xff) < 16 and (op >> 8) < 16


def test_rate_adaptation_known_values():
    """genetic_algorithm_core.cpp:33-41 on hand-computed values."""
    p = _params(pop=4, elite=4, mutation_rate=0.3)
    seqs, fit, rs = _pop(p)
    hi = 10 ** 9
    cases = [  # (no_improvement before, rate before) -> rate after (float32)
        ((0, 0.3), 0.3),        # max(0.3, 0.285); the best improves: counter 0
        ((5, 0.5), 0.65),       # counter 6: min(0.6, 0.65) -> 0.6
        ((10, 0.5), 0.75),      # counter 11: min(0.8, 0.75) -> 0.75
        ((20, 0.7), 1.05),      # counter 21: min(0.8, 1.05) -> 0.8
    ]
    expect = [F(0.3), F(0.6), F(0.75), F(0.8)]
    for ((ni, r0), _), e in zip(cases, expect):
        st0 = _state(best=hi if ni == 0 else 0, rate=r0, noimp=ni)
        _, _, st, *_ = breed(p, seqs, fit, rs, st0)
        assert st["mutation_rate"] == e, (ni, r0, st["mutation_rate"], e)


def test_catastrophe_keeps_survivors_and_randomizes_the_rest():
    p = _params(pop=20, elite=2, catastrophe_threshold=3, catastrophe_survival=0.25,
                rows=8, max_ops=16)
    seqs, fit, rs = _pop(p)
    kids, _, st, _, mode = breed(p, seqs, fit, rs, _state(best=0, cat=2, noimp=2))
    assert mode == 1 and st["catastrophe"] == 0 and st["no_improvement"] == 3
    order = np.lexsort((np.arange(20), fit))
    for k in range(5):                                 # floor(20 * 0.25)
        assert kids[k] == seqs[order[k]]
    for k in kids[5:]:
        assert 1 <= len(k) <= 16 and all((o & 0xff) != (o >> 8) for o in k)
    # threshold not reached: ordinary step
    _, _, st2, _, mode2 = breed(p, seqs, fit, rs, _state(best=0, cat=1, noimp=1))
    assert mode2 == 0 and st2["catastrophe"] == 2


def test_restarts_only_after_the_stale_limit():
    p = _params(pop=64, elite=0, restart_after=15, restart_p=0.2, rows=8, max_ops=24)
    seqs, fit, rs = _pop(p)
    seqs = [[] for _ in seqs]   # crossover children of empty parents stay empty
    kids, *_ = breed(p, seqs, fit, rs, _state(best=0, noimp=14))    # counter 15: not > 15
    assert all(len(k) == 0 for k in kids)
    kids, *_ = breed(p, seqs, fit, rs, _state(best=0, noimp=15))    # counter 16
    n = sum(len(k) > 0 for k in kids)
    assert 3 <= n <= 25                                 # about 0.2 * 64


def test_random_sequence_is_identity_plus_3n_row_xors():
    r = _Rng(99)
    seq = random_sequence(8, 1000, r)
    assert 15 <= len(seq) <= 24 and all((o & 0xff) != (o >> 8) for o in seq)
    assert len(random_sequence(64, 50, _Rng(5))) == 50   # capped at max_ops


def test_offspring_groups_share_parents_and_cuts():
    """num_offspring children of one group are identical before mutation
    (the reference's crossover builds every offspring of a parent set the
    same way, operators.cpp:84-98)."""
    p = _params(pop=24, elite=0, num_offspring=3, num_parents=3)
    seqs, fit, rs = _pop(p)
    kids, *_ = breed(p, seqs, fit, rs, _state())
    for g in range(8):
        assert kids[3 * g] == kids[3 * g + 1] == kids[3 * g + 2]

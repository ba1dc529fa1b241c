"""GPU island-model GA: exactness of the reported fitness, invertibility of
the found transform, monotone progress, ring migration and determinism."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _weight(m):
    return int(sum(bin(int(v)).count("1") for v in m.reshape(-1)))


def _small(lib, **kw):
    S, A, M = lib.gen_planted(0x5EED0005, 64, 2048, 0.02)
    p = lib.ga_default_params(islands=4, population=64, max_ops=64, migrate_every=5,
                              check_every=5, **kw)
    return S, A, p


@pytest.mark.parametrize("learned", [0, 1])
def test_ga_fitness_exact_and_transform_invertible(lib, ora, learned):
    _say("ga small", learned)
    S, A, p = _small(lib, learned=learned, train_steps=4, n_candidates=32)
    ga = lib.GpuGA(A, 2048, p)
    st0 = ga.run(1)
    st = ga.run(60)
    best, w = ga.best()
    ga.close()
    assert st.generations == 61
    assert st.best_weight == w == _weight(best)
    assert w <= _weight(A)
    assert w < _weight(A)                          # some progress in 61 generations
    # row operations only: same row space, full rank kept
    assert ora.rank(best, 2048) == ora.rank(A, 2048) == 64
    assert ora.rank(np.concatenate([A, best]), 2048) == 64
    assert st.evaluations == 61 * 4 * 64
    assert st0.best_weight >= st.best_weight


def test_ga_deterministic_for_a_seed(lib):
    _say("ga determinism")
    S, A, p = _small(lib, learned=0)
    r = []
    for _ in range(2):
        ga = lib.GpuGA(A, 2048, p)
        st = ga.run(30)
        r.append((st.best_weight, ga.best()[0].tobytes()))
        ga.close()
    assert r[0] == r[1]


def test_ga_graph_equals_eager(lib):
    _say("ga graph vs eager")
    S, A, p = _small(lib, learned=0)
    out = []
    for g in (1, 0):
        p.use_graph = g
        ga = lib.GpuGA(A, 2048, p)
        st = ga.run(25)
        out.append((st.best_weight, ga.best()[0].tobytes()))
        ga.close()
    assert out[0] == out[1]


def test_ga_reaches_target_stops(lib):
    S, A, p = _small(lib, learned=0)
    ga = lib.GpuGA(A, 2048, p)
    st = ga.run(1000, target=_weight(A))            # already at target
    ga.close()
    assert st.reached_target == 1 and st.ran <= p.check_every


def test_ga_rejects_bad_shapes(lib):
    A = np.zeros((8, 16), np.uint64)
    with pytest.raises(ValueError):
        lib.GpuGA(A, 1000, lib.ga_default_params())
    with pytest.raises(ValueError):
        lib.GpuGA(np.zeros((300, 32), np.uint64), 2048, lib.ga_default_params())


def _li_reference(rows_words, sweeps=1):
    """Parallel-round pairwise improvement, restated in Python: per round the
    rows with index bit b == phase are targets, the others sources; a target
    i takes the source j minimising (wt_i + wt_j - 2 |r_i & r_j|, j) if that
    beats wt_i (genetic_algorithm_operators.cpp:131-192 with rounds)."""
    rows = [int.from_bytes(r.tobytes(), "little") for r in rows_words]
    n = len(rows)
    bits = max(1, (n - 1).bit_length())
    pc = lambda x: bin(x).count("1")   # noqa: E731
    for _ in range(sweeps):
        for b in range(bits):
            for phase in (0, 1):
                w = [pc(r) for r in rows]
                moves = {}
                for i in range(n):
                    if (i >> b) & 1 != phase or w[i] == 0:
                        continue
                    best = None
                    for j in range(n):
                        if (j >> b) & 1 == phase or w[j] >= 2 * w[i]:
                            continue
                        cw = w[i] + w[j] - 2 * pc(rows[i] & rows[j])
                        if cw < w[i] and (best is None or (cw, j) < best):
                            best = (cw, j)
                    if best:
                        moves[i] = best[1]
                for i, j in moves.items():
                    rows[i] ^= rows[j]
    W = rows_words.shape[1]
    return np.array([[(r >> (64 * k)) & ((1 << 64) - 1) for k in range(W)] for r in rows],
                    dtype=np.uint64)


def test_ga_local_improvement_matches_restatement(lib):
    _say("ga local improvement")
    S, A, M = lib.gen_planted(0x5EED0005, 32, 2048, 0.05)
    p = lib.ga_default_params(islands=1, population=4, max_ops=8, migrate_every=1,
                              mutations=0, li_sweeps=1, check_every=1, learned=0)
    ga = lib.GpuGA(A, 2048, p)
    st = ga.run(1)
    best, w = ga.best()
    ga.close()
    exp = _li_reference(A, 1)
    assert st.li_moves > 0
    assert np.array_equal(best, exp)
    assert w == st.best_weight == _weight(exp)

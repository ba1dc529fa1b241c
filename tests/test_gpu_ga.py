"""GPU island-model GA: exactness of the reported fitness, invertibility of
the found transform, monotone progress, ring migration and determinism."""
import numpy as np
import pytest

from oracle.ga_li import li_rounds

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


@pytest.mark.parametrize("init_rounds", [0, 1])
def test_ga_local_improvement_matches_restatement(lib, init_rounds):
    """One generation with an epoch after it: one pairwise sweep at the
    epoch, and with init_rounds one more before the first generation (no
    mutations, so the generation itself changes nothing)."""
    _say("ga local improvement", init_rounds)
    S, A, M = lib.gen_planted(0x5EED0005, 32, 2048, 0.05)
    p = lib.ga_default_params(islands=1, population=4, max_ops=8, migrate_every=1,
                              mutations=0, li_sweeps=1, check_every=1, learned=0, win_rows=0,
                              isd_trials=0, init_rounds=init_rounds)
    ga = lib.GpuGA(A, 2048, p)
    st = ga.run(1)
    best, w = ga.best()
    ga.close()
    exp = li_rounds(A, 1 + init_rounds)
    assert st.li_moves > 0
    assert np.array_equal(best, exp)
    assert w == st.best_weight == _weight(exp)


def _win_reference(ora, A, rounds_bits):
    """Window improvement restated: for each round (bit, phase), every row i
    with index bit `bit` == phase takes its exact coset minimum over the
    subsets of the other half (all of it as the window, in index order; the
    oracle's coset minimum, lowest mask on ties) when that beats wt(r_i)."""
    rows = A.copy()
    n = rows.shape[0]
    for bit, phase in rounds_bits:
        src = [j for j in range(n) if (j >> bit) & 1 != phase]
        new = rows.copy()
        for i in range(n):
            if (i >> bit) & 1 != phase:
                continue
            sub = np.concatenate([rows[src], rows[i:i + 1]])
            w, idx = ora.coset_min(sub, rows.shape[1] * 64)
            if int(w[-1]) < _weight(rows[i:i + 1]):
                acc = rows[i].copy()
                for k, j in enumerate(src):
                    if (int(idx[-1]) >> k) & 1:
                        acc ^= rows[j]
                new[i] = acc
        rows = new
    return rows


@pytest.mark.parametrize("learned,cols", [(0, 2048), (1, 2048), (0, 6144)])
def test_ga_window_improvement_matches_restatement(lib, ora, learned, cols):
    """One epoch of window improvement alone (no mutation, no pairwise
    improvement) on 32 x 2048: with 16-row halves a 16-row window is the whole
    other half, so the exact move is determined (plain: in index order, equal
    rows; learned: windows in model order, so equal weights, not rows)."""
    _say("ga window improvement", learned, cols)
    S, A, M = lib.gen_planted(0x5EED0006, 32, cols, 0.05)
    rounds = 1 if learned else 2   # ties may pick other rows in model order
    p = lib.ga_default_params(islands=2, population=4, max_ops=8, migrate_every=1,
                              mutations=0, li_sweeps=0, check_every=1, learned=learned,
                              win_rows=16, win_rounds=rounds, win_candidates=32, isd_trials=0,
                              init_rounds=0)
    ga = lib.GpuGA(A, cols, p)
    st = ga.run(1)
    best, w = ga.best()
    ga.close()
    exp = _win_reference(ora, A, [(0, 0), (1, 0)][:rounds])
    assert st.win_moves > 0
    assert w == st.best_weight == _weight(exp)
    if not learned:
        assert np.array_equal(best, exp)
    # Row operations only: same row space (full rank of [A; best] equals rank A).
    both = np.concatenate([A, best])
    assert ora.rank(both, cols) == ora.rank(A, cols) == 32


def _isd_reference(A, seed, isl, cols):
    """One information-set trial restated: column draws from XorShift64
    seeded as the kernel seeds it, Gauss-Jordan elimination pivoting on the
    lowest unpivoted row that has the column, then GlobalGreedyApplier's pool
    (eliminated rows with their masks, then the island rows with unit masks),
    stable by weight, kept while independent."""
    M64 = (1 << 64) - 1
    n = A.shape[0]
    rows = [int(sum(int(A[i, k]) << (64 * k) for k in range(A.shape[1]))) for i in range(n)]
    R = list(rows)
    T = [1 << i for i in range(n)]
    rng = (seed ^ ((0x9E3779B97F4A7C15 * (isl + 1)) & M64)) & M64 or 1
    pivoted = [False] * n
    npiv, draws = 0, 0
    while npiv < n and draws < 8 * cols + 64:
        draws += 1
        rng ^= (rng << 13) & M64
        rng ^= rng >> 7
        rng ^= (rng << 17) & M64
        c = rng % cols
        has = [(R[i] >> c) & 1 for i in range(n)]
        cand = [i for i in range(n) if has[i] and not pivoted[i]]
        if not cand:
            continue
        r = cand[0]
        pivoted[r] = True
        npiv += 1
        for i in range(n):
            if has[i] and i != r:
                R[i] ^= R[r]
                T[i] ^= T[r]
    pool = [(bin(R[i]).count("1"), i) for i in range(n)] + \
           [(bin(rows[j]).count("1"), n + j) for j in range(n)]
    pool.sort()
    piv, out = {}, []
    for _, idx in pool:
        v = T[idx] if idx < n else 1 << (idx - n)
        while v:
            hb = v.bit_length() - 1
            if hb in piv:
                v ^= piv[hb]
            else:
                piv[hb] = v
                out.append(R[idx] if idx < n else rows[idx - n])
                break
        if len(out) == n:
            break
    W = A.shape[1]
    return np.array([[(r >> (64 * k)) & M64 for k in range(W)] for r in out], dtype=np.uint64)


@pytest.mark.parametrize("islands,init_rounds", [(1, 0), (2, 0), (1, 1), (2, 1)])
def test_ga_information_set_trial_matches_restatement(lib, ora, islands, init_rounds):
    """One information-set trial per island and nothing else (no mutation,
    no pairwise or window improvement), either at the epoch after one
    generation or, with init_rounds and no epoch, before the first
    generation: each island base equals the restated trial; the best island
    is reported."""
    _say("ga information-set trial", islands, init_rounds)
    cols = 2048
    S, A, M = lib.gen_planted(0x5EED0008, 40, cols, 0.03)
    p = lib.ga_default_params(islands=islands, population=4, max_ops=8,
                              migrate_every=1000 if init_rounds else 1, mutations=0,
                              li_sweeps=0, check_every=1, learned=0, win_rows=0, isd_trials=1,
                              init_rounds=init_rounds)
    ga = lib.GpuGA(A, cols, p)
    st = ga.run(1)
    best, w = ga.best()
    ga.close()
    seed = (0x5EED0005 ^ 0xA0761D6478BD642F) & ((1 << 64) - 1)
    exps = [_isd_reference(A, seed, i, cols) for i in range(islands)]
    ws = [_weight(e) for e in exps]
    k = ws.index(min(ws))
    assert w == st.best_weight == ws[k] < _weight(A)
    assert np.array_equal(best, exps[k])
    assert st.isd_rows > 0
    both = np.concatenate([A, best])
    assert ora.rank(both, cols) == ora.rank(A, cols) == 40


def test_ga_information_set_trials_recover_planted(lib):
    """On a planted 64 x 2048 instance (Bernoulli(0.01) rows under a random
    invertible 64 x 64 transform) information-set trials reach the planted
    weight within a few epochs."""
    _say("ga information-set planted")
    S, A, M = lib.gen_planted(0x5EED0009, 64, 2048, 0.01)
    planted = _weight(S)
    p = lib.ga_default_params(islands=4, population=16, max_ops=16, migrate_every=2,
                              check_every=2, learned=0, isd_trials=2)
    ga = lib.GpuGA(A, 2048, p)
    st = ga.run(40, target=int(planted * 1.01))
    best, w = ga.best()
    ga.close()
    assert st.reached_target and w <= 1.01 * planted
    assert _weight(best) == w

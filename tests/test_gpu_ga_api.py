"""The GA classes' reference API on the GPU GA: the reference's own tests
(tests/python/test_ga_api.py:52-118) and what their results must satisfy:
each candidate's mask reproduces its output row from the window rows, the
weights add up to the best score, the transform is invertible, and the best
score never exceeds the window's weight."""
import numpy as np
import pytest

import bmmpy as bmm
from bmmpy import ga as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def make_ga_matrix():
    return bmm.matrix_from_rows(["11001000", "01100100", "00100110", "10010001"])


def _spec(island_id, interval):
    mig = bmm.MigrationPolicy()
    mig.interval_generations = interval
    mig.export_count = 1
    mig.import_count = 1
    spec = bmm.IslandSpec()
    spec.island_id = island_id
    spec.migration = mig
    return spec


def _native_config(seed):
    c = bmm.ga.GeneticAlgorithmConfig()
    c.population_size = 12
    c.elite_count = 2
    c.tournament_size = 2
    c.mutation_rate = 0.25
    c.enable_catastrophe = True
    c.catastrophe_threshold = 2
    c.catastrophe_survival_rate = 0.5
    stop = bmm.StopCriteria()
    stop.max_generations = 4
    stop.max_stale_generations = 4
    c.stop = stop
    c.seed = seed
    return c


def _check_candidates(window, cands):
    words, cols = G._window_words(window)
    n = len(window)
    assert len(cands) == n
    masks = np.zeros((n, n), np.uint8)
    for i, c in enumerate(cands):
        acc = np.zeros(words.shape[1], np.uint64)
        for j in range(n):
            if c.has_row(j):
                acc ^= words[j]
                masks[i, j] = 1
        assert int(np.unpackbits(acc.view(np.uint8)).sum()) == c.weight
    # invertible when the window's rows are independent
    a = masks.copy()
    r = 0
    for col in range(n):
        piv = [i for i in range(r, n) if a[i, col]]
        if not piv:
            continue
        a[[r, piv[0]]] = a[[piv[0], r]]
        for i in range(n):
            if i != r and a[i, col]:
                a[i] ^= a[r]
        r += 1
    return r, int(np.unpackbits(words.view(np.uint8)).sum())


def test_genetic_algorithm_stepwise_wrapper():
    _say("ga api stepwise")
    matrix = make_ga_matrix()
    ga = bmm.GeneticAlgorithm(population_size=16, elite_count=2, tournament_size=2,
                              mutation_rate=0.3, max_generations=4, max_stale_generations=4,
                              seed=7, num_parents=2, num_offspring=1, enable_catastrophe=True,
                              catastrophe_threshold=2, catastrophe_survival_rate=0.5)
    window = matrix.row_window([0, 1, 2, 3])
    ga.initialize(window)
    while not ga.done():
        ga.step()
    individual = ga.best_individual()
    stats = ga.stats()
    assert ga.name() == "ga"
    assert len(individual) == 4
    assert ga.best_score() == sum(c.weight for c in individual)
    assert stats.generations == ga.generation() == 4
    assert stats.evaluations > 0
    assert stats.seed == 7
    assert ga.num_parents == 2 and ga.num_offspring == 1 and ga.enable_catastrophe
    assert ga.catastrophe_threshold == 2 and ga.catastrophe_survival_rate == 0.5
    rank, start = _check_candidates(window, individual)
    assert rank == 4 and ga.best_score() <= start


def test_island_model_wrapper():
    _say("ga api island model")
    matrix = make_ga_matrix()
    window = matrix.row_window([0, 1, 2, 3])
    model = bmm.IslandModel(islands=[_spec(0, 4), _spec(1, 6)],
                            ga_configs=[_native_config(11), _native_config(19)],
                            channel_capacity=32, overflow_policy="drop_oldest")
    best = model.run_to_completion(window)
    snapshot = model.snapshot()
    assert model.name() == "island_model"
    assert len(best) == 4
    assert snapshot.island_count == 2
    assert len(snapshot.islands) == 2
    assert snapshot.best_score == sum(c.weight for c in best)
    assert not snapshot.running
    assert [s.island_id for s in snapshot.islands] == [0, 1]
    assert all(s.finished and s.best_score >= snapshot.best_score for s in snapshot.islands)
    assert snapshot.islands[1].stats.seed == 19
    rank, start = _check_candidates(window, best)
    assert rank == 4 and snapshot.best_score <= start


def test_planted_window_improves_and_stops_on_target(lib):
    _say("ga api planted")
    S, A, _ = lib.gen_planted(0x5EED0006, 32, 2500, 0.05)
    rows = ["".join("1" if (int(r[c // 64]) >> (c % 64)) & 1 else "0" for c in range(2500))
            for r in A]
    mat = bmm.matrix_from_rows(rows)
    window = mat.row_window(list(range(32)))
    planted = int(sum(bin(int(v)).count("1") for v in S.reshape(-1)))
    ga = bmm.GeneticAlgorithm(population_size=64, seed=3, max_generations=200,
                              target_total_weight=int(planted * 1.01))
    best = ga.optimize(window)
    rank, start = _check_candidates(window, best)
    assert rank == 32
    assert ga.best_score() <= int(planted * 1.01) < start
    assert ga.done() and ga.generation() <= 200


def test_single_row_dependent_rows_and_request_stop():
    _say("ga api edge windows")
    one = bmm.matrix_from_rows(["1011"])
    ga = bmm.GeneticAlgorithm(population_size=8, max_generations=3)
    best = ga.optimize(one.row_window([0]))
    assert [c.weight for c in best] == [3] and ga.generation() == 3
    dep = bmm.matrix_from_rows(["110011", "011110", "101101", "000000"])
    w = dep.row_window([0, 1, 2, 3])
    best = bmm.GeneticAlgorithm(population_size=8, max_generations=5).optimize(w)
    _check_candidates(w, best)
    assert sum(c.weight for c in best) <= w.weight()
    # No stop criteria: runs until request_stop.
    model = bmm.IslandModel(islands=[_spec(0, 5)], population_size=8)
    model.initialize(make_ga_matrix().row_window([0, 1, 2, 3]))
    model.start()
    assert model.running() or model.snapshot().total_generations >= 0
    model.request_stop()
    model.wait()
    assert not model.running() and model.stop_requested()
    assert len(model.best_individual()) == 4


def test_island_model_lifecycle_errors_and_island_seeds():
    """initialize() or start() on a running model raise as the reference's
    IslandModel (src/ga/island_model.cpp:32-35, 56-66); islands seeded
    seed + i * seed_stride draw from their own streams."""
    _say("ga api lifecycle")
    w = make_ga_matrix().row_window([0, 1, 2, 3])
    model = bmm.IslandModel(islands=[_spec(0, 5), _spec(1, 5)], population_size=8, seed=4)
    model.initialize(w)
    pops = [model._run.ga.population(i, 8, 512)["rng"] for i in range(2)]
    assert not np.array_equal(pops[0], pops[1])
    model.start()
    try:
        with pytest.raises(RuntimeError, match="model is already running"):
            model.initialize(w)
        with pytest.raises(RuntimeError, match="model is already running"):
            model.start()
    finally:
        model.request_stop()
        model.wait()
    assert not model.running()
    model.initialize(w)      # allowed again after wait()
    same = bmm.IslandModel(islands=[_spec(0, 5), _spec(1, 5)], population_size=8, seed=4,
                           seed_stride=0)
    same.initialize(w)
    pops = [same._run.ga.population(i, 8, 512)["rng"] for i in range(2)]
    assert np.array_equal(pops[0], pops[1])
    # export_count = 0 on island 0 and import_count = 0 on island 1 reach the device
    s0, s1 = _spec(0, 5), _spec(1, 5)
    s0.migration.export_count = 0
    s1.migration.import_count = 0
    flags = bmm.IslandModel(islands=[s0, s1], population_size=8)
    flags.initialize(w)
    assert [flags._run.ga.population(i, 8, 512)["migrate"] for i in range(2)] == [2, 1]

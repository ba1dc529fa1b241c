"""The GA classes' reference API (bmmpy/ga.py, python/bindings/ga.cpp) without
a device: record types and defaults, argument validation in the reference's
order and messages, and the host-side GF(2) step that turns the best rows
into one Candidate per window row."""
import numpy as np
import pytest

import bmmpy as bmm
from bmmpy import ga as G


def test_records_and_defaults():
    c = bmm.ga.GeneticAlgorithmConfig()
    assert (c.population_size, c.elite_count, c.num_parents, c.num_offspring,
            c.tournament_size, c.enable_catastrophe, c.catastrophe_threshold,
            c.catastrophe_survival_rate, c.mutation_rate, c.seed) == (
        300, 3, 2, 1, 3, False, 30, 0.2, 0.3, 0)
    s = bmm.StopCriteria()
    assert (s.max_generations, s.max_stale_generations, s.target_total_weight) == (None,) * 3
    m = bmm.MigrationPolicy()
    assert (m.interval_generations, m.export_count, m.import_count,
            m.shared_pool_capacity) == (32, 2, 2, 128)
    spec = bmm.IslandSpec()
    assert spec.island_id == 0 and spec.migration.interval_generations == 32
    r = bmm.RunStats()
    assert (r.generations, r.evaluations, r.stale_generations, r.seed) == (0, 0, 0, 0)
    snap = bmm.IslandModelSnapshot()
    assert snap.islands == [] and not snap.running
    assert bmm.ga.IslandModelConfig().islands == []
    with pytest.raises(AttributeError):
        c.no_such_field = 1
    for name in ("GeneticAlgorithm", "IslandModel", "IslandModelSnapshot", "IslandSnapshot",
                 "IslandSpec", "MigrationPolicy", "OverflowPolicy", "RunStats", "StopCriteria"):
        assert name in bmm.__all__ and getattr(bmm, name) is getattr(bmm.ga, name)


def test_genetic_algorithm_surface_without_device():
    ga = bmm.GeneticAlgorithm(population_size=16, seed=7, max_generations=4,
                              enable_catastrophe=True, catastrophe_threshold=2)
    assert ga.name() == "ga" and not ga.done() and ga.generation() == 0
    assert ga.best_individual() == [] and ga.best_score() == 0
    assert ga.stats().generations == 0
    ga.step()   # before initialize: no-op, as the reference
    assert ga.generation() == 0
    assert "population_size=16" in repr(ga) and ga.enable_catastrophe
    m = bmm.BitMatrix(3, 0)
    with pytest.raises(ValueError, match="empty window is not supported"):
        ga.initialize(m.row_window([0, 1, 2]))
    m = bmm.matrix_from_rows(["1100", "0110"])
    with pytest.raises(ValueError, match="population_size must be greater than zero"):
        bmm.GeneticAlgorithm(population_size=0).initialize(m.row_window([0, 1]))
    big = bmm.BitMatrix(257, 8)
    with pytest.raises(ValueError, match="at most 256 rows"):
        ga.initialize(big.row_window(list(range(257))))


def test_island_model_validation():
    spec = bmm.IslandSpec()
    with pytest.raises(ValueError, match="requires at least one island"):
        bmm.IslandModel(islands=[])
    with pytest.raises(ValueError, match="channel_capacity must be greater than zero"):
        bmm.IslandModel(islands=[spec], channel_capacity=0)
    with pytest.raises(ValueError, match="seed_stride must be non-negative"):
        bmm.IslandModel(islands=[spec], seed_stride=-1)
    with pytest.raises(ValueError, match="ga_configs length must match"):
        bmm.IslandModel(islands=[spec], ga_configs=[])
    with pytest.raises(ValueError, match="not both"):
        bmm.IslandModel(islands=[spec], ga_configs=[bmm.ga.GeneticAlgorithmConfig()],
                        population_size=10)
    with pytest.raises(ValueError, match="Unsupported overflow_policy"):
        bmm.IslandModel(islands=[spec], overflow_policy="sideways")
    m = bmm.IslandModel(islands=[spec, spec], overflow_policy="DROP_NEWEST", channel_capacity=8)
    assert m.overflow_policy is bmm.OverflowPolicy.DropNewest
    assert m.name() == "island_model" and m.island_count == 2
    assert "channel_capacity=8" in repr(m)
    snap = m.snapshot()
    assert snap.island_count == 2 and not snap.running and snap.best_individual == []
    with pytest.raises(RuntimeError, match="initialize"):
        m.start()
    hetero = [bmm.ga.GeneticAlgorithmConfig(), bmm.ga.GeneticAlgorithmConfig()]
    hetero[1].population_size = 12
    model = bmm.IslandModel(islands=[spec, spec], ga_configs=hetero)
    with pytest.raises(ValueError, match="share population_size"):
        model.initialize(bmm.matrix_from_rows(["1100", "0110"]).row_window([0, 1]))


def test_combination_masks_reconstruct_rows():
    rng = np.random.default_rng(5)
    for n, w, rank in ((6, 2, 6), (9, 3, 4), (5, 1, 1), (12, 4, 12)):
        base = rng.integers(0, 2**63, size=(rank, w), dtype=np.uint64)
        mix = rng.integers(0, 2, size=(n, rank)).astype(bool)
        mix[:rank] |= np.eye(rank, dtype=bool)[:min(n, rank)]
        rows = np.zeros((n, w), np.uint64)
        for i in range(n):
            for j in np.flatnonzero(mix[i]):
                rows[i] ^= base[j]
        targets = rows[rng.permutation(n)] ^ rows[rng.permutation(n)]
        masks = G._combination_masks(targets, rows)
        for t, mk in zip(targets, masks):
            acc = np.zeros(w, np.uint64)
            for j in range(n):
                if (int(mk[j // 64]) >> (j % 64)) & 1:
                    acc ^= rows[j]
            assert np.array_equal(acc, t)
    with pytest.raises(RuntimeError, match="row space"):
        G._combination_masks(np.array([[1]], np.uint64), np.array([[2]], np.uint64))


def test_unsupported_breeding_settings_are_refused():
    """Settings the GPU breeding step cannot take raise ValueError before any
    device work, instead of being ignored."""
    w = bmm.matrix_from_rows(["1100", "0110", "0011"]).row_window([0, 1, 2])
    for kw, msg in (({"num_parents": 5}, "num_parents in \\[1, 4\\]"),
                    ({"num_parents": 0}, "num_parents in \\[1, 4\\]"),
                    ({"num_offspring": 0}, "num_offspring must be at least 1"),
                    ({"mutation_rate": -0.5}, "mutation_rate must be in"),
                    ({"enable_catastrophe": True, "catastrophe_survival_rate": 1.5},
                     "catastrophe_survival_rate"),
                    ({"population_size": 2000}, "at most 1024")):
        with pytest.raises(ValueError, match=msg):
            bmm.GeneticAlgorithm(**kw).initialize(w)
    spec = bmm.IslandSpec()
    cfgs = [bmm.ga.GeneticAlgorithmConfig(), bmm.ga.GeneticAlgorithmConfig()]
    cfgs[1].mutation_rate = 0.5
    with pytest.raises(ValueError, match="share population_size"):
        bmm.IslandModel(islands=[spec, spec], ga_configs=cfgs).initialize(w)
    cfgs[1].mutation_rate = 0.3
    cfgs[1].num_offspring = 2
    with pytest.raises(ValueError, match="share population_size"):
        bmm.IslandModel(islands=[spec, spec], ga_configs=cfgs).initialize(w)

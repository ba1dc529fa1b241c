"""Genetic optimizers with the reference's Python API (bmmpy/ga.py:1-446,
python/bindings/ga.cpp:82-198) on the GPU island GA (bmm_hip_ga_*).

The names, keyword arguments, defaults, record types and lifecycle
(initialize / step / done / optimize; start / request_stop / wait /
snapshot / run_to_completion) are the reference's.  The search itself is the
GPU GA: op-sequence individuals over island bases (an individual is a list
of row XORs r_dst ^= r_src, its fitness the exact total weight after replay),
exact local improvement and information-set trials every epoch, ring
migration.  The reference's operator settings drive the GPU breeding step:
elitism (the elite_count best), tournaments of tournament_size, num_parents
parents per crossover and num_offspring children per parent set (op-list
crossover), mutation as max(1, floor(rows * mutation_rate)) row XORs per
child with the reference's rate adaptation, restarts after 15 stale
generations with probability 0.2, the catastrophe when enabled, per-island
seeds, and export / import counts (zero turns an island's export or import
off).  In a ring each island receives at most one migrant per interval (its
predecessor's best, the one individual the reference exports), so
channel_capacity never binds and the overflow policy never applies; both are
validated as in the reference.  Values the GPU GA cannot take raise
ValueError (num_parents outside [1, 4], num_offspring < 1, a negative
mutation_rate, islands that differ in their breeding settings).

best_individual() is a list of one Candidate per window row: candidate i's
mask selects the window rows whose XOR is output row i of the best transform
found, its weight that row's weight.
"""
from __future__ import annotations

import enum
import threading
from collections.abc import Sequence

import numpy as np

from . import _lib
from ._bmmpy import Candidate, RowWindow


class OverflowPolicy(enum.Enum):
    """Migration channel overflow behaviour (migration/channel.hpp)."""
    DropOldest = 0
    DropNewest = 1
    RejectPublish = 2


class StopCriteria:
    """Stopping rules for the genetic optimizers."""
    __slots__ = ("max_generations", "max_stale_generations", "target_total_weight")

    def __init__(self) -> None:
        self.max_generations: int | None = None
        self.max_stale_generations: int | None = None
        self.target_total_weight: int | None = None


class RunStats:
    """Runtime counters collected during a genetic run."""
    __slots__ = ("generations", "evaluations", "stale_generations", "seed")

    def __init__(self) -> None:
        self.generations = 0
        self.evaluations = 0
        self.stale_generations = 0
        self.seed = 0


class MigrationPolicy:
    """Per-island migration settings."""
    __slots__ = ("interval_generations", "export_count", "import_count", "shared_pool_capacity")

    def __init__(self) -> None:
        self.interval_generations = 32
        self.export_count = 2
        self.import_count = 2
        self.shared_pool_capacity = 128


class IslandSpec:
    """Static configuration for one island in the island model."""
    __slots__ = ("island_id", "migration")

    def __init__(self) -> None:
        self.island_id = 0
        self.migration = MigrationPolicy()


class IslandSnapshot:
    """Per-island runtime snapshot."""
    __slots__ = ("island_id", "running", "stop_requested", "finished", "stats", "best_score",
                 "best_individual")

    def __init__(self) -> None:
        self.island_id = 0
        self.running = False
        self.stop_requested = False
        self.finished = False
        self.stats = RunStats()
        self.best_score = 0
        self.best_individual: list[Candidate] = []


class IslandModelSnapshot:
    """Whole-model runtime snapshot."""
    __slots__ = ("running", "stop_requested", "island_count", "total_generations", "best_score",
                 "best_individual", "islands")

    def __init__(self) -> None:
        self.running = False
        self.stop_requested = False
        self.island_count = 0
        self.total_generations = 0
        self.best_score = 0
        self.best_individual: list[Candidate] = []
        self.islands: list[IslandSnapshot] = []


class GeneticAlgorithmConfig:
    """Low-level GA configuration (include/bmmpy/ga/genetic_algorithm.hpp:17-29)."""
    __slots__ = ("population_size", "elite_count", "num_parents", "num_offspring",
                 "tournament_size", "enable_catastrophe", "catastrophe_threshold",
                 "catastrophe_survival_rate", "mutation_rate", "stop", "seed")

    def __init__(self) -> None:
        self.population_size = 300
        self.elite_count = 3
        self.num_parents = 2
        self.num_offspring = 1
        self.tournament_size = 3
        self.enable_catastrophe = False
        self.catastrophe_threshold = 30
        self.catastrophe_survival_rate = 0.2
        self.mutation_rate = 0.3
        self.stop = StopCriteria()
        self.seed = 0


class IslandModelConfig:
    """Low-level island-model configuration."""
    __slots__ = ("islands",)

    def __init__(self) -> None:
        self.islands: list[IslandSpec] = []


def _build_stop_criteria(max_generations, max_stale_generations, target_total_weight):
    stop = StopCriteria()
    stop.max_generations = max_generations
    stop.max_stale_generations = max_stale_generations
    stop.target_total_weight = target_total_weight
    return stop


def _build_ga_config(*, population_size, elite_count, tournament_size, num_parents, num_offspring,
                     enable_catastrophe, catastrophe_threshold, catastrophe_survival_rate,
                     mutation_rate, max_generations, max_stale_generations, target_total_weight,
                     seed) -> GeneticAlgorithmConfig:
    c = GeneticAlgorithmConfig()
    c.population_size = population_size
    c.elite_count = elite_count
    c.tournament_size = tournament_size
    c.num_parents = num_parents
    c.num_offspring = num_offspring
    c.enable_catastrophe = enable_catastrophe
    c.catastrophe_threshold = catastrophe_threshold
    c.catastrophe_survival_rate = catastrophe_survival_rate
    c.mutation_rate = mutation_rate
    c.stop = _build_stop_criteria(max_generations, max_stale_generations, target_total_weight)
    c.seed = seed
    return c


def _clone_ga_config(config: GeneticAlgorithmConfig) -> GeneticAlgorithmConfig:
    c = GeneticAlgorithmConfig()
    for name in GeneticAlgorithmConfig.__slots__:
        setattr(c, name, getattr(config, name))
    c.stop = _build_stop_criteria(config.stop.max_generations, config.stop.max_stale_generations,
                                  config.stop.target_total_weight)
    return c


def _resolve_overflow_policy(policy) -> OverflowPolicy:
    if isinstance(policy, str):
        names = {"drop_oldest": OverflowPolicy.DropOldest,
                 "drop_newest": OverflowPolicy.DropNewest,
                 "reject_publish": OverflowPolicy.RejectPublish}
        key = policy.strip().lower()
        if key in names:
            return names[key]
        raise ValueError(f"Unsupported overflow_policy: {policy!r}")
    if not isinstance(policy, OverflowPolicy):
        raise ValueError(f"Unsupported overflow_policy: {policy!r}")
    return policy


def _breeding_key(c: GeneticAlgorithmConfig) -> tuple:
    return (c.population_size, c.elite_count, c.tournament_size, c.num_parents, c.num_offspring,
            float(c.mutation_rate), bool(c.enable_catastrophe),
            int(c.catastrophe_threshold) if c.enable_catastrophe else 0,
            float(c.catastrophe_survival_rate) if c.enable_catastrophe else 0.0)


def _check_gpu_config(configs: list[GeneticAlgorithmConfig]) -> None:
    c0 = configs[0]
    if any(_breeding_key(c) != _breeding_key(c0) for c in configs):
        raise ValueError("IslandModel: the GPU islands share population_size, elite_count, "
                         "tournament_size, num_parents, num_offspring, mutation_rate and the "
                         "catastrophe settings")
    if c0.population_size > 1024:
        raise ValueError("GeneticAlgorithm: population_size must be at most 1024")
    if not 1 <= c0.num_parents <= 4:
        raise ValueError("GeneticAlgorithm: the GPU GA takes num_parents in [1, 4]")
    if c0.num_offspring < 1:
        raise ValueError("GeneticAlgorithm: num_offspring must be at least 1")
    if not 0.0 <= float(c0.mutation_rate) <= 1000.0:
        raise ValueError("GeneticAlgorithm: mutation_rate must be in [0, 1000]")
    if c0.enable_catastrophe and not 0.0 <= float(c0.catastrophe_survival_rate) <= 1.0:
        raise ValueError("GeneticAlgorithm: catastrophe_survival_rate must be in [0, 1]")
    if c0.enable_catastrophe and c0.catastrophe_threshold < 0:
        raise ValueError("GeneticAlgorithm: catastrophe_threshold must be non-negative")


def _should_stop(cfg: GeneticAlgorithmConfig, st: RunStats, best: int) -> bool:
    # genetic_algorithm_internal.hpp:22-31.
    s = cfg.stop
    if s.max_generations is not None and st.generations >= s.max_generations:
        return True
    if s.max_stale_generations is not None and st.stale_generations >= s.max_stale_generations:
        return True
    if s.target_total_weight is not None and best <= s.target_total_weight:
        return True
    return False


# ------------------------------------------------------------------ engine --
_GPU_COLS = 2048   # the GPU GA works on column slices of 2048


def _window_words(window: RowWindow) -> tuple[np.ndarray, int]:
    """The window's rows as uint64 words, tail bits cleared."""
    mat = window.materialize()
    n, cols = len(window), window.cols
    w = (cols + 63) // 64
    words = np.asarray(mat.to_words(), np.uint64).reshape(n, -1)[:, :w].copy()
    if cols % 64 and w:
        words[:, -1] &= np.uint64((1 << (cols % 64)) - 1)
    return words, cols


def _popcount(words: np.ndarray) -> np.ndarray:
    return np.unpackbits(words.view(np.uint8), axis=-1).sum(axis=-1, dtype=np.int64)


def _combination_masks(targets: np.ndarray, basis: np.ndarray) -> np.ndarray:
    """GF(2): for each target row, a mask over the basis rows whose XOR is
    that row (targets lie in the row space of basis)."""
    n = basis.shape[0]
    mw = (n + 63) // 64
    rows = basis.copy()
    comb = np.zeros((n, mw), np.uint64)
    for i in range(n):
        comb[i, i // 64] = np.uint64(1) << np.uint64(i % 64)
    pivots = []   # (row index, word, bit)
    used = np.zeros(n, bool)
    for wd in range(rows.shape[1]):
        for b in range(64):
            bit = np.uint64(1) << np.uint64(b)
            has = ((rows[:, wd] & bit) != 0) & ~used
            idx = np.flatnonzero(has)
            if idx.size == 0:
                continue
            p = idx[0]
            used[p] = True
            others = np.flatnonzero((rows[:, wd] & bit) != 0)
            others = others[others != p]
            rows[others] ^= rows[p]
            comb[others] ^= comb[p]
            pivots.append((p, wd, bit))
            if len(pivots) == n:
                break
        if len(pivots) == n:
            break
    t = targets.copy()
    out = np.zeros((t.shape[0], mw), np.uint64)
    for p, wd, bit in pivots:
        sel = (t[:, wd] & bit) != 0
        t[sel] ^= rows[p]
        out[sel] ^= comb[p]
    if t.any():
        raise RuntimeError("GeneticAlgorithm: result left the window's row space")
    return out


def _candidates(best: np.ndarray, basis: np.ndarray) -> list[Candidate]:
    masks = _combination_masks(best, basis)
    weights = _popcount(best)
    return [Candidate([int(v) for v in masks[i]], int(weights[i])) for i in range(best.shape[0])]


class _Run:
    """One GPU GA over a window: `islands` islands with the given configs."""

    def __init__(self, window: RowWindow, configs: list[GeneticAlgorithmConfig],
                 migrate_every: int, migrate_flags: list[int] | None = None) -> None:
        n, cols = len(window), window.cols
        if n == 0 or cols == 0:
            raise ValueError("GeneticAlgorithm::initialize: empty window is not supported")
        for c in configs:
            if c.population_size == 0:
                raise ValueError(
                    "GeneticAlgorithm::initialize: population_size must be greater than zero")
        if n > 256:
            raise ValueError("GeneticAlgorithm: the GPU GA takes windows of at most 256 rows")
        _check_gpu_config(configs)
        c0 = configs[0]
        self.words, self.cols = _window_words(window)
        self.n = n
        self.configs = configs
        self.start_weight = int(_popcount(self.words).sum())
        self.ga = None
        if n >= 2:
            pad = -(-cols // _GPU_COLS) * _GPU_COLS
            m = np.zeros((n, pad // 64), np.uint64)
            m[:, :self.words.shape[1]] = self.words
            pop = c0.population_size
            # A rate of 0 still mutates once per child and adapts to 0.3 at the
            # first step (genetic_algorithm_core.cpp:33-41); 0 itself selects the
            # GPU's own mutation draws, so the smallest positive rate stands in.
            rate = max(float(c0.mutation_rate), 1e-30)
            p = _lib.ga_default_params(
                islands=len(configs), population=pop,
                tournament=max(1, c0.tournament_size), elite=min(c0.elite_count, pop),
                migrate_every=max(1, migrate_every), seed=c0.seed & ((1 << 64) - 1),
                num_parents=c0.num_parents, num_offspring=min(c0.num_offspring, pop),
                mutation_rate=rate,
                catastrophe_threshold=c0.catastrophe_threshold if c0.enable_catastrophe else 0,
                catastrophe_survival=float(c0.catastrophe_survival_rate),
                restart_after=15, restart_p=0.2)
            self.ga = _lib.GpuGA(m, pad, p)
            self.ga.configure_islands(
                seeds=[c.seed & ((1 << 64) - 1) for c in configs],
                migrate=migrate_flags if migrate_flags is not None else [3] * len(configs))
        self.generations = 0
        self.best_score = self.start_weight
        self.stale = 0

    def step(self, generations: int = 1) -> None:
        if self.ga is not None:
            st = self.ga.run(generations)
            best = int(st.best_weight)
        else:
            best = self.best_score
        self.generations += generations
        if best < self.best_score:
            self.best_score = best
            self.stale = 0
        else:
            self.stale += generations

    def evaluations(self, island: int = 0) -> int:
        return self.configs[island].population_size * (self.generations + 1)

    def stats(self, island: int = 0) -> RunStats:
        s = RunStats()
        s.generations, s.evaluations = self.generations, self.evaluations(island)
        s.stale_generations, s.seed = self.stale, self.configs[island].seed
        return s

    def best(self, island: int = -1) -> tuple[list[Candidate], int]:
        if self.ga is None:
            return _candidates(self.words, self.words), self.start_weight
        m, w = self.ga.best() if island < 0 else self.ga.island_best(island)
        rows = m[:, :self.words.shape[1]]
        return _candidates(rows, self.words), int(w)

    def close(self) -> None:
        if self.ga is not None:
            self.ga.close()
            self.ga = None


# --------------------------------------------------------------- public --
class GeneticAlgorithm:
    """
    Genetic optimizer over a row window.

    The returned value is a full list of Candidate objects with length equal to
    the input window size. Each candidate describes one output row of the found
    transform.
    """

    __slots__ = ("population_size", "elite_count", "tournament_size", "num_parents",
                 "num_offspring", "enable_catastrophe", "catastrophe_threshold",
                 "catastrophe_survival_rate", "mutation_rate", "max_generations",
                 "max_stale_generations", "target_total_weight", "seed", "_config", "_run",
                 "_done", "_best")

    def __init__(self, *, population_size: int = 300, elite_count: int = 3,
                 tournament_size: int = 3, num_parents: int = 2, num_offspring: int = 1,
                 enable_catastrophe: bool = False, catastrophe_threshold: int = 30,
                 catastrophe_survival_rate: float = 0.2, mutation_rate: float = 0.3,
                 max_generations: int | None = None, max_stale_generations: int | None = None,
                 target_total_weight: int | None = None, seed: int = 0) -> None:
        self.population_size = population_size
        self.elite_count = elite_count
        self.tournament_size = tournament_size
        self.num_parents = num_parents
        self.num_offspring = num_offspring
        self.enable_catastrophe = enable_catastrophe
        self.catastrophe_threshold = catastrophe_threshold
        self.catastrophe_survival_rate = catastrophe_survival_rate
        self.mutation_rate = mutation_rate
        self.max_generations = max_generations
        self.max_stale_generations = max_stale_generations
        self.target_total_weight = target_total_weight
        self.seed = seed
        self._config = _build_ga_config(
            population_size=population_size, elite_count=elite_count,
            tournament_size=tournament_size, num_parents=num_parents,
            num_offspring=num_offspring, enable_catastrophe=enable_catastrophe,
            catastrophe_threshold=catastrophe_threshold,
            catastrophe_survival_rate=catastrophe_survival_rate, mutation_rate=mutation_rate,
            max_generations=max_generations, max_stale_generations=max_stale_generations,
            target_total_weight=target_total_weight, seed=seed)
        self._run: _Run | None = None
        self._done = False
        self._best: list[Candidate] = []

    def __repr__(self) -> str:
        return (f"GeneticAlgorithm(population_size={self.population_size}, "
                f"elite_count={self.elite_count}, tournament_size={self.tournament_size}, "
                f"num_parents={self.num_parents}, num_offspring={self.num_offspring}, "
                f"enable_catastrophe={self.enable_catastrophe}, "
                f"catastrophe_threshold={self.catastrophe_threshold}, "
                f"catastrophe_survival_rate={self.catastrophe_survival_rate}, "
                f"mutation_rate={self.mutation_rate}, max_generations={self.max_generations}, "
                f"max_stale_generations={self.max_stale_generations}, "
                f"target_total_weight={self.target_total_weight}, seed={self.seed})")

    def name(self) -> str:
        return "ga"

    def initialize(self, window: RowWindow) -> None:
        if self._run is not None:
            self._run.close()
        self._run = _Run(window, [self._config], 20)
        self._best = []
        self._done = _should_stop(self._config, self._run.stats(), self._run.best_score)

    def step(self) -> None:
        if self._run is None or self._done:
            return
        self._run.step(1)
        self._best = []
        self._done = _should_stop(self._config, self._run.stats(), self._run.best_score)

    def done(self) -> bool:
        return self._done

    def generation(self) -> int:
        return 0 if self._run is None else self._run.generations

    def best_individual(self) -> list[Candidate]:
        if self._run is None:
            return []
        if not self._best:
            self._best, w = self._run.best()
            self._run.best_score = min(self._run.best_score, w)
        return list(self._best)

    def stats(self) -> RunStats:
        return RunStats() if self._run is None else self._run.stats()

    def best_score(self) -> int:
        if self._run is None:
            return 0
        return sum(c.weight for c in self.best_individual())

    def optimize(self, window: RowWindow) -> list[Candidate]:
        self.initialize(window)
        while not self.done():
            self.step()
        return self.best_individual()


class IslandModel:
    """
    Island-model wrapper around GeneticAlgorithm.

    By default this class builds one GA configuration per island from the common
    keyword arguments. For heterogeneous islands you can pass ga_configs with
    one GeneticAlgorithmConfig per IslandSpec.
    """

    __slots__ = ("island_count", "channel_capacity", "overflow_policy", "_specs", "_configs",
                 "_run", "_thread", "_lock", "_stop", "_running", "_finished")

    def __init__(self, *, islands: Sequence[IslandSpec],
                 ga_configs: Sequence[GeneticAlgorithmConfig] | None = None,
                 population_size: int = 300, elite_count: int = 3, tournament_size: int = 3,
                 num_parents: int = 2, num_offspring: int = 1, enable_catastrophe: bool = False,
                 catastrophe_threshold: int = 30, catastrophe_survival_rate: float = 0.2,
                 mutation_rate: float = 0.3, max_generations: int | None = None,
                 max_stale_generations: int | None = None,
                 target_total_weight: int | None = None, seed: int = 0, seed_stride: int = 1,
                 channel_capacity: int = 256, overflow_policy="drop_oldest") -> None:
        island_list = list(islands)
        if not island_list:
            raise ValueError("IslandModel requires at least one island")
        if channel_capacity <= 0:
            raise ValueError("channel_capacity must be greater than zero")
        if seed_stride < 0:
            raise ValueError("seed_stride must be non-negative")
        common = dict(population_size=population_size, elite_count=elite_count,
                      tournament_size=tournament_size, num_parents=num_parents,
                      num_offspring=num_offspring, enable_catastrophe=enable_catastrophe,
                      catastrophe_threshold=catastrophe_threshold,
                      catastrophe_survival_rate=catastrophe_survival_rate,
                      mutation_rate=mutation_rate, max_generations=max_generations,
                      max_stale_generations=max_stale_generations,
                      target_total_weight=target_total_weight)
        if ga_configs is not None:
            if len(ga_configs) != len(island_list):
                raise ValueError("ga_configs length must match the number of islands")
            defaults = dict(population_size=300, elite_count=3, tournament_size=3, num_parents=2,
                            num_offspring=1, enable_catastrophe=False, catastrophe_threshold=30,
                            catastrophe_survival_rate=0.2, mutation_rate=0.3,
                            max_generations=None, max_stale_generations=None,
                            target_total_weight=None)
            if common != defaults or seed != 0 or seed_stride != 1:
                raise ValueError("Pass either ga_configs or common GA keyword arguments, not both")
            configs = [_clone_ga_config(c) for c in ga_configs]
        else:
            configs = [_build_ga_config(**common, seed=seed + i * seed_stride)
                       for i in range(len(island_list))]
        self.island_count = len(island_list)
        self.channel_capacity = channel_capacity
        self.overflow_policy = _resolve_overflow_policy(overflow_policy)
        self._specs = island_list
        self._configs = configs
        self._run: _Run | None = None
        self._thread: threading.Thread | None = None
        self._lock = threading.Lock()
        self._stop = False
        self._running = False
        self._finished = False

    def __repr__(self) -> str:
        return (f"IslandModel(island_count={self.island_count}, "
                f"channel_capacity={self.channel_capacity}, "
                f"overflow_policy={self.overflow_policy})")

    def name(self) -> str:
        return "island_model"

    def _all_done(self) -> bool:
        st = self._run.stats()
        return all(_should_stop(c, st, self._run.best_score) for c in self._configs)

    def initialize(self, window: RowWindow) -> None:
        if self._running:
            raise RuntimeError("IslandModel::initialize: model is already running")
        if self._run is not None:
            self._run.close()
        # Epochs (migration, base folding, local improvement) run at the
        # smallest positive interval; an island with interval 0 or a zero
        # export_count exports nothing, one with a zero import_count imports
        # nothing (src/ga/island.cpp:149-228).
        intervals = [int(s.migration.interval_generations) for s in self._specs]
        positive = [i for i in intervals if i > 0]
        interval = min(positive) if positive else 20
        flags = [(1 if (iv > 0 and s.migration.export_count > 0) else 0)
                 | (2 if s.migration.import_count > 0 else 0)
                 for iv, s in zip(intervals, self._specs)]
        self._run = _Run(window, self._configs, interval, flags)
        self._stop = False
        self._finished = self._all_done()

    def _loop(self) -> None:
        try:
            while not self._stop and not self._all_done():
                with self._lock:
                    self._run.step(1)
        finally:
            self._finished = True

    def start(self) -> None:
        # src/ga/island_model.cpp:56-72: running stays set until wait().
        if self._run is None:
            raise RuntimeError("IslandModel::start: model must be initialized first")
        if self._running:
            raise RuntimeError("IslandModel::start: model is already running")
        self._running = True
        self._stop = False
        self._finished = False
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def request_stop(self) -> None:
        self._stop = True

    def wait(self) -> None:
        if self._thread is not None:
            self._thread.join()
            self._thread = None
        self._running = False

    def running(self) -> bool:
        return self._running

    def stop_requested(self) -> bool:
        return self._stop

    def best_individual(self) -> list[Candidate]:
        if self._run is None:
            return []
        with self._lock:
            return self._run.best()[0]

    def best_score(self) -> int:
        return self.snapshot().best_score

    def snapshot(self) -> IslandModelSnapshot:
        snap = IslandModelSnapshot()
        snap.running = self._running
        snap.stop_requested = self._stop
        snap.island_count = self.island_count
        if self._run is None:
            return snap
        with self._lock:
            snap.total_generations = self._run.generations * self.island_count
            snap.best_individual = self._run.best()[0]
            snap.best_score = sum(c.weight for c in snap.best_individual)
            for i, spec in enumerate(self._specs):
                isl = IslandSnapshot()
                isl.island_id = spec.island_id
                isl.running = self._running
                isl.stop_requested = self._stop
                isl.finished = self._finished
                isl.stats = self._run.stats(i)
                isl.best_individual = (self._run.best(i)[0] if self._run.ga is not None
                                       else snap.best_individual)
                isl.best_score = sum(c.weight for c in isl.best_individual)
                snap.islands.append(isl)
        return snap

    def run_to_completion(self, window: RowWindow) -> list[Candidate]:
        self.initialize(window)
        self.start()
        self.wait()
        return self.best_individual()


__all__ = [
    "GeneticAlgorithm", "GeneticAlgorithmConfig", "IslandModel", "IslandModelConfig",
    "IslandModelSnapshot", "IslandSnapshot", "IslandSpec", "MigrationPolicy", "OverflowPolicy",
    "RunStats", "StopCriteria",
]

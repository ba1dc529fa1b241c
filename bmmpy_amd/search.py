"""Keyword-only searcher wrappers (bmmpy/search.py:55-414).

Each wraps the native searcher of the same name; config structs stay hidden.
All searches run on the GPU.
"""
from __future__ import annotations

from ._bmmpy import (
    BruteforceSearch as _NativeBruteforceSearch,
    BruteforceSearchConfig as _NativeBruteforceSearchConfig,
    Candidate,
    CudaBruteforceSearch as _NativeCudaBruteforceSearch,
    CudaBruteforceSearchConfig as _NativeCudaBruteforceSearchConfig,
    CudaMitmFwhtSearch as _NativeCudaMitmFwhtSearch,
    CudaMitmFwhtSearchConfig as _NativeCudaMitmFwhtSearchConfig,
    FwhtSearch as _NativeFwhtSearch,
    FwhtSearchConfig as _NativeFwhtSearchConfig,
    MitmFwhtSearch as _NativeMitmFwhtSearch,
    MitmFwhtSearchConfig as _NativeMitmFwhtSearchConfig,
    RowWindow,
    CoolingPolicyKind as _NativeCoolingPolicyKind,
    SASelectionResult,
    SASelector as _NativeSASelector,
    SASelectorConfig as _NativeSASelectorConfig,
    WindowScorePolicyKind as _NativeWindowScorePolicyKind,
)


class BruteforceSearch:
    """Exact Gray-order brute force over all row subsets of a window."""

    __slots__ = ("max_candidates", "chunk_bits", "_impl")

    def __init__(self, *, max_candidates: int = 64, chunk_bits: int = 0) -> None:
        config = _NativeBruteforceSearchConfig()
        config.max_candidates = max_candidates
        config.chunk_bits = chunk_bits
        self.max_candidates = max_candidates
        self.chunk_bits = chunk_bits
        self._impl = _NativeBruteforceSearch(config)

    def __repr__(self) -> str:
        return (f"BruteforceSearch(max_candidates={self.max_candidates}, "
                f"chunk_bits={self.chunk_bits})")

    def name(self) -> str:
        return self._impl.name()

    def search(self, window: RowWindow) -> list[Candidate]:
        return self._impl.search(window)


class FwhtSearch:
    """Exact top-k via the Walsh-Hadamard transform of the column histogram.

    Raises ValueError when the window has more than ``max_rows`` rows.
    """

    __slots__ = ("max_rows", "max_candidates", "_impl")

    def __init__(self, *, max_rows: int = 16, max_candidates: int = 64) -> None:
        config = _NativeFwhtSearchConfig()
        config.max_rows = max_rows
        config.max_candidates = max_candidates
        self.max_rows = max_rows
        self.max_candidates = max_candidates
        self._impl = _NativeFwhtSearch(config)

    def __repr__(self) -> str:
        return f"FwhtSearch(max_rows={self.max_rows}, max_candidates={self.max_candidates})"

    def name(self) -> str:
        return self._impl.name()

    def search(self, window: RowWindow) -> list[Candidate]:
        window_size = len(window)
        if window_size > self.max_rows:
            raise ValueError(
                f"Window has {window_size} rows, which exceeds the configured "
                f"max_rows={self.max_rows} limit for this searcher.")
        return self._impl.search(window)


class MitmFwhtSearch:
    """Exact top-k via the meet-in-the-middle split transform."""

    __slots__ = ("max_candidates", "_impl")

    def __init__(self, *, max_candidates: int = 64) -> None:
        config = _NativeMitmFwhtSearchConfig()
        config.max_candidates = max_candidates
        self.max_candidates = max_candidates
        self._impl = _NativeMitmFwhtSearch(config)

    def __repr__(self) -> str:
        return f"MitmFwhtSearch(max_candidates={self.max_candidates})"

    def name(self) -> str:
        return self._impl.name()

    def search(self, window: RowWindow) -> list[Candidate]:
        return self._impl.search(window)


class CudaMitmFwhtSearch:
    """The reference's GPU split-FWHT searcher name and limits (t in [16, 64],
    max_candidates <= 128, total weight <= 32767), run by the HIP kernels."""

    __slots__ = ("max_candidates", "low_bits", "_impl")

    def __init__(self, *, max_candidates: int = 64, low_bits: int = 0) -> None:
        config = _NativeCudaMitmFwhtSearchConfig()
        config.max_candidates = max_candidates
        config.low_bits = low_bits
        self.max_candidates = max_candidates
        self.low_bits = low_bits
        self._impl = _NativeCudaMitmFwhtSearch(config)

    def __repr__(self) -> str:
        return (f"CudaMitmFwhtSearch(max_candidates={self.max_candidates}, "
                f"low_bits={self.low_bits})")

    def name(self) -> str:
        return self._impl.name()

    def search(self, window: RowWindow) -> list[Candidate]:
        return self._impl.search(window)


class CudaBruteforceSearch:
    """The reference's GPU brute-force searcher name and limits (t in [16, 64],
    widths 512/1024/4096, max_candidates <= 128), run by the HIP kernels."""

    __slots__ = ("max_candidates", "chunk_bits", "_impl")

    def __init__(self, *, max_candidates: int = 64, chunk_bits: int = 0) -> None:
        config = _NativeCudaBruteforceSearchConfig()
        config.max_candidates = max_candidates
        config.chunk_bits = chunk_bits
        self.max_candidates = max_candidates
        self.chunk_bits = chunk_bits
        self._impl = _NativeCudaBruteforceSearch(config)

    def __repr__(self) -> str:
        return (f"CudaBruteforceSearch(max_candidates={self.max_candidates}, "
                f"chunk_bits={self.chunk_bits})")

    def name(self) -> str:
        return self._impl.name()

    def search(self, window: RowWindow) -> list[Candidate]:
        return self._impl.search(window)


def _resolve_score_policy(name: str):
    if name == "pairwise_synergy":
        return _NativeWindowScorePolicyKind.PairwiseSynergy
    if name == "higher_order_synergy":
        return _NativeWindowScorePolicyKind.HigherOrderSynergy
    raise ValueError(f"Unsupported score_policy: {name!r}")


def _resolve_cooling_policy(name: str):
    if name == "adaptive_geometric":
        return _NativeCoolingPolicyKind.AdaptiveGeometric
    raise ValueError(f"Unsupported cooling_policy: {name!r}")


class SASelector:
    """Select a row window by simulated annealing (bmmpy/search.py:416-538 of
    the reference: same keyword arguments, policies and results).  The
    pairwise synergy table is built on the GPU; the annealing runs on the
    host with the reference's random stream, so a seed gives the same window.
    """

    __slots__ = ("iterations", "restarts", "seed", "score_policy", "cooling_policy",
                 "temperature_probe_samples", "initial_acceptance_probability", "cooling_rate",
                 "min_temperature", "_impl")

    def __init__(self, *, iterations: int = 10000, restarts: int = 8, seed: int = 0,
                 score_policy: str = "pairwise_synergy",
                 cooling_policy: str = "adaptive_geometric",
                 temperature_probe_samples: int = 64,
                 initial_acceptance_probability: float = 0.8, cooling_rate: float = 0.99,
                 min_temperature: float = 1e-6) -> None:
        config = _NativeSASelectorConfig()
        config.iterations = iterations
        config.restarts = restarts
        config.seed = seed
        config.score_policy = _resolve_score_policy(score_policy)
        config.cooling_policy = _resolve_cooling_policy(cooling_policy)
        config.temperature_probe_samples = temperature_probe_samples
        config.initial_acceptance_probability = initial_acceptance_probability
        config.cooling_rate = cooling_rate
        config.min_temperature = min_temperature
        self.iterations = iterations
        self.restarts = restarts
        self.seed = seed
        self.score_policy = score_policy
        self.cooling_policy = cooling_policy
        self.temperature_probe_samples = temperature_probe_samples
        self.initial_acceptance_probability = initial_acceptance_probability
        self.cooling_rate = cooling_rate
        self.min_temperature = min_temperature
        self._impl = _NativeSASelector(config)

    def __repr__(self) -> str:
        return (f"SASelector(iterations={self.iterations}, restarts={self.restarts}, "
                f"seed={self.seed}, score_policy={self.score_policy!r}, "
                f"cooling_policy={self.cooling_policy!r}, "
                f"temperature_probe_samples={self.temperature_probe_samples}, "
                f"initial_acceptance_probability={self.initial_acceptance_probability}, "
                f"cooling_rate={self.cooling_rate}, min_temperature={self.min_temperature})")

    def name(self) -> str:
        return self._impl.name()

    def select(self, matrix, window_size: int) -> SASelectionResult:
        return self._impl.select(matrix, window_size)

    def select_window(self, matrix, window_size: int) -> RowWindow:
        return self._impl.select_window(matrix, window_size)


__all__ = ["BruteforceSearch", "CudaBruteforceSearch", "CudaMitmFwhtSearch", "FwhtSearch",
           "MitmFwhtSearch", "SASelectionResult", "SASelector"]

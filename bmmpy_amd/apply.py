"""GreedyApplier wrapper (bmmpy/apply.py:46-131)."""
from __future__ import annotations

from typing import Iterable

from ._bmmpy import ApplyResult, Candidate, RowWindow
from ._bmmpy import GreedyApplier as _NativeGreedyApplier


class GreedyApplier:
    """Apply candidates in nondecreasing weight order, each replacing the row
    with the largest gain; the resulting row XORs are replayed on the GPU."""

    __slots__ = ("min_gain", "stochastic", "seed", "_impl")

    def __init__(self, min_gain: int, *, stochastic: bool = False, seed: int = 0) -> None:
        self.min_gain = min_gain
        self.stochastic = stochastic
        self.seed = seed
        self._impl = _NativeGreedyApplier(min_gain, stochastic, seed)

    def __repr__(self) -> str:
        return (f"GreedyApplier(min_gain={self.min_gain}, stochastic={self.stochastic}, "
                f"seed={self.seed})")

    def apply(self, window: RowWindow, candidates: Iterable[Candidate]) -> ApplyResult:
        return self._impl.apply(window, list(candidates))


__all__ = ["GreedyApplier"]

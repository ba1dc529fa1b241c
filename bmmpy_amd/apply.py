"""GreedyApplier and GlobalGreedyApplier wrappers (bmmpy/apply.py:46-202)."""
from __future__ import annotations

from typing import Iterable

from ._bmmpy import ApplyResult, Candidate, RowWindow
from ._bmmpy import GlobalGreedyApplier as _NativeGlobalGreedyApplier
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


class GlobalGreedyApplier:
    """Assemble one full invertible transform from the candidates plus the
    window's unit rows (stably sorted by weight, greedy pivot basis over the
    masks) and apply it in one GPU row combination when the total weight does
    not grow (strictly drops when ``require_improvement``)."""

    __slots__ = ("require_improvement", "_impl")

    def __init__(self, *, require_improvement: bool = True) -> None:
        self.require_improvement = require_improvement
        self._impl = _NativeGlobalGreedyApplier(require_improvement)

    def __repr__(self) -> str:
        return f"GlobalGreedyApplier(require_improvement={self.require_improvement})"

    def name(self) -> str:
        return self._impl.name()

    def apply(self, window: RowWindow, candidates: Iterable[Candidate]) -> ApplyResult:
        return self._impl.apply(window, list(candidates))


__all__ = ["GreedyApplier", "GlobalGreedyApplier"]

"""search_apply (bmmpy/workflows.py:33-90) and the structural types of its
searcher and applier arguments."""
from __future__ import annotations

from collections.abc import Iterable
from typing import Protocol

from ._bmmpy import ApplyResult, Candidate, RowWindow


class SupportsSearch(Protocol):
    """Anything with ``search(window) -> list[Candidate]``."""

    def search(self, window: RowWindow) -> list[Candidate]: ...


class SupportsApply(Protocol):
    """Anything with ``apply(window, candidates) -> ApplyResult``."""

    def apply(self, window: RowWindow, candidates: Iterable[Candidate]) -> ApplyResult: ...


def search_apply(window: RowWindow, *, searcher: SupportsSearch,
                 applier: SupportsApply) -> ApplyResult:
    """Run ``searcher.search(window)`` then ``applier.apply(window, candidates)``."""
    candidates = searcher.search(window)
    return applier.apply(window, candidates)


__all__ = ["SupportsApply", "SupportsSearch", "search_apply"]

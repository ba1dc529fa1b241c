"""search_apply (bmmpy/workflows.py:45-90)."""
from __future__ import annotations


def search_apply(window, *, searcher, applier):
    """Run ``searcher.search(window)`` then ``applier.apply(window, candidates)``."""
    candidates = searcher.search(window)
    return applier.apply(window, candidates)


__all__ = ["search_apply"]

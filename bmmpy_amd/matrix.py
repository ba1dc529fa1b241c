"""BitMatrix / RowWindow helpers (bmmpy/matrix.py)."""
from __future__ import annotations

from typing import Sequence

from ._bmmpy import BitMatrix, RowWindow  # noqa: F401


def matrix_from_rows(rows: Sequence[str]) -> BitMatrix:
    """Build a BitMatrix from equal-width strings of 0 and 1 (ValueError otherwise)."""
    return BitMatrix.from_rows(list(rows))


__all__ = ["BitMatrix", "RowWindow", "matrix_from_rows"]

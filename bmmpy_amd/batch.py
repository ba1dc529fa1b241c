"""Batched coset-minimum inputs stored as BMMPYB1 files.

A batch is the packed layout the batched kernels read from HBM: each
matrix's rows of ceil(cols / 64) words back to back in one uint64 array,
``word_offset[i]`` = first word of matrix i, with ``rows[i]`` and
``cols[i]``.  On disk every matrix is one BMMPYB1 file
(src/core/bit_matrix_io.cpp:49-67: magic ``BMMPYB1\\0``, little-endian u64
rows and cols, unpadded row words), so a batch written by the reference's
``BitMatrix.save_binary_file`` loads here unchanged, and ``BitMatrix.hash``
identifies each matrix independently of padding.
"""
from __future__ import annotations

import os
from typing import Iterable, Sequence

import numpy as np

from ._bmmpy import BitMatrix


def load_batch(paths: Iterable[str | os.PathLike]):
    """Read BMMPYB1 files into (words, word_offset, rows, cols)."""
    chunks, offs, rows, cols = [], [], [], []
    at = 0
    for p in paths:
        m = BitMatrix.load_binary_file(p)
        w = np.asarray(m.to_words(), np.uint64).reshape(-1)
        offs.append(at)
        rows.append(m.rows)
        cols.append(m.cols)
        chunks.append(w)
        at += w.size
    words = np.concatenate(chunks) if chunks else np.zeros(0, np.uint64)
    return (words, np.array(offs, np.int64), np.array(rows, np.int32),
            np.array(cols, np.int32))


def save_batch(directory: str | os.PathLike, words, word_offset: Sequence[int],
               rows: Sequence[int], cols: Sequence[int]) -> list[str]:
    """Write each matrix of a packed batch as ``<directory>/mNNNNN.bmm``;
    returns the paths in batch order."""
    words = np.asarray(words, np.uint64)
    paths = []
    for i, (off, r, c) in enumerate(zip(word_offset, rows, cols)):
        w = (int(c) + 63) // 64
        m = BitMatrix.from_words(words[int(off):int(off) + int(r) * w].reshape(int(r), w), int(c))
        path = os.path.join(os.fspath(directory), f"m{i:05d}.bmm")
        m.save_binary_file(path)
        paths.append(path)
    return paths


def batch_hashes(words, word_offset: Sequence[int], rows: Sequence[int],
                 cols: Sequence[int]) -> list[str]:
    """Canonical SHA-256 of every matrix of a packed batch."""
    words = np.asarray(words, np.uint64)
    out = []
    for off, r, c in zip(word_offset, rows, cols):
        w = (int(c) + 63) // 64
        out.append(BitMatrix.from_words(words[int(off):int(off) + int(r) * w].reshape(int(r), w),
                                        int(c)).hash())
    return out


__all__ = ["load_batch", "save_batch", "batch_hashes"]

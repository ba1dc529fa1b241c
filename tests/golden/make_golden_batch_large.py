"""Golden coset minima for the large matrices of the batched uneven workload
(4096 matrices, seed 0x5EED0004: rows 8 + index(21), cols 64 (1 + index(32)),
density 0.1 + 0.4 unit()): the first matrix of each row count 21 .. 28, every
row against the matrix's other rows, computed by the oracle's coset minimum
(one histogram + int32 FWHT of 2^(rows-1) entries per target; pinned to the
reference FwhtSearch ranking and brute force in tests/test_oracle.py).

Run from the repository root:  python tests/golden/make_golden_batch_large.py
Writes tests/golden/golden_batch_large.npz (a few KiB).  Takes a few minutes
on 8 cores.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED_BATCH_4096 = 0x5EED0004
N_MAT = 4096


def _pick(rows):
    first = {}
    for m, r in enumerate(rows):
        if 21 <= r <= 28 and int(r) not in first:
            first[int(r)] = m
    return [first[r] for r in sorted(first)]


def _solve(args):
    import oracle
    mat, cols = args
    w, idx = oracle.oracle().coset_min(mat, cols)
    return np.asarray(w, np.uint32), np.asarray(idx, np.uint64)


def main() -> None:
    import oracle
    ora = oracle.oracle()
    rows, cols, dens, words, offs = ora.batch(SEED_BATCH_4096, N_MAT)
    picks = _pick(rows)
    jobs = []
    for m in picks:
        W = (int(cols[m]) + 63) // 64
        jobs.append((words[offs[m]:offs[m] + rows[m] * W].reshape(rows[m], W).copy(),
                     int(cols[m])))
    with mp.Pool(min(8, len(jobs))) as pool:
        res = pool.map(_solve, jobs)
    out = {"matrices": np.array(picks, np.int64),
           "rows": np.array([rows[m] for m in picks], np.int32),
           "cols": np.array([cols[m] for m in picks], np.int32),
           "weights": np.concatenate([r[0] for r in res]),
           "indices": np.concatenate([r[1] for r in res])}
    np.savez_compressed(os.path.join(HERE, "golden_batch_large.npz"), **out)
    print({k: v.tolist() for k, v in out.items() if k != "indices" and k != "weights"})


if __name__ == "__main__":
    main()

"""GPU parity of SASelector's pairwise policy (synergy table on the GPU,
annealing on the host) against the reference SASelector's outputs
(golden_sa.json, written from the reference build), and of the overlap
kernel against direct AND-popcounts."""
import numpy as np
import pytest

from test_api import _bitmatrix, _sa_cases

import bmmpy as bmm

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,cols,p", [(1, 64, 0.5), (5, 70, 0.5), (64, 1000, 0.2),
                                      (65, 2048, 0.3), (200, 333, 0.4), (300, 4100, 0.05)])
def test_pair_overlaps_match_direct(lib, ora, n, cols, p):
    _say("pair overlaps", n, cols)
    m = ora.bernoulli(900 + n, n, cols, p)
    got = lib.pair_overlaps(m)
    bits = np.unpackbits(m.view(np.uint8), bitorder="little").reshape(n, -1).astype(np.int64)
    exp = bits @ bits.T
    np.fill_diagonal(exp, 0)
    assert np.array_equal(got, exp.astype(np.uint32))


def test_sa_selector_pairwise_matches_reference_golden(lib):
    for case in _sa_cases("pairwise_synergy"):
        _say("sa pairwise", len(case["rows"]), case["window_size"], case["seed"])
        rows = [[int(v) for v in r] for r in case["rows"]]
        mat = _bitmatrix(rows, case["cols"])
        sel = bmm.SASelector(iterations=case["iterations"], restarts=case["restarts"],
                             seed=case["seed"])
        res = sel.select(mat, case["window_size"])
        e = case["expect"]
        assert list(res.rows) == e["rows"] and res.score == e["score"]
        assert (res.accepted_moves, res.iterations_run, res.best_iteration, res.restart_index,
                res.seed) == (e["accepted_moves"], e["iterations_run"], e["best_iteration"],
                              e["restart_index"], e["seed"])
        if case["window_size"]:
            win = sel.select_window(mat, case["window_size"])
            assert len(win) == case["window_size"]

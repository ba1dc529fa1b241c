"""The Python surface, mirroring the reference's tests/python/test_search_api.py,
test_cuda_api.py and test_public_api.py.  CPU tests cover names, argument
validation (raised before any device query) and the data structures; GPU
tests run the searchers and the applier against the oracle."""
import random

import numpy as np
import pytest

import bmmpy as bmm
import oracle

ROWS_6X12 = ["110010101001", "101100101100", "011010011001", "111000010111", "000111100011",
             "101011110000"]


def make_search_matrix():
    m = bmm.BitMatrix(2, 5)
    for col in (0, 2, 4):
        m[0, col] = True
        m[1, col] = True
    return m


def pairs(cands):
    return [(c.weight, c.mask_u64()) for c in cands]


def words_of(m: "bmm.BitMatrix"):
    return m.to_words()


# ------------------------------------------------------------------ CPU ----
def test_public_names():
    for name in ("BitMatrix", "RowWindow", "Candidate", "FwhtSearch", "MitmFwhtSearch",
                 "BruteforceSearch", "CudaMitmFwhtSearch", "CudaBruteforceSearch",
                 "GreedyApplier", "GlobalGreedyApplier", "ApplyResult", "search_apply",
                 "matrix_from_rows", "SASelector", "solver",
                 "get_runtime_features", "get_version", "fwht_i16", "fwht_i32", "MatrixError"):
        assert hasattr(bmm, name), name
    # config structs stay hidden (tests/python/test_public_api.py:9-26)
    for hidden in ("FwhtSearchConfig", "MitmFwhtSearchConfig", "BruteforceSearchConfig"):
        assert not hasattr(bmm, hidden)


def test_global_greedy_applier_surface():
    ap = bmm.GlobalGreedyApplier()
    assert ap.name() == "global_greedy" and ap.require_improvement is True
    assert repr(bmm.GlobalGreedyApplier(require_improvement=False)) == \
        "GlobalGreedyApplier(require_improvement=False)"
    with pytest.raises(TypeError):
        bmm.GlobalGreedyApplier(False)   # keyword-only, as bmmpy/apply.py:152
    # An empty window returns before any device call.
    m = bmm.BitMatrix(3, 5)
    res = ap.apply(m.row_window([]), [])
    assert (res.applied_count, res.weight_improvement) == (0, 0)


def test_searcher_names_and_repr():
    assert bmm.FwhtSearch().name() == "fwht"
    assert bmm.MitmFwhtSearch().name() == "mitm_fwht"
    assert bmm.BruteforceSearch().name() == "bruteforce"
    assert bmm.CudaMitmFwhtSearch(max_candidates=8, low_bits=0).name() == "cuda_mitm_fwht"
    assert bmm.CudaBruteforceSearch(max_candidates=8, chunk_bits=0).name() == "cuda_bruteforce"
    assert repr(bmm.FwhtSearch(max_rows=3, max_candidates=2)) == \
        "FwhtSearch(max_rows=3, max_candidates=2)"


def test_runtime_features():
    f = bmm.get_runtime_features()
    assert f.hip_compiled and not f.cuda_compiled and not f.cuda_available
    assert f.hip_available == (f.device_count > 0)
    assert isinstance(f.avx2_available, bool) and not f.avx2_compiled


def test_bitmatrix_layout_and_ops():
    m = bmm.BitMatrix(3, 130)
    assert m.stride_words == 4 and m.words_per_row == 3 and m.shape == (3, 130)
    m[0, 0] = True
    m[0, 129] = True
    m[1, 129] = True
    m.row_xor(1, 0)
    assert m.row_popcount(1) == 1 and m[1, 0] and not m[1, 129]
    m.row_xor(1, 1)                          # no-op on itself
    assert m.row_popcount(1) == 1
    with pytest.raises(IndexError):
        m[3, 0]
    with pytest.raises(IndexError):
        m.row_window([0, 7])
    with pytest.raises(ValueError):
        m.row_window([0, 0])
    w = m.row_window([2, 0])
    assert len(w) == 2 and w.rows == [2, 0] and w.cols == 130
    assert w.materialize().to_rows()[1] == m.to_rows()[0]
    assert bmm.matrix_from_rows(["101", "010"]).to_rows() == ["101", "010"]
    with pytest.raises(ValueError):
        bmm.matrix_from_rows(["10", "1"])
    i3 = bmm.BitMatrix.identity(3)
    assert i3.rank() == 3   # products run on the GPU (tests/test_gpu_global_apply.py)
    with pytest.raises(bmm.MatrixError):   # shapes are checked before any device call
        i3.mul(bmm.BitMatrix(2, 3))


def test_row_window_metadata_and_assignment():
    """total_weight counts the columns any row of the window sets (an OR of
    the rows, src/core/row_window.cpp:67-88), not the sum of row weights;
    assign_materialized writes rows back in window order (:109-115)."""
    m = bmm.BitMatrix(4, 130)
    for r, c in [(0, 0), (0, 129), (1, 129), (1, 64), (2, 3), (3, 100)]:
        m[r, c] = True
    w = m.row_window([1, 0])
    assert [w.row_popcount(i) for i in range(2)] == [2, 2] and w.weight() == 4
    assert w.total_weight == 3                 # columns 0, 64, 129
    assert m.row_window([0, 1, 2, 3]).total_weight == 5
    assert m.row_window([]).total_weight == 0
    assert w.shape == (2, 130) and repr(w) == "RowWindow(size=2, cols=130)"
    new = bmm.BitMatrix(2, 130)
    new[0, 5] = True                           # goes to global row 1
    new[1, 7] = True                           # goes to global row 0
    w.assign_materialized(new)
    assert m.to_rows()[1][5] == "1" and m.row_popcount(1) == 1
    assert m.to_rows()[0][7] == "1" and m.row_popcount(0) == 1
    assert w.materialize().to_rows() == new.to_rows()
    assert m.row_popcount(2) == 1 and m.row_popcount(3) == 1   # untouched
    with pytest.raises(bmm.MatrixError):
        w.assign_materialized(bmm.BitMatrix(3, 130))
    with pytest.raises(bmm.MatrixError):
        w.assign_materialized(bmm.BitMatrix(2, 129))


def test_candidate_api():
    c = bmm.Candidate.from_u64(0b1011, 7)
    assert c.rows == [0, 1, 3] and len(c) == 3 and 3 in c and 2 not in c
    assert list(c) == [0, 1, 3] and sorted(iter(c)) == c.selected_rows()
    wide = bmm.Candidate.from_words([0b10, 1 << 5], 9)
    assert list(wide) == [1, 69] and wide.weight == 9 and 69 in wide and 68 not in wide
    big = bmm.Candidate([1, 2], 3)
    with pytest.raises(RuntimeError):
        big.mask_u64()


def test_validation_before_device():
    m = bmm.BitMatrix(3, 5)
    w3 = m.row_window([0, 1, 2])
    with pytest.raises(ValueError):
        bmm.FwhtSearch(max_rows=2).search(w3)
    assert bmm.FwhtSearch(max_candidates=0).search(w3) == []
    assert bmm.MitmFwhtSearch(max_candidates=0).search(w3) == []
    assert bmm.BruteforceSearch(max_candidates=0).search(w3) == []
    with pytest.raises(ValueError):
        bmm.BruteforceSearch(chunk_bits=4).search(w3)
    # GPU-class limits (tests/cpp/search_cuda_tests.cpp:29-70, 113-135)
    with pytest.raises(ValueError):
        bmm.CudaMitmFwhtSearch().search(w3)
    m28 = bmm.BitMatrix(28, 4)
    m28[0, 0] = True
    with pytest.raises(ValueError):
        bmm.CudaMitmFwhtSearch(max_candidates=64, low_bits=28).search(m28.row_window(list(range(28))))
    heavy = bmm.BitMatrix(28, 32768)
    for c in range(32768):
        heavy[0, c] = True
    with pytest.raises(ValueError):
        bmm.CudaMitmFwhtSearch().search(heavy.row_window(list(range(28))))
    eq15 = bmm.BitMatrix.from_words(oracle.equivalence_matrix(15, 512), 512)
    with pytest.raises(ValueError):
        bmm.CudaBruteforceSearch().search(eq15.row_window(list(range(15))))
    eq16 = bmm.BitMatrix.from_words(oracle.equivalence_matrix(16, 256), 256)
    with pytest.raises(ValueError):
        bmm.CudaBruteforceSearch().search(eq16.row_window(list(range(16))))
    eq16w = bmm.BitMatrix.from_words(oracle.equivalence_matrix(16, 512), 512)
    with pytest.raises(ValueError):
        bmm.CudaBruteforceSearch(max_candidates=129).search(eq16w.row_window(list(range(16))))


def test_no_device_raises_runtime_error():
    if bmm.get_runtime_features().hip_available:
        pytest.skip("a device is visible")
    w = make_search_matrix().row_window([0, 1])
    for s in (bmm.FwhtSearch(), bmm.MitmFwhtSearch(), bmm.BruteforceSearch()):
        with pytest.raises(RuntimeError):
            s.search(w)
    with pytest.raises(RuntimeError):
        bmm.fwht_i32([1, 2, 3, 4])


def test_compact_split_window_kat():
    """tests/cpp/search_cpu_tests.cpp:128-152."""
    from bmmpy_amd import _bmmpy
    m = bmm.matrix_from_rows(["11001", "00100", "01101", "01001"])
    c = _bmmpy.build_compact_split_window(m.row_window([0, 1, 2, 3]), 2)
    assert (c.t, c.low_bits, c.high_bits, c.total_weight) == (4, 2, 2, 4)
    assert c.q == [0, 1, 3] and c.r == [1, 2, 1] and c.multiplicity == [1, 1, 2]


# ------------------------------------------------------------------ GPU ----
@pytest.mark.gpu
def test_gpu_fwht_search_wrapper():
    w = make_search_matrix().row_window([0, 1])
    s = bmm.FwhtSearch(max_rows=16, max_candidates=1)
    c = s.search(w)
    assert len(c) == 1 and c[0].mask_u64() == 0x3 and c[0].weight == 0
    assert len(bmm.FwhtSearch().search(w)) == 3


@pytest.mark.gpu
def test_gpu_mitm_and_brute_match_fwht():
    m = bmm.matrix_from_rows(ROWS_6X12)
    w = m.row_window([0, 1, 2, 3, 4, 5])
    exp = pairs(bmm.FwhtSearch(max_rows=16, max_candidates=8).search(w))
    assert pairs(bmm.MitmFwhtSearch(max_candidates=8).search(w)) == exp
    assert pairs(bmm.BruteforceSearch(max_candidates=8, chunk_bits=0).search(w)) == exp


@pytest.mark.gpu
def test_gpu_search_apply():
    m = make_search_matrix()
    r = bmm.search_apply(m.row_window([0, 1]), searcher=bmm.FwhtSearch(max_rows=16,
                                                                        max_candidates=1),
                         applier=bmm.GreedyApplier(min_gain=1))
    assert r.applied_count >= 1 and r.weight_improvement > 0
    assert 0 in (m.row_popcount(0), m.row_popcount(1))


@pytest.mark.gpu
def test_gpu_cuda_named_searchers_match_cpu_semantics():
    rng = random.Random(42)
    m = bmm.BitMatrix(16, 512)
    for r in range(16):
        for c in range(512):
            if rng.getrandbits(1):
                m[r, c] = True
    w = m.row_window(list(range(16)))
    ora = oracle.oracle()
    exp = ora.bruteforce_topk(m.to_words(), 512, 8)
    assert pairs(bmm.CudaBruteforceSearch(max_candidates=8).search(w)) == exp
    e28 = bmm.BitMatrix.from_words(oracle.equivalence_matrix(28, 96), 96)
    w28 = e28.row_window(list(range(28)))
    for k in (8, 128):
        assert pairs(bmm.CudaMitmFwhtSearch(max_candidates=k).search(w28)) == \
            ora.mitm_topk(e28.to_words(), 96, k)


@pytest.mark.gpu
def test_gpu_greedy_apply_matches_oracle():
    ora = oracle.oracle()
    base = ora.bernoulli(32, 12, 256, 0.3)
    cands = ora.fwht_topk(base, 256, 40)
    for mg, st, sd in ((0, False, 0), (1, False, 0), (3, True, 7)):
        m = bmm.BitMatrix.from_words(base, 256)
        res = bmm.GreedyApplier(mg, stochastic=st, seed=sd).apply(
            m.row_window(list(range(12))), [bmm.Candidate.from_u64(mk, w) for w, mk in cands])
        exp_rows, ap, imp = ora.greedy_apply(base, 256, cands, mg, st, sd)
        assert (res.applied_count, res.weight_improvement) == (ap, imp)
        assert np.array_equal(m.to_words(), exp_rows)


@pytest.mark.gpu
def test_gpu_fwht_math_wrappers():
    assert bmm.fwht_i32([1, 2, 3, 4]).tolist() == [10, -2, -4, 0]
    assert bmm.fwht_i16([32767, 1]).tolist() == [-32768, 32766]


def test_exact_basis_caps_and_empty():
    """ExactBasisSolver's caps and early returns (src/solver/exact_basis_solver.cpp:
    58-96) are decided before any device work."""
    mat = bmm.matrix_from_rows(["1" * 10])
    empty = bmm.solver.ExactBasisSolver().solve(mat.row_window([]))
    assert empty.rank == 0 and empty.input_rows == 0 and empty.basis_matrix.shape == (0, 10)
    zero = bmm.matrix_from_rows(["000", "000"])
    r0 = bmm.solver.ExactBasisSolver().solve(zero.row_window([0, 1]))
    assert r0.rank == 0 and r0.enumerated_states == 3 and r0.transform_matrix.shape == (0, 2)
    big = bmm.matrix_from_rows(["1"] * 5)
    with pytest.raises(ValueError, match="max_rows"):
        bmm.solver.ExactBasisSolver(max_rows=4).solve(big.row_window(list(range(5))))
    with pytest.raises(ValueError, match="max_states"):
        bmm.solver.ExactBasisSolver(max_states=30).solve(big.row_window(list(range(5))))
    with pytest.raises(ValueError, match="storage"):
        bmm.solver.ExactBasisSolver(max_storage_bytes=100).solve(big.row_window(list(range(5))))


# ---------------------------------------------------------- SASelector ---
def _sa_cases(policy):
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "golden_sa.json")
    with open(path) as f:
        return [c for c in json.load(f)["cases"] if c["policy"] == policy]


def _bitmatrix(rows_u64, cols):
    strs = ["".join("1" if (int(r[c // 64]) >> (c % 64)) & 1 else "0" for c in range(cols))
            for r in rows_u64]
    return bmm.matrix_from_rows(strs)


def test_sa_selector_higher_order_matches_reference_golden():
    """The higher-order synergy policy (column counts, host annealing) against
    the reference SASelector's outputs for the same seeds (golden_sa.json)."""
    for case in _sa_cases("higher_order_synergy"):
        rows = [[int(v) for v in r] for r in case["rows"]]
        mat = _bitmatrix(rows, case["cols"])
        sel = bmm.SASelector(iterations=case["iterations"], restarts=case["restarts"],
                             seed=case["seed"], score_policy="higher_order_synergy")
        res = sel.select(mat, case["window_size"])
        e = case["expect"]
        assert list(res.rows) == e["rows"] and res.score == e["score"]
        assert (res.accepted_moves, res.iterations_run, res.best_iteration, res.restart_index,
                res.seed) == (e["accepted_moves"], e["iterations_run"], e["best_iteration"],
                              e["restart_index"], e["seed"])


def test_sa_selector_validation():
    mat = bmm.matrix_from_rows(["1100", "0110", "0011"])
    for kw, msg in ((dict(restarts=0), "restarts"),
                    (dict(initial_acceptance_probability=1.0), "initial_acceptance_probability"),
                    (dict(cooling_rate=0.0), "cooling_rate"),
                    (dict(min_temperature=0.0), "min_temperature")):
        with pytest.raises(ValueError, match=msg):
            bmm.SASelector(**kw).select(mat, 2)
    with pytest.raises(ValueError, match="window_size"):
        bmm.SASelector().select(mat, 4)
    with pytest.raises(ValueError, match="score_policy"):
        bmm.SASelector(score_policy="nope")
    r = bmm.SASelector(seed=0).select(mat, 0)
    assert list(r.rows) == [] and r.seed == 0x9E3779B97F4A7C15
    assert bmm.SASelector().name() == "sa"


def test_reference_import_name():
    """``import bmmpy`` and its submodules resolve to this package, with the
    reference's public names (bmmpy/__init__.py:65-103, less the algebra
    package, which this build leaves out)."""
    import importlib

    import bmmpy_amd

    ref_all = {"ApplyResult", "BitMatrix", "BruteforceSearch", "Candidate", "FwhtSearch",
               "GreedyApplier", "GlobalGreedyApplier", "MatrixErr", "MatrixError",
               "MitmFwhtSearch", "CudaMitmFwhtSearch", "CudaBruteforceSearch", "RowWindow",
               "RuntimeFeatures", "SASelectionResult", "SASelector", "GeneticAlgorithm",
               "IslandModel", "IslandModelSnapshot", "IslandSnapshot", "IslandSpec",
               "MigrationPolicy", "OverflowPolicy", "RunStats", "StopCriteria", "ga", "add",
               "solver", "fixed_weight_masks_u32", "fixed_weight_masks_u64", "fwht_i16",
               "fwht_i32", "get_runtime_features", "get_version", "matrix_from_rows",
               "search_apply"}
    assert ref_all <= set(bmm.__all__)
    for name in ref_all:
        assert getattr(bmm, name) is getattr(bmmpy_amd, name)
    assert importlib.import_module("bmmpy._bmmpy") is bmmpy_amd._bmmpy
    for mod in ("search", "solver", "ga", "apply", "matrix", "workflows"):
        assert importlib.import_module(f"bmmpy.{mod}") is getattr(bmmpy_amd, mod)
    from bmmpy._bmmpy import BitMatrix
    from bmmpy.candidate import Candidate
    from bmmpy.search import FwhtSearch
    assert BitMatrix is bmm.BitMatrix and Candidate is bmm.Candidate
    assert FwhtSearch is bmm.FwhtSearch
    assert bmm.__version__ == bmmpy_amd.__version__

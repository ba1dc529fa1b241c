"""Test infrastructure only: ctypes front-end for the CPU checkers.

* ``liboracle.so`` - this project's C restatement of the reference's CPU
  algorithms (``oracle.c``), the checker for the HIP path.
* ``_ref/libbmmref.so`` - the reference's own CPU sources compiled where they
  lie (``Makefile`` target ``ref``; needs the reference tree, so it exists only
  where it was built).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbmmref.so")
# The same reference sources without the optional AVX2 backend (its
# ExactBasisSolver passes 16-byte-aligned std::vector buffers to 32-byte
# aligned AVX2 loads and crashes on rows of >= 4 words).
REF_SCALAR_SO = os.path.join(HERE, "_ref", "libbmmref_scalar.so")
REF_TREE = "/root/reference"

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_szp = np.ctypeslib.ndpointer(np.uintp, flags="C_CONTIGUOUS")


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (always) and _ref (when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if ref is None:
        ref = os.path.isdir(REF_TREE)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "-j8", "ref"], check=True)


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise OracleError(f"{what} failed with code {rc}")
    return rc


class Oracle:
    """Typed wrappers around liboracle.so."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        L = self.lib = C.CDLL(path)
        L.or_xs_seed.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
        L.or_xs_next.argtypes = [C.POINTER(C.c_uint64)]
        L.or_xs_next.restype = C.c_uint64
        L.or_fill_bernoulli.argtypes = [C.POINTER(C.c_uint64), _u64p, C.c_int, C.c_int, C.c_int,
                                        C.c_double]
        L.or_rank.argtypes = [_u64p, C.c_int, C.c_int, C.c_int]
        L.or_mul.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u64p, C.c_int, C.c_int, _u64p]
        L.or_gen_planted.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_double, _u64p, _u64p,
                                     _u64p]
        L.or_gen_batch.argtypes = [C.c_uint64, C.c_int, _i32p, _i32p, _f64p, _u64p, C.c_size_t,
                                   _szp]
        L.or_gen_batch.restype = C.c_longlong
        L.or_fwht_i32.argtypes = [_i32p, C.c_size_t]
        L.or_fwht_i16.argtypes = [_i16p, C.c_size_t]
        L.or_histogram.argtypes = [_u64p, C.c_int, C.c_int, _i32p, C.c_int, _i32p,
                                   C.POINTER(C.c_int64)]
        for fn in ("or_fwht_topk", "or_mitm_topk"):
            f = getattr(L, fn)
            f.argtypes = [_u64p, C.c_int, C.c_int, _i32p, C.c_int, C.c_size_t, _u64p, _u32p]
            f.restype = C.c_longlong
        L.or_bruteforce_topk.argtypes = [_u64p, C.c_int, C.c_int, _i32p, C.c_int, C.c_size_t,
                                         C.c_int, _u64p, _u32p]
        L.or_bruteforce_topk.restype = C.c_longlong
        L.or_split_window.argtypes = [_u64p, C.c_int, C.c_int, _i32p, C.c_int, C.c_int,
                                      C.POINTER(C.c_int32), _u64p, _u64p, _i32p, C.c_size_t]
        L.or_split_window.restype = C.c_longlong
        L.or_coset_min.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u32p, _u64p]
        L.or_exact_basis.argtypes = [_u64p, C.c_int, C.c_int, _i32p, C.c_int, _u32p, _u32p,
                                     C.c_size_t, C.POINTER(C.c_uint64)]
        L.or_weight_order.argtypes = [_u64p, C.c_int, C.c_int, _i32p, C.c_int, C.c_size_t,
                                      _u32p, _u32p]
        L.or_weight_order.restype = C.c_longlong
        L.or_gray_min.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]
        L.or_replay.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u16p, C.c_size_t]
        L.or_replay.restype = C.c_longlong
        L.or_greedy_apply.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u64p, _u32p, C.c_size_t,
                                      C.c_uint64, C.c_int, C.c_uint64, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.or_scores_order_i32.argtypes = [_i32p, C.c_size_t, C.c_int32, _i32p, _i32p, _i32p,
                                          _i32p]
        L.or_scores_order_i16.argtypes = [_i16p, C.c_size_t, C.c_int32, _i16p, _i32p, _i32p,
                                          _i32p]
        L.or_fixed_weight_masks.argtypes = [C.c_uint32, C.c_uint32, C.c_int, _u64p, C.c_size_t]
        L.or_fixed_weight_masks.restype = C.c_longlong
        L.or_global_greedy.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u64p, C.c_int, _u32p,
                                       C.c_size_t, C.c_int, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]

    # -- RNG / generators -------------------------------------------------
    def xs_stream(self, seed: int, n: int) -> np.ndarray:
        st = C.c_uint64()
        self.lib.or_xs_seed(C.byref(st), seed)
        return np.array([self.lib.or_xs_next(C.byref(st)) for _ in range(n)], dtype=np.uint64)

    def bernoulli(self, seed: int, n_rows: int, cols: int, p: float) -> np.ndarray:
        st = C.c_uint64()
        self.lib.or_xs_seed(C.byref(st), seed)
        w = (cols + 63) // 64
        out = np.zeros((n_rows, w), dtype=np.uint64)
        self.lib.or_fill_bernoulli(C.byref(st), out, n_rows, w, cols, p)
        return out

    def rank(self, rows: np.ndarray, cols: int) -> int:
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        return _check(self.lib.or_rank(rows, rows.shape[0], rows.shape[1], cols), "rank")

    def planted(self, seed: int, n_rows: int, cols: int, p: float):
        w = (cols + 63) // 64
        wn = (n_rows + 63) // 64
        S = np.zeros((n_rows, w), dtype=np.uint64)
        A = np.zeros((n_rows, w), dtype=np.uint64)
        M = np.zeros((n_rows, wn), dtype=np.uint64)
        self.lib.or_gen_planted(seed, n_rows, cols, p, S, A, M)
        return S, A, M

    def batch(self, seed: int, n_mat: int):
        rows = np.zeros(n_mat, dtype=np.int32)
        cols = np.zeros(n_mat, dtype=np.int32)
        dens = np.zeros(n_mat, dtype=np.float64)
        cap = n_mat * 28 * 32
        words = np.zeros(cap, dtype=np.uint64)
        offs = np.zeros(n_mat, dtype=np.uintp)
        n = _check(self.lib.or_gen_batch(seed, n_mat, rows, cols, dens, words, cap, offs),
                   "gen_batch")
        return rows, cols, dens, words[:n].copy(), offs.astype(np.int64)

    # -- FWHT / histogram -------------------------------------------------
    def fwht_i32(self, x) -> np.ndarray:
        a = np.array(x, dtype=np.int32)
        _check(self.lib.or_fwht_i32(a, a.size), "fwht_i32")
        return a

    def fwht_i16(self, x) -> np.ndarray:
        a = np.array(x, dtype=np.int16)
        _check(self.lib.or_fwht_i16(a, a.size), "fwht_i16")
        return a

    def mul(self, a, a_cols: int, b, b_cols: int) -> np.ndarray:
        """GF(2) product a * b (rows of words; a has a_cols columns)."""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        out = np.zeros((a.shape[0], b.shape[1]), np.uint64)
        self.lib.or_mul(a, a.shape[0], a.shape[1], a_cols, b, b.shape[1], b_cols, out)
        return out

    def histogram(self, m: np.ndarray, cols: int, idx=None):
        m = np.ascontiguousarray(m, dtype=np.uint64)
        idx = np.arange(m.shape[0], dtype=np.int32) if idx is None else np.asarray(idx, np.int32)
        b = np.zeros(1 << len(idx), dtype=np.int32)
        tot = C.c_int64()
        _check(self.lib.or_histogram(m, m.shape[1], cols, idx, len(idx), b, C.byref(tot)),
               "histogram")
        return b, tot.value

    # -- searches -----------------------------------------------------------
    def _topk(self, fn, m, cols, k, idx, *extra, cap=None):
        m = np.ascontiguousarray(m, dtype=np.uint64)
        idx = np.arange(m.shape[0], dtype=np.int32) if idx is None else np.asarray(idx, np.int32)
        t = len(idx)
        cap = cap if cap is not None else max(1, min(k, (1 << min(t, 40))))
        masks = np.zeros(cap, dtype=np.uint64)
        ws = np.zeros(cap, dtype=np.uint32)
        n = _check(fn(m, m.shape[1], cols, idx, t, k, *extra, masks, ws), fn.__name__)
        return [(int(ws[i]), int(masks[i])) for i in range(n)]

    def fwht_topk(self, m, cols, k, idx=None):
        return self._topk(self.lib.or_fwht_topk, m, cols, k, idx)

    def mitm_topk(self, m, cols, k, idx=None):
        return self._topk(self.lib.or_mitm_topk, m, cols, k, idx)

    def bruteforce_topk(self, m, cols, k, chunk_bits=0, idx=None):
        return self._topk(self.lib.or_bruteforce_topk, m, cols, k, idx, chunk_bits)

    def split_window(self, m, cols, low_bits, idx=None):
        m = np.ascontiguousarray(m, dtype=np.uint64)
        idx = np.arange(m.shape[0], dtype=np.int32) if idx is None else np.asarray(idx, np.int32)
        cap = cols + 1
        q = np.zeros(cap, np.uint64)
        r = np.zeros(cap, np.uint64)
        mult = np.zeros(cap, np.int32)
        tot = C.c_int32()
        n = _check(self.lib.or_split_window(m, m.shape[1], cols, idx, len(idx), low_bits,
                                            C.byref(tot), q, r, mult, cap), "split_window")
        return tot.value, q[:n].copy(), r[:n].copy(), mult[:n].copy()

    def coset_min(self, m, cols):
        m = np.ascontiguousarray(m, dtype=np.uint64)
        n = m.shape[0]
        w = np.zeros(n, np.uint32)
        a = np.zeros(n, np.uint64)
        _check(self.lib.or_coset_min(m, m.shape[1], cols, n, w, a), "coset_min")
        return w, a

    def exact_basis(self, m, cols):
        """ExactBasisSolver on all rows of m: (masks, weights, total, rank)."""
        m = np.ascontiguousarray(m, dtype=np.uint64)
        t = m.shape[0]
        idx = np.arange(t, dtype=np.int32)
        masks, ws = np.zeros(t, np.uint32), np.zeros(t, np.uint32)
        tot = C.c_uint64()
        rank = _check(self.lib.or_exact_basis(m, m.shape[1], cols, idx, t, masks, ws, t,
                                              C.byref(tot)), "exact_basis")
        return masks[:rank], ws[:rank], tot.value, rank

    def weight_order(self, m, cols, k):
        """First k nonzero masks in (weight, Gray index) order."""
        m = np.ascontiguousarray(m, dtype=np.uint64)
        t = m.shape[0]
        idx = np.arange(t, dtype=np.int32)
        masks, ws = np.zeros(max(k, 1), np.uint32), np.zeros(max(k, 1), np.uint32)
        n = _check(self.lib.or_weight_order(m, m.shape[1], cols, idx, t, k, masks, ws),
                   "weight_order")
        return masks[:n], ws[:n]

    def gray_min(self, m, block_rows, target, words=None):
        m = np.ascontiguousarray(m, dtype=np.uint64)
        words = m.shape[1] if words is None else words
        w = C.c_uint32()
        s = C.c_uint64()
        lm = C.c_uint64()
        _check(self.lib.or_gray_min(m, m.shape[1], words, block_rows, target, C.byref(w),
                                    C.byref(s), C.byref(lm)), "gray_min")
        return w.value, s.value, lm.value

    def replay(self, m, cols, ops):
        m = np.array(m, dtype=np.uint64, copy=True, order="C")
        ops = np.ascontiguousarray(ops, dtype=np.uint16).reshape(-1)
        wt = _check(self.lib.or_replay(m, m.shape[0], m.shape[1], cols, ops, ops.size // 2),
                    "replay")
        return m, wt

    def greedy_apply(self, m, cols, cands, min_gain=0, stochastic=False, seed=0):
        m = np.array(m, dtype=np.uint64, copy=True, order="C")
        masks = np.array([c[1] for c in cands], dtype=np.uint64)
        ws = np.array([c[0] for c in cands], dtype=np.uint32)
        ap, imp = C.c_uint64(), C.c_uint64()
        _check(self.lib.or_greedy_apply(m, m.shape[0], m.shape[1], cols, masks, ws, len(cands),
                                        min_gain, int(stochastic), seed, C.byref(ap),
                                        C.byref(imp)), "greedy_apply")
        return m, ap.value, imp.value

    def scores_order(self, h, n):
        """calc_scores_and_order on int32 or int16 h: (s, order, cnt, off)."""
        return _scores_order_call(self.lib.or_scores_order_i32, self.lib.or_scores_order_i16,
                                  h, n, lambda f, *a: _check(f(*a), "scores_order"))

    def fixed_weight_masks(self, n, k, width=32):
        """fixed_weight_masks_u32 / _u64 as a uint64 array."""
        cnt = self.lib.or_fixed_weight_masks(n, k, width, np.zeros(1, np.uint64), 0)
        _check(min(cnt, 0), "fixed_weight_masks")
        out = np.zeros(max(cnt, 1), np.uint64)
        self.lib.or_fixed_weight_masks(n, k, width, out, out.size)
        return out[:cnt]

    def global_greedy(self, m, cols, masks, weights, require_improvement=True):
        """GlobalGreedyApplier on all rows of m; masks: (n_cand, mask_words)
        uint64.  Returns (rows, applied, improvement)."""
        m = np.array(m, dtype=np.uint64, copy=True, order="C")
        masks, ws = _cand_arrays(masks, weights)
        ap, imp = C.c_uint64(), C.c_uint64()
        _check(self.lib.or_global_greedy(m, m.shape[0], m.shape[1], cols, masks.reshape(-1),
                                         masks.shape[1], ws, ws.size, int(require_improvement),
                                         C.byref(ap), C.byref(imp)), "global_greedy")
        return m, ap.value, imp.value


class RefError(RuntimeError):
    pass


class Reference:
    """The reference's CPU library (compiled from its sources) via ref_driver."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        ep = [C.c_char_p, C.c_int]
        L.ref_fwht_search.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_size_t,
                                      _u64p, _u32p, C.c_size_t] + ep
        L.ref_mitm_search.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_size_t, _u64p, _u32p,
                                      C.c_size_t] + ep
        L.ref_bruteforce_search.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_size_t,
                                            C.c_size_t, _u64p, _u32p, C.c_size_t] + ep
        L.ref_fwht_i32.argtypes = [_i32p, C.c_size_t] + ep
        L.ref_fwht_i16.argtypes = [_i16p, C.c_size_t] + ep
        L.ref_compact_split.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int), _u64p, _u64p,
                                        _i32p, C.c_size_t] + ep
        L.ref_replay.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u16p, C.c_size_t] + ep
        L.ref_replay.restype = C.c_longlong
        L.ref_greedy_apply.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u64p, _u32p,
                                       C.c_size_t, C.c_uint64, C.c_int, C.c_uint64,
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)] + ep
        L.ref_exact_basis.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u32p, _u32p,
                                      C.c_size_t, C.POINTER(C.c_uint64), _u64p] + ep
        L.ref_sa_select.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_size_t,
                                    C.c_size_t, C.c_uint64, C.c_int, C.c_size_t, C.c_double,
                                    C.c_double, C.c_double, _u64p, _u64p,
                                    C.POINTER(C.c_longlong)] + ep
        L.ref_scores_order_i32.argtypes = [_i32p, C.c_size_t, C.c_int32, _i32p, _i32p, _i32p,
                                           _i32p] + ep
        L.ref_scores_order_i16.argtypes = [_i16p, C.c_size_t, C.c_int32, _i16p, _i32p, _i32p,
                                           _i32p] + ep
        L.ref_global_greedy.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u64p, C.c_int, _u32p,
                                        C.c_size_t, C.c_int, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)] + ep
        L.ref_matrix_io.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p,
                                    C.c_size_t, C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t,
                                    C.POINTER(C.c_size_t)] + ep
        L.ref_matrix_load.argtypes = [C.c_char_p, C.c_size_t, C.c_int, _u64p, C.c_size_t,
                                      _u64p] + ep
        L.ref_mul_rank.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, _u64p, C.c_int, C.c_int,
                                   _u64p] + ep
        L.ref_fixed_weight_masks.argtypes = [C.c_uint32, C.c_uint32, C.c_int, _u64p,
                                             C.c_size_t] + ep
        L.ref_fixed_weight_masks.restype = C.c_longlong

    @staticmethod
    def _err():
        return C.create_string_buffer(512)

    def _run(self, fn, *args):
        e = self._err()
        rc = fn(*args, e, 512)
        if rc < 0:
            raise RefError(e.value.decode())
        return rc

    @staticmethod
    def _rows(m):
        return np.ascontiguousarray(m, dtype=np.uint64)

    def sa_select(self, m, cols, window_size, iterations=10000, restarts=8, seed=0,
                  policy="pairwise_synergy", probe_samples=64, p0=0.8, cooling=0.99,
                  min_t=1e-6):
        """Reference SASelector::select: dict of the result fields."""
        m = self._rows(m)
        t, W = m.shape
        rows = np.zeros(max(window_size, 1), np.uint64)
        stats = np.zeros(5, np.uint64)
        score = C.c_longlong()
        n = self._run(self.lib.ref_sa_select, m, t, W, cols, window_size, iterations, restarts,
                      seed, 1 if policy == "higher_order_synergy" else 0, probe_samples, p0,
                      cooling, min_t, rows, stats, C.byref(score))
        return {"rows": [int(x) for x in rows[:n]], "score": score.value,
                "accepted_moves": int(stats[0]), "iterations_run": int(stats[1]),
                "best_iteration": int(stats[2]), "restart_index": int(stats[3]),
                "seed": int(stats[4])}

    def exact_basis(self, m, cols):
        """Reference ExactBasisSolver: (masks, weights, total, rank, basis rows)."""
        m = self._rows(m)
        t, W = m.shape
        masks, ws = np.zeros(max(t, 1), np.uint32), np.zeros(max(t, 1), np.uint32)
        basis = np.zeros((max(t, 1), W), np.uint64)
        tot = C.c_uint64()
        rank = self._run(self.lib.ref_exact_basis, m, t, W, cols, masks, ws, t, C.byref(tot),
                         basis)
        return masks[:rank], ws[:rank], tot.value, rank, basis[:rank]

    def fwht_search(self, m, cols, k, max_rows=64, cap=None):
        m = self._rows(m)
        t = m.shape[0]
        cap = cap or max(1, min(k, 1 << t))
        masks, ws = np.zeros(cap, np.uint64), np.zeros(cap, np.uint32)
        n = self._run(self.lib.ref_fwht_search, m, t, m.shape[1], cols, max_rows, k, masks, ws,
                      cap)
        n = min(n, cap)
        return [(int(ws[i]), int(masks[i])) for i in range(n)]

    def fwht_search_arrays(self, m, cols, k, max_rows=64):
        m = self._rows(m)
        t = m.shape[0]
        cap = max(1, min(k, 1 << t))
        masks, ws = np.zeros(cap, np.uint64), np.zeros(cap, np.uint32)
        n = self._run(self.lib.ref_fwht_search, m, t, m.shape[1], cols, max_rows, k, masks, ws,
                      cap)
        return ws[:n], masks[:n]

    def mitm_search(self, m, cols, k):
        m = self._rows(m)
        t = m.shape[0]
        cap = max(1, k)
        masks, ws = np.zeros(cap, np.uint64), np.zeros(cap, np.uint32)
        n = self._run(self.lib.ref_mitm_search, m, t, m.shape[1], cols, k, masks, ws, cap)
        return [(int(ws[i]), int(masks[i])) for i in range(min(n, cap))]

    def bruteforce_search(self, m, cols, k, chunk_bits=0):
        m = self._rows(m)
        t = m.shape[0]
        cap = max(1, k)
        masks, ws = np.zeros(cap, np.uint64), np.zeros(cap, np.uint32)
        n = self._run(self.lib.ref_bruteforce_search, m, t, m.shape[1], cols, k, chunk_bits,
                      masks, ws, cap)
        return [(int(ws[i]), int(masks[i])) for i in range(min(n, cap))]

    def fwht_i32(self, x):
        a = np.array(x, dtype=np.int32)
        self._run(self.lib.ref_fwht_i32, a, a.size)
        return a

    def fwht_i16(self, x):
        a = np.array(x, dtype=np.int16)
        self._run(self.lib.ref_fwht_i16, a, a.size)
        return a

    def compact_split(self, m, cols, low_bits):
        m = self._rows(m)
        cap = cols + 1
        q, r = np.zeros(cap, np.uint64), np.zeros(cap, np.uint64)
        mult = np.zeros(cap, np.int32)
        tot, hb = C.c_int32(), C.c_int()
        n = self._run(self.lib.ref_compact_split, m, m.shape[0], m.shape[1], cols, low_bits,
                      C.byref(tot), C.byref(hb), q, r, mult, cap)
        return tot.value, q[:n].copy(), r[:n].copy(), mult[:n].copy()

    def replay(self, m, cols, ops):
        m = np.array(m, dtype=np.uint64, copy=True, order="C")
        ops = np.ascontiguousarray(ops, dtype=np.uint16).reshape(-1)
        e = self._err()
        wt = self.lib.ref_replay(m, m.shape[0], m.shape[1], cols, ops, ops.size // 2, e, 512)
        if wt < 0:
            raise RefError(e.value.decode())
        return m, wt

    def greedy_apply(self, m, cols, cands, min_gain=0, stochastic=False, seed=0):
        m = np.array(m, dtype=np.uint64, copy=True, order="C")
        masks = np.array([c[1] for c in cands], dtype=np.uint64)
        ws = np.array([c[0] for c in cands], dtype=np.uint32)
        ap, imp = C.c_uint64(), C.c_uint64()
        self._run(self.lib.ref_greedy_apply, m, m.shape[0], m.shape[1], cols, masks, ws,
                  len(cands), min_gain, int(stochastic), seed, C.byref(ap), C.byref(imp))
        return m, ap.value, imp.value

    def scores_order(self, h, n):
        return _scores_order_call(self.lib.ref_scores_order_i32, self.lib.ref_scores_order_i16,
                                  h, n, self._run)

    def fixed_weight_masks(self, n, k, width=32):
        """Reference fixed_weight_masks_u32 / _u64 as a uint64 array."""
        cnt = self._run(self.lib.ref_fixed_weight_masks, n, k, width, np.zeros(1, np.uint64), 0)
        out = np.zeros(max(cnt, 1), np.uint64)
        self._run(self.lib.ref_fixed_weight_masks, n, k, width, out, out.size)
        return out[:cnt]

    def mul_rank(self, a, a_cols, b, b_cols):
        """Reference (a * b, rank(a))."""
        a, b = self._rows(a), self._rows(b)
        out = np.zeros((a.shape[0], b.shape[1]), np.uint64)
        rank = self._run(self.lib.ref_mul_rank, a, a.shape[0], a.shape[1], a_cols, b, b.shape[1],
                         b_cols, out)
        return out, rank

    def matrix_io(self, m, cols):
        """Reference (hash hex, save_binary bytes, save_text str) of rows m."""
        m = self._rows(m)
        hexbuf = C.create_string_buffer(65)
        bl, tl = C.c_size_t(), C.c_size_t()
        self._run(self.lib.ref_matrix_io, m, m.shape[0], m.shape[1], cols, hexbuf, None, 0,
                  C.byref(bl), None, 0, C.byref(tl))
        b = C.create_string_buffer(max(bl.value, 1))
        t = C.create_string_buffer(max(tl.value, 1))
        self._run(self.lib.ref_matrix_io, m, m.shape[0], m.shape[1], cols, hexbuf, b, bl.value,
                  C.byref(bl), t, tl.value, C.byref(tl))
        return hexbuf.value.decode(), b.raw[:bl.value], t.raw[:tl.value].decode()

    def matrix_load(self, data: bytes, text: bool = False, cap: int = 1 << 20):
        """Reference load_binary / load_text: (rows as words, rows, cols)."""
        out = np.zeros(max(cap, 1), np.uint64)
        shape = np.zeros(2, np.uint64)
        self._run(self.lib.ref_matrix_load, data, len(data), int(text), out, cap, shape)
        r, c = int(shape[0]), int(shape[1])
        w = (c + 63) // 64
        return out[:r * w].reshape(r, w), r, c

    def global_greedy(self, m, cols, masks, weights, require_improvement=True):
        m = np.array(m, dtype=np.uint64, copy=True, order="C")
        masks, ws = _cand_arrays(masks, weights)
        ap, imp = C.c_uint64(), C.c_uint64()
        self._run(self.lib.ref_global_greedy, m, m.shape[0], m.shape[1], cols, masks.reshape(-1),
                  masks.shape[1], ws, ws.size, int(require_improvement), C.byref(ap),
                  C.byref(imp))
        return m, ap.value, imp.value


def _scores_order_call(f32, f16, h, n, run):
    h = np.ascontiguousarray(h)
    f = {np.dtype(np.int32): f32, np.dtype(np.int16): f16}[h.dtype]
    m = h.size
    s = np.zeros(max(m, 1), h.dtype)
    order = np.zeros(max(m - 1, 1), np.int32)
    cnt = np.zeros(max(n + 1, 1), np.int32)
    off = np.zeros(max(n + 1, 1), np.int32)
    run(f, h if m else np.zeros(1, h.dtype), m, n, s, order, cnt, off)
    return s[:m], order[:max(m - 1, 0)], cnt[:n + 1], off[:n + 1]


def _cand_arrays(masks, weights):
    masks = np.ascontiguousarray(masks, dtype=np.uint64)
    if masks.ndim == 1:
        masks = masks.reshape(-1, 1)
    if masks.size == 0:
        masks = np.zeros((0, 1), np.uint64)
    ws = np.ascontiguousarray(weights, dtype=np.uint32)
    return np.ascontiguousarray(masks), ws


def oracle() -> Oracle:
    return Oracle()


def reference(scalar: bool = False) -> Reference | None:
    """The compiled reference (scalar: without its AVX2 backend), or None
    where it was not built."""
    path = REF_SCALAR_SO if scalar else REF_SO
    return Reference(path) if os.path.exists(path) else None


# -- small pure-numpy helpers shared by tests ------------------------------
def matrix_from_rows(rows):
    """Rows given as '0'/'1' strings, char c = column c (tests/support/test_common.hpp:31-53)."""
    cols = len(rows[0]) if rows else 0
    w = (cols + 63) // 64
    m = np.zeros((len(rows), max(w, 1)), dtype=np.uint64)
    for r, s in enumerate(rows):
        assert len(s) == cols
        for c, ch in enumerate(s):
            if ch == "1":
                m[r, c // 64] |= np.uint64(1) << np.uint64(c % 64)
    return m, cols


def equivalence_matrix(rows=28, cols=96):
    """Deterministic hash fixture (tests/support/test_common.hpp:173-190)."""
    M = (1 << 64) - 1
    w = (cols + 63) // 64
    m = np.zeros((rows, w), dtype=np.uint64)
    for r in range(rows):
        for c in range(cols):
            mix = (((r + 1) * 0x9E3779B185EBCA87) & M) ^ (((c + 3) * 0xC2B2AE3D27D4EB4F) & M) ^ \
                  (((r + c + 11) * 0x165667B19E3779F9) & M)
            folded = mix ^ (mix >> 17) ^ (mix >> 33)
            if folded & 1:
                m[r, c // 64] |= np.uint64(1) << np.uint64(c % 64)
    return m

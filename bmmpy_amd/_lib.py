"""ctypes binding of ``libbmmpy_hip.so`` (the C ABI in ``include/bmmpy_hip.h``).

The library is required: importing a compute wrapper without it raises, there
is no CPU fallback.  Status codes map onto the reference's exception types
(invalid_argument -> ValueError, overflow_error -> OverflowError,
out_of_range -> IndexError, runtime_error -> RuntimeError).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbmmpy_hip.so")

BMM_OK, BMM_ERR_INVALID, BMM_ERR_OVERFLOW, BMM_ERR_RANGE, BMM_ERR_DEVICE = range(5)
TOPK_MAX = 4096

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_sz = C.c_size_t


class TopKEntry(C.Structure):
    _fields_ = [("mask", C.c_uint64), ("weight", C.c_uint32), ("pad", C.c_uint32)]


TOPK_DTYPE = np.dtype([("mask", np.uint64), ("weight", np.uint32), ("pad", np.uint32)])
_topkp = np.ctypeslib.ndpointer(TOPK_DTYPE, flags="C_CONTIGUOUS")


class DeviceInfo(C.Structure):
    _fields_ = [("device_count", C.c_int), ("device", C.c_int), ("compute_units", C.c_int),
                ("wave_size", C.c_int), ("lds_per_block", C.c_int), ("clock_khz", C.c_int),
                ("arch", C.c_char * 64), ("name", C.c_char * 128)]


class BatchInfo(C.Structure):
    _fields_ = [("n_matrices", C.c_size_t), ("n_slots", C.c_size_t),
                ("total_blocks", C.c_size_t), ("persistent_wgs", C.c_int), ("chunk", C.c_int),
                ("max_block_bits", C.c_int), ("subset_evals", C.c_double),
                ("transform_elements", C.c_double)]


class GaParams(C.Structure):
    _fields_ = [("islands", C.c_int), ("population", C.c_int), ("max_ops", C.c_int),
                ("tournament", C.c_int), ("elite", C.c_int), ("migrate_every", C.c_int),
                ("mutations", C.c_int), ("learned", C.c_int), ("n_candidates", C.c_int),
                ("train_steps", C.c_int), ("check_every", C.c_int), ("use_graph", C.c_int),
                ("li_sweeps", C.c_int), ("rows", C.c_int), ("p_append", C.c_float),
                ("p_replace", C.c_float), ("p_delete", C.c_float), ("learning_rate", C.c_float),
                ("p_learned", C.c_float), ("seed", C.c_uint64), ("win_rows", C.c_int),
                ("win_rounds", C.c_int), ("win_candidates", C.c_int), ("isd_trials", C.c_int),
                ("num_parents", C.c_int), ("num_offspring", C.c_int),
                ("mutation_rate", C.c_float), ("catastrophe_threshold", C.c_int),
                ("catastrophe_survival", C.c_float), ("restart_after", C.c_int),
                ("restart_p", C.c_float), ("init_rounds", C.c_int)]


class GaIslandState(C.Structure):
    _fields_ = [("best_weight", C.c_uint32), ("best_len", C.c_int32),
                ("no_improvement", C.c_uint32), ("catastrophe", C.c_uint32),
                ("mutation_rate", C.c_float), ("migrate", C.c_int32)]


class GaStatus(C.Structure):
    _fields_ = [("generations", C.c_size_t), ("ran", C.c_size_t), ("evaluations", C.c_uint64),
                ("best_weight", C.c_uint64), ("best_island", C.c_int),
                ("reached_target", C.c_int), ("elapsed_s", C.c_double),
                ("train_loss", C.c_double), ("li_moves", C.c_uint64),
                ("win_moves", C.c_uint64), ("isd_rows", C.c_uint64),
                ("replayed_ops", C.c_uint64)]


# name -> (argtypes, restype)
SIGNATURES = {
    "bmm_hip_last_error": ([], C.c_char_p),
    "bmm_hip_version": ([], C.c_char_p),
    "bmm_hip_device_info": ([C.POINTER(DeviceInfo)], C.c_int),
    "bmm_hip_timing_begin": ([], C.c_int),
    "bmm_hip_timing_end": ([C.POINTER(C.c_double), C.POINTER(_sz)], C.c_int),
    "bmm_hip_fwht_i32": ([_i32p, _sz], C.c_int),
    "bmm_hip_fwht_i16": ([_i16p, _sz], C.c_int),
    "bmm_hip_fwht_i32_dev": ([_vp, _sz, _vp], C.c_int),
    "bmm_hip_fwht_i16_dev": ([_vp, _sz, _vp], C.c_int),
    "bmm_hip_fwht_topk": ([_u64p, _sz, _sz, _sz, _sz, _topkp, C.POINTER(_sz)], C.c_int),
    "bmm_hip_mitm_topk": ([_u64p, _u64p, _i32p, _sz, _sz, _sz, C.c_int32, _sz, _topkp,
                           C.POINTER(_sz)], C.c_int),
    "bmm_hip_bruteforce_topk": ([_u64p, _sz, _sz, _sz, _sz, _topkp, C.POINTER(_sz)], C.c_int),
    "bmm_hip_coset_min": ([_u64p, _sz, _sz, _sz, _u32p, _u64p], C.c_int),
    "bmm_hip_coset_min_batch": ([_u64p, _i64p, _i32p, _i32p, _sz, _i64p, _u32p, _u64p],
                                C.c_int),
    "bmm_hip_coset_min_batch_dev": ([_vp, _i64p, _i32p, _i32p, _sz, _i64p, _vp, _vp, _sz, _vp],
                                    C.c_int),
    "bmm_hip_batch_plan_create": ([_vp, _i64p, _i32p, _i32p, _sz, _i64p, _vp, _vp, _sz, C.c_int,
                                   C.POINTER(C.c_void_p)], C.c_int),
    "bmm_hip_batch_plan_run": ([_vp, C.c_int, _vp], C.c_int),
    "bmm_hip_batch_plan_info": ([_vp, C.POINTER(BatchInfo)], C.c_int),
    "bmm_hip_batch_plan_set_shape": ([_vp, C.c_int, _sz, _sz], C.c_int),
    "bmm_hip_batch_plan_destroy": ([_vp], C.c_int),
    "bmm_hip_window_coset_min": ([_u64p, _sz, _sz, _sz, _i32p, _sz, _i32p, _sz, _u32p, _u64p],
                                 C.c_int),
    "bmm_hip_gray_coset_min": ([_u64p, _sz, _sz, _sz, _i32p, _sz, _u32p, _u64p, _u64p],
                               C.c_int),
    "bmm_hip_gray_coset_min_dev": ([_vp, _sz, _sz, _sz, _i32p, _sz, _sz, _vp, _vp, _vp],
                                   C.c_int),
    "bmm_hip_gray_coset_min_noalloc": ([_vp, _sz, _sz, _sz, _i32p, _sz, _sz, _vp, _vp, _vp],
                                       C.c_int),
    "bmm_hip_replay": ([_u64p, _sz, _sz, _u16p, _sz, C.POINTER(C.c_uint64)], C.c_int),
    "bmm_hip_ga_default_params": ([C.POINTER(GaParams)], C.c_int),
    "bmm_hip_ga_create": ([_u64p, _sz, _sz, C.POINTER(GaParams), C.POINTER(C.c_void_p)], C.c_int),
    "bmm_hip_ga_run": ([_vp, _sz, C.c_uint64, C.c_double, C.POINTER(GaStatus)], C.c_int),
    "bmm_hip_ga_best": ([_vp, _u64p, C.POINTER(C.c_uint64)], C.c_int),
    "bmm_hip_ga_island_best": ([_vp, C.c_int, _u64p, C.POINTER(C.c_uint64)], C.c_int),
    "bmm_hip_ga_destroy": ([_vp], C.c_int),
    "bmm_hip_ga_configure_islands": ([_vp, _vp, _vp, _sz], C.c_int),
    "bmm_hip_ga_best_ops": ([_vp, C.c_int, _u64p, _u16p, C.POINTER(_sz)], C.c_int),
    "bmm_hip_ga_set_island_base": ([_vp, C.c_int, _u64p], C.c_int),
    "bmm_hip_ga_proposals": ([_vp, C.c_int, _u16p, _sz, C.POINTER(_sz)], C.c_int),
    "bmm_hip_ga_population": ([_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                               C.POINTER(GaIslandState)], C.c_int),
    "bmm_hip_mlp_num_params": ([], C.c_size_t),
    "bmm_hip_mlp_create": ([C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "bmm_hip_mlp_destroy": ([_vp], C.c_int),
    "bmm_hip_mlp_set_params": ([_vp, _f32p], C.c_int),
    "bmm_hip_mlp_get_params": ([_vp, _f32p], C.c_int),
    "bmm_hip_mlp_forward": ([_vp, _f32p, _sz, _f32p], C.c_int),
    "bmm_hip_mlp_score_pairs": ([_vp, _u64p, _sz, _u16p, _sz, _f32p], C.c_int),
    "bmm_hip_pair_overlaps": ([_u64p, _sz, _sz, _u32p], C.c_int),
    "bmm_hip_weight_order": ([_u64p, _sz, _sz, _sz, _sz, _u32p, _u32p, C.POINTER(_sz)], C.c_int),
    "bmm_hip_row_combine": ([_u64p, _sz, _sz, _u64p, _sz, _sz, _u64p, C.POINTER(C.c_uint64)],
                            C.c_int),
    "bmm_hip_scores_order_i32": ([_i32p, _sz, C.c_int32, _i32p, _i32p, _i32p, _i32p], C.c_int),
    "bmm_hip_scores_order_i16": ([_i16p, _sz, C.c_int32, _i16p, _i32p, _i32p, _i32p], C.c_int),
    "bmm_hip_fixed_weight_masks_u32": ([C.c_uint32, C.c_uint32, _vp, _sz, C.POINTER(_sz)],
                                       C.c_int),
    "bmm_hip_fixed_weight_masks_u64": ([C.c_uint32, C.c_uint32, _vp, _sz, C.POINTER(_sz)],
                                       C.c_int),
    "bmm_hip_mlp_gradients": ([_vp, _f32p, _f32p, _sz, _f32p, C.POINTER(C.c_float)], C.c_int),
    "bmm_hip_mlp_train_step": ([_vp, _f32p, _f32p, _sz, C.c_float, C.POINTER(C.c_float)],
                               C.c_int),
    "bmm_gen_bernoulli": ([C.c_uint64, _sz, _sz, C.c_double, _u64p], C.c_int),
    "bmm_gen_planted": ([C.c_uint64, _sz, _sz, C.c_double, _u64p, _u64p, _u64p], C.c_int),
    "bmm_gen_batch": ([C.c_uint64, _sz, _i32p, _i32p, _f64p, _vp, _sz, _i64p, C.POINTER(_sz)],
                      C.c_int),
}

_lib = None


def load() -> C.CDLL:
    """Load libbmmpy_hip.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        from . import _runtime  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        lib = C.CDLL(LIB_PATH)
        for name, (argt, rest) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.argtypes = argt
            fn.restype = rest
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    lib = load()
    return [n for n in SIGNATURES if hasattr(lib, n)]


_EXC = {BMM_ERR_INVALID: ValueError, BMM_ERR_OVERFLOW: OverflowError, BMM_ERR_RANGE: IndexError,
        BMM_ERR_DEVICE: RuntimeError}


def check(rc: int) -> None:
    if rc != BMM_OK:
        msg = load().bmm_hip_last_error().decode()
        raise _EXC.get(rc, RuntimeError)(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def device_info() -> DeviceInfo:
    d = DeviceInfo()
    call("bmm_hip_device_info", C.byref(d))
    return d


def timing_begin() -> None:
    call("bmm_hip_timing_begin")


def timing_end() -> tuple[float, int]:
    """(summed kernel milliseconds, launches) of the timed dominant kernels."""
    ms, n = C.c_double(), _sz()
    call("bmm_hip_timing_end", C.byref(ms), C.byref(n))
    return ms.value, n.value


def _rows(m) -> np.ndarray:
    return np.ascontiguousarray(m, dtype=np.uint64)


# ---- thin typed wrappers ---------------------------------------------------
def fwht_i32(x) -> np.ndarray:
    a = np.array(x, dtype=np.int32, copy=True)
    call("bmm_hip_fwht_i32", a, a.size)
    return a


def fwht_i16(x) -> np.ndarray:
    a = np.array(x, dtype=np.int16, copy=True)
    call("bmm_hip_fwht_i16", a, a.size)
    return a


def _cap(k: int, t: int) -> int:
    """Result capacity: min(k, 2^t - 1) records (at least one)."""
    return max(1, k if t >= 63 else min(k, (1 << t) - 1))


def _entries(out, n):
    return [(int(out["weight"][i]), int(out["mask"][i])) for i in range(n)]


def fwht_topk(rows, cols: int, k: int):
    rows = _rows(rows)
    out = np.zeros(_cap(k, rows.shape[0]), dtype=TOPK_DTYPE)
    n = _sz()
    call("bmm_hip_fwht_topk", rows, rows.shape[0], rows.shape[1], cols, k, out, C.byref(n))
    return _entries(out, n.value)


def mitm_topk(q, r, mult, low_bits: int, high_bits: int, total: int, k: int):
    q = np.ascontiguousarray(q, dtype=np.uint64)
    r = np.ascontiguousarray(r, dtype=np.uint64)
    mult = np.ascontiguousarray(mult, dtype=np.int32)
    out = np.zeros(_cap(k, low_bits + high_bits), dtype=TOPK_DTYPE)
    n = _sz()
    call("bmm_hip_mitm_topk", q, r, mult, q.size, low_bits, high_bits, total, k, out,
         C.byref(n))
    return _entries(out, n.value)


def bruteforce_topk(rows, k: int, chunk_bits: int = 0):
    rows = _rows(rows)
    out = np.zeros(_cap(k, rows.shape[0]), dtype=TOPK_DTYPE)
    n = _sz()
    call("bmm_hip_bruteforce_topk", rows, rows.shape[0], rows.shape[1], chunk_bits, k, out,
         C.byref(n))
    return _entries(out, n.value)


def coset_min(rows, cols: int):
    rows = _rows(rows)
    w = np.zeros(rows.shape[0], np.uint32)
    i = np.zeros(rows.shape[0], np.uint64)
    call("bmm_hip_coset_min", rows, rows.shape[0], rows.shape[1], cols, w, i)
    return w, i


def coset_min_batch(words, word_offset, rows, cols, out_offset=None):
    words = np.ascontiguousarray(words, dtype=np.uint64)
    word_offset = np.ascontiguousarray(word_offset, dtype=np.int64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    if out_offset is None:
        out_offset = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    out_offset = np.ascontiguousarray(out_offset, dtype=np.int64)
    n_slots = int((out_offset + rows).max()) if rows.size else 0
    w = np.zeros(max(n_slots, 1), np.uint32)
    i = np.zeros(max(n_slots, 1), np.uint64)
    call("bmm_hip_coset_min_batch", words, word_offset, rows, cols, rows.size, out_offset, w, i)
    return w[:n_slots], i[:n_slots]


def window_coset_min(rows, cols: int, targets, windows):
    """Per target t (targets[i]) and window rows windows[i]: min weight of
    row_t ^ XOR(subset of window) and its lowest subset mask."""
    rows = _rows(rows)
    targets = np.ascontiguousarray(targets, dtype=np.int32)
    windows = np.ascontiguousarray(windows, dtype=np.int32)
    n, width = windows.shape
    w = np.zeros(n, np.uint32)
    m = np.zeros(n, np.uint64)
    call("bmm_hip_window_coset_min", rows, rows.shape[0], rows.shape[1], cols,
         windows.reshape(-1), width, targets, n, w, m)
    return w, m


def pair_overlaps(rows):
    """popcount(row_i & row_j) for every pair (0 on the diagonal)."""
    rows = _rows(rows)
    n = rows.shape[0]
    out = np.zeros((n, n), np.uint32)
    call("bmm_hip_pair_overlaps", rows, n, rows.shape[1], out.reshape(-1))
    return out


def weight_order(rows, cols: int, k: int):
    """First k nonzero masks of the window `rows` in (weight, Gray index)
    order: (masks, weights)."""
    rows = _rows(rows)
    m = np.zeros(max(k, 1), np.uint32)
    w = np.zeros(max(k, 1), np.uint32)
    n = _sz()
    call("bmm_hip_weight_order", rows, rows.shape[0], rows.shape[1], cols, k, m, w, C.byref(n))
    return m[:n.value], w[:n.value]


def row_combine(rows, masks):
    """GF(2) product: out row i = XOR of the rows selected by masks[i]
    (uint64 words).  Returns (out rows, total weight)."""
    rows = _rows(rows)
    masks = np.ascontiguousarray(masks, dtype=np.uint64)
    if masks.ndim == 1:
        masks = masks.reshape(-1, 1)
    out = np.zeros((masks.shape[0], rows.shape[1]), np.uint64)
    wt = C.c_uint64()
    call("bmm_hip_row_combine", rows.reshape(-1), rows.shape[0], rows.shape[1],
         masks.reshape(-1), masks.shape[1], masks.shape[0], out.reshape(-1), C.byref(wt))
    return out, wt.value


def fixed_weight_masks(n: int, k: int, bits: int = 32) -> np.ndarray:
    """fixed_weight_masks_u32 / _u64 as a numpy array (count query, then the
    masks)."""
    name = f"bmm_hip_fixed_weight_masks_u{bits}"
    cnt = _sz()
    call(name, n, k, None, 0, C.byref(cnt))
    out = np.zeros(max(cnt.value, 1), np.uint32 if bits == 32 else np.uint64)
    if cnt.value:
        call(name, n, k, out.ctypes.data, out.size, C.byref(cnt))
    return out[:cnt.value]


def calc_scores_and_order(h, n: int):
    """calc_scores_and_order on the dtype of h (int32 or int16):
    (s_by_mask, order, cnt, off)."""
    h = np.ascontiguousarray(h)
    if h.dtype not in (np.int32, np.int16):
        raise TypeError("h must be int32 or int16")
    m = h.size
    s = np.zeros(max(m, 1), h.dtype)
    order = np.zeros(max(m - 1, 1), np.int32)
    cnt = np.zeros(max(n + 1, 1), np.int32)
    off = np.zeros(max(n + 1, 1), np.int32)
    name = "bmm_hip_scores_order_i32" if h.dtype == np.int32 else "bmm_hip_scores_order_i16"
    call(name, h if m else np.zeros(1, h.dtype), m, n, s, order, cnt, off)
    return s[:m], order[:max(m - 1, 0)], cnt[:n + 1], off[:n + 1]


def gray_coset_min(rows, block_rows: int, targets):
    rows = _rows(rows)
    targets = np.ascontiguousarray(targets, dtype=np.int32)
    n = targets.size
    w = np.zeros(n, np.uint32)
    f = np.zeros(n, np.uint64)
    m = np.zeros(n, np.uint64)
    call("bmm_hip_gray_coset_min", rows, rows.shape[0], rows.shape[1], block_rows, targets, n,
         w, f, m)
    return w, f, m


def replay(rows, ops):
    """Row-XOR replay on the GPU; returns (new rows, total weight)."""
    rows = np.array(rows, dtype=np.uint64, copy=True, order="C")
    ops = np.ascontiguousarray(ops, dtype=np.uint16).reshape(-1)
    wt = C.c_uint64()
    call("bmm_hip_replay", rows, rows.shape[0], rows.shape[1], ops, ops.size // 2, C.byref(wt))
    return rows, wt.value


def gen_bernoulli(seed: int, n_rows: int, cols: int, p: float) -> np.ndarray:
    out = np.zeros((n_rows, (cols + 63) // 64), dtype=np.uint64)
    call("bmm_gen_bernoulli", seed, n_rows, cols, p, out)
    return out


def gen_planted(seed: int, n_rows: int, cols: int, p: float):
    w, wn = (cols + 63) // 64, (n_rows + 63) // 64
    S = np.zeros((n_rows, w), np.uint64)
    A = np.zeros((n_rows, w), np.uint64)
    M = np.zeros((n_rows, wn), np.uint64)
    call("bmm_gen_planted", seed, n_rows, cols, p, S, A, M)
    return S, A, M


def gen_batch(seed: int, n_mat: int):
    rows = np.zeros(n_mat, np.int32)
    cols = np.zeros(n_mat, np.int32)
    dens = np.zeros(n_mat, np.float64)
    offs = np.zeros(n_mat, np.int64)
    nw = _sz()
    cap = n_mat * 28 * 32
    words = np.zeros(cap, np.uint64)
    call("bmm_gen_batch", seed, n_mat, rows, cols, dens, words.ctypes.data, cap, offs,
         C.byref(nw))
    return rows, cols, dens, words[:nw.value].copy(), offs


# ---- GA / MLP ---------------------------------------------------------------
def ga_default_params(**overrides) -> GaParams:
    p = GaParams()
    call("bmm_hip_ga_default_params", C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class GpuGA:
    """Island-model GA over row-XOR sequences (bmm_hip_ga_*)."""

    def __init__(self, matrix, cols: int, params: GaParams):
        m = _rows(matrix)
        self.rows, self.cols = m.shape[0], cols
        self.h = C.c_void_p()
        call("bmm_hip_ga_create", m, m.shape[0], cols, C.byref(params), C.byref(self.h))

    def run(self, generations: int, target: int = 0, budget_s: float = 0.0) -> GaStatus:
        st = GaStatus()
        call("bmm_hip_ga_run", self.h, generations, target, budget_s, C.byref(st))
        return st

    def best(self):
        out = np.zeros((self.rows, self.cols // 64), np.uint64)
        w = C.c_uint64()
        call("bmm_hip_ga_best", self.h, out, C.byref(w))
        return out, w.value

    def island_best(self, island: int):
        """Island `island`'s best matrix and weight."""
        out = np.zeros((self.rows, self.cols // 64), np.uint64)
        w = C.c_uint64()
        call("bmm_hip_ga_island_best", self.h, island, out, C.byref(w))
        return out, w.value

    def configure_islands(self, seeds=None, migrate=None):
        """Per-island seeds and migration flags (bit 0 export, bit 1 import)."""
        n = None
        sp = mp = None
        if seeds is not None:
            sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
            n, sp = sd.size, sd.ctypes.data
        if migrate is not None:
            mg = np.ascontiguousarray(np.asarray(migrate, dtype=np.int32))
            n, mp = mg.size, mg.ctypes.data
        call("bmm_hip_ga_configure_islands", self.h, sp, mp, n or 0)

    def best_ops(self, island: int):
        """(base matrix, best op sequence as (src, dst) rows) of an island."""
        base = np.zeros((self.rows, self.cols // 64), np.uint64)
        ops = np.zeros(4096, np.uint16)
        n = _sz()
        call("bmm_hip_ga_best_ops", self.h, island, base, ops, C.byref(n))
        o = ops[:n.value]
        return base, np.stack([o & 0xff, o >> 8], axis=1).astype(np.uint16)

    def proposals(self, island: int) -> np.ndarray:
        """The island's ranked candidate pairs (src | dst << 8)."""
        out = np.zeros(1024, np.uint16)
        n = _sz()
        call("bmm_hip_ga_proposals", self.h, island, out, out.size, C.byref(n))
        return out[:n.value].copy()

    def set_island_base(self, island: int, matrix):
        call("bmm_hip_ga_set_island_base", self.h, island, _rows(matrix))

    def population(self, island: int, pop: int, max_ops: int) -> dict:
        """Next / previous population of an island, the previous one's
        fitness, the individual random streams and the island counters."""
        on = np.zeros((pop, max_ops), np.uint16)
        ln = np.zeros(pop, np.int32)
        op = np.zeros((pop, max_ops), np.uint16)
        lp = np.zeros(pop, np.int32)
        fp = np.zeros(pop, np.uint32)
        rng = np.zeros(pop, np.uint64)
        st = GaIslandState()
        call("bmm_hip_ga_population", self.h, island, on.ctypes.data, ln.ctypes.data,
             op.ctypes.data, lp.ctypes.data, fp.ctypes.data, rng.ctypes.data, C.byref(st))
        return {"ops_next": on, "len_next": ln, "ops_prev": op, "len_prev": lp, "fit_prev": fp,
                "rng": rng, "best_weight": st.best_weight, "best_len": st.best_len,
                "no_improvement": st.no_improvement, "catastrophe": st.catastrophe,
                "mutation_rate": st.mutation_rate, "migrate": st.migrate}

    def close(self):
        if self.h:
            call("bmm_hip_ga_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuMlp:
    """Proposal MLP with host-side features (bmm_hip_mlp_*)."""

    def __init__(self, seed: int = 1):
        self.h = C.c_void_p()
        call("bmm_hip_mlp_create", seed, C.byref(self.h))
        self.n = load().bmm_hip_mlp_num_params()

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n
        call("bmm_hip_mlp_set_params", self.h, p)

    def get_params(self):
        p = np.zeros(self.n, np.float32)
        call("bmm_hip_mlp_get_params", self.h, p)
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[0], np.float32)
        call("bmm_hip_mlp_forward", self.h, x, x.shape[0], y)
        return y

    def score_pairs(self, sketch, pairs):
        """sketch: rows x 2 uint64 (128-bit row sketches); pairs: uint16
        src | dst << 8.  Returns the model output per pair (float32)."""
        sketch = np.ascontiguousarray(sketch, dtype=np.uint64).reshape(-1, 2)
        pairs = np.ascontiguousarray(pairs, dtype=np.uint16).reshape(-1)
        y = np.zeros(pairs.size, np.float32)
        call("bmm_hip_mlp_score_pairs", self.h, sketch, sketch.shape[0], pairs, pairs.size, y)
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n, np.float32)
        loss = C.c_float()
        call("bmm_hip_mlp_gradients", self.h, x, t, x.shape[0], g, C.byref(loss))
        return loss.value, g

    def train_step(self, x, t, lr=1e-3):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        call("bmm_hip_mlp_train_step", self.h, x, t, x.shape[0], lr, C.byref(loss))
        return loss.value

    def close(self):
        if self.h:
            call("bmm_hip_mlp_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Benchmark: subset evaluations/s of the exact row-XOR searches on one MI355X.

Workload (default, --workload gray): 32 x 1024 bit matrix, Bernoulli(0.3) bits
from XorShift64(0x5EED0002); for every target row t, all subsets of the pivot
block rows [0, 24) minus t are visited in reflected-Gray order, one row XOR
plus one popcount per step; per target the minimum weight, the first Gray step
reaching it and the lowest block mask reaching it are produced.  Subset
evaluations per pass = 8 * 2^24 + 24 * 2^23 = 335,544,320.

Other workloads (one JSON line each, same fields):
  --workload coset   20 x 128 Bernoulli(0.5), seed 0x5EED0001: every row
                     against the other 19 (20 * 2^19 subset evaluations).
  --workload batch   4096 matrices (rows 8..28, cols 64..2048, density
                     0.1..0.5, seed 0x5EED0004): every row of every matrix
                     against the matrix's other rows, one persistent launch.
  --workload window  planted 48 x 4096 (Bernoulli(0.02) times a random
                     invertible 48 x 48, seed 0x5EED0003): per target, the
                     target plus the 40 lowest-index other rows (2^40 subsets);
                     a step is one target.
  --workload basis   ExactBasisSolver on a 24 x 1024 Bernoulli(0.3) window
                     (seed 0x5EED0007): all 2^24 - 1 row combinations weighed,
                     minimum-total-weight basis; a step is one solve.
  --workload scores  calc_scores_and_order on 2^28 int32 entries uniform in
                     [-1023, 1023] (seed 0x5EED0010), n = 1023: scores,
                     histogram and stable order of 2^28 - 1 masks; a step is
                     one call from host arrays, the roofline on the kernels.
  --workload sa      SASelector's pairwise synergy table (pair_overlaps) on
                     4096 x 8192 Bernoulli(0.3) (seed 0x5EED0011); value in
                     row pairs/s.
  --workload comb    fixed_weight_masks_u32(32, 16): 601,080,390 masks;
                     value in masks/s from host arrays, roofline on the kernel.
  --workload ga      plain vs learned island GA on the planted 256 x 8192.

A step is one full pass of the workload with the inputs already resident in
HBM.  Timing uses HIP events on the stream the kernels are launched on; the
roofline figure divides the dominant kernel's algorithmic work by its own
average duration, measured with events the library records around that
kernel on the same stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "subset evals/s (FWHT, Gray); GA time-to-1%-of-planted, learned vs plain"
SEED_GRAY_32X1024 = 0x5EED0002
ROWS, COLS, P, BLOCK = 32, 1024, 0.3, 24
GRAY_OPS_PER_EVAL = 66      # 32 v_xor_b32 + 32 v_bcnt_u32_b32 + 2 key minima
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # INT32 VALU, non-packed
VALU_PK16_PEAK_TOPS = 2 * VALU_PEAK_TOPS   # packed int16 (v_pk_add_u16: 2 ops per lane)

# HBM bytes per launch of each workload's dominant kernel, from rocprofv3
# FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch), each scaled by what the
# same counter reads for a plain 256 MiB device copy (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0).  Re-measure when the kernel changes.
TRAFFIC_BYTES = {
    "gray_coset_sweep": (192.5 * 2.0 + 160.0) * 1024,
    "coset_batch_4096": (53729.0 * 2.0 + 29953.0) * 1024,
    "window_coset_planted_48x4096": (118.4 * 2.0 + 131137.0) * 1024,
    # The kernels of one calc_scores_and_order call on 2^28 entries (scores,
    # the three table-scan kernels, the tile order), FETCH_SIZE x 2.0 plus
    # WRITE_SIZE x 1.0 in KiB.
    # The array FWHT of 2^28 int32 (fwht_lo16_kernel + fwht_hi12_kernel).
    "fwht_array_2p28": ((524350.0 + 527800.0) * 2.0 + 1048576.0 + 1057000.0) * 1024,
    "scores_order_2p28": ((524412.9 + 16406.5 + 261.3 + 16670.0 + 557224.6) * 2.0
                          + 1085440.1 + 512.0 + 512.0 + 65536.0 + 1076899.8) * 1024,
}


def gray_sweep_evals() -> int:
    return sum((1 << (BLOCK - 1)) if t < BLOCK else (1 << BLOCK) for t in range(ROWS))


def host_cpu() -> dict:
    """The host the CPU baseline ran on: CPU model (as lscpu prints it, from
    /proc/cpuinfo), logical CPUs in the machine and in this process's
    affinity mask."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)),
            "logical_cpus": os.cpu_count()}


def _omp_threads(n: int) -> None:
    """Sets the OpenMP team size of the reference build (libgomp is shared by
    the process)."""
    import ctypes

    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(matrix: np.ndarray, budget_s: float = 4.0, one_budget_s: float = 8.0) -> dict:
    """Reference CPU brute force (BruteforceSearch, compiled from the
    reference sources into oracle/_ref by oracle/Makefile) on windows
    [block rows..., target], k = 1: each window visits all 2^25 masks, OpenMP
    over prefixes.  Timed with all the process's cores (windows repeated for
    about `budget_s` seconds) and with one thread (`one_budget_s`).  Where
    the reference build is absent the project's C restatement (one thread)
    runs instead, and `kind` / `path` say so."""
    import oracle  # checker / baseline only

    host = host_cpu()
    ref = oracle.reference()
    wins = [np.concatenate([matrix[:BLOCK], matrix[t:t + 1]]) for t in range(BLOCK, ROWS)]
    if ref is not None:
        def run(threads: int, budget: float):
            _omp_threads(threads)
            evals, done = 0, 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < budget:
                ref.bruteforce_search(wins[done % len(wins)], COLS, 1)
                evals += 1 << (BLOCK + 1)
                done += 1
            return evals / (time.perf_counter() - t0), done

        threads = host["nproc"]
        v_all, n_all = run(threads, budget_s)
        v_one, n_one = run(1, one_budget_s)
        _omp_threads(threads)
        return {"value": v_all, "unit": "subset evals/s", "cores": threads, "kind": "reference",
                "path": "oracle/_ref/libbmmref.so: reference BruteforceSearch",
                "sample": f"reference BruteforceSearch(k=1) on windows of 25 rows (24-row pivot "
                          f"block + target) x 1024 cols, 2^25 masks each, OpenMP over prefixes: "
                          f"{n_all} windows on {threads} threads, {n_one} on 1 thread",
                "one_thread": {"value": v_one, "windows": n_one}, **host}
    ora = oracle.oracle()
    t0 = time.perf_counter()
    evals = 0
    done = 0
    while time.perf_counter() - t0 < one_budget_s:
        ora.gray_min(matrix, BLOCK, BLOCK + done % (ROWS - BLOCK))
        evals += 1 << BLOCK
        done += 1
    dt = time.perf_counter() - t0
    return {"value": evals / dt, "unit": "subset evals/s", "cores": 1, "kind": "port",
            "path": "oracle/liboracle.so: C restatement (reference build absent on this host)",
            "sample": f"C restatement Gray sweep, {done} targets x 2^24 subsets, one thread",
            "one_thread": {"value": evals / dt, "windows": done}, **host}


def _time(torch, stream, fn, reps):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def batch_inputs(_lib, name: str):
    """(rows, cols, words, word offsets) of the coset-minimum workloads."""
    if name == "coset":
        m = _lib.gen_bernoulli(0x5EED0001, 20, 128, 0.5)
        return (np.array([20], np.int32), np.array([128], np.int32), m.reshape(-1),
                np.array([0], np.int64))
    rows, cols, _, words, offs = _lib.gen_batch(0x5EED0004, 4096)
    return rows, cols, words, offs


def transform_ops(rows) -> float:
    """Algorithmic int32 ops of one pass: one shared 2^r Walsh-Hadamard
    transform per matrix, r add/sub per element (r radix-2 stages)."""
    return float(sum(int(r) * (1 << int(r)) for r in rows))


def run_batch_plan(torch, _lib, dev, stream, name: str, steps: int, warmup: int,
                   persistent_wgs: int = 0) -> dict:
    """Every row of every matrix against the matrix's other rows: one plan
    (autotuned launch shape), replayed as a captured hipGraph per pass.
    Subset evaluations are counted as sum over matrices of rows * 2^(rows-1)."""
    import ctypes as C

    sp = stream.cuda_stream
    rows, cols, words, offs = batch_inputs(_lib, name)
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    n_slots = int(rows.sum())
    d_words = torch.from_numpy(words.view(np.int64).copy()).to(dev)
    d_w = torch.empty(n_slots, dtype=torch.int32, device=dev)
    d_i = torch.empty(n_slots, dtype=torch.int64, device=dev)
    plan = C.c_void_p()
    t0 = time.perf_counter()
    _lib.call("bmm_hip_batch_plan_create", d_words.data_ptr(), offs, rows, cols, rows.size,
              out_off, d_w.data_ptr(), d_i.data_ptr(), persistent_wgs, int(persistent_wgs == 0),
              C.byref(plan))
    plan_s = time.perf_counter() - t0
    info = _lib.BatchInfo()
    _lib.call("bmm_hip_batch_plan_info", plan, C.byref(info))
    try:
        for _ in range(warmup):
            _lib.call("bmm_hip_batch_plan_run", plan, 1, sp)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            _lib.call("bmm_hip_batch_plan_run", plan, 1, sp)
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        # The same passes launched directly (no graph), back to back.
        a.record(stream)
        for _ in range(steps):
            _lib.call("bmm_hip_batch_plan_run", plan, 0, sp)
        b.record(stream)
        torch.cuda.synchronize()
        ms_eager = a.elapsed_time(b) / steps
        # Kernel-level time of the persistent transform kernel (eager launches,
        # events around that kernel only).
        _lib.timing_begin()
        for _ in range(max(1, min(steps, 10))):
            _lib.call("bmm_hip_batch_plan_run", plan, 0, sp)
        torch.cuda.synchronize()
        kms, nk = _lib.timing_end()
        weights = d_w.cpu().numpy()
    finally:
        _lib.call("bmm_hip_batch_plan_destroy", plan)
    kernel_avg_ms = kms / max(nk, 1)
    ops = transform_ops(rows)
    achieved = ops / (kernel_avg_ms / 1e3) / 1e12
    # 2^14 blocks run packed int16 (two add/sub per lane per instruction).
    packed = info.max_block_bits == 14
    peak = VALU_PK16_PEAK_TOPS if packed else VALU_PEAK_TOPS
    return {
        "ms": ms, "ms_eager": ms_eager, "evals": info.subset_evals, "weights": weights,
        "config": {"workload": "coset_20x128" if name == "coset" else "coset_batch_4096",
                   "matrices": int(rows.size), "rows_total": n_slots,
                   "seed": "0x5eed0001" if name == "coset" else "0x5eed0004",
                   "subset_evals_per_step": info.subset_evals,
                   "transform_elements_per_step": info.transform_elements,
                   "persistent_wgs": info.persistent_wgs, "chunk": info.chunk,
                   "max_block_bits": info.max_block_bits, "total_blocks": info.total_blocks,
                   "plan_create_s": plan_s},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "TOPS (packed int16 VALU)" if packed else "TOPS (int32 VALU)",
                     "frac": achieved / peak,
                     "traffic": TRAFFIC_BYTES.get("coset_batch_4096") if name == "batch" else None,
                     "ops_per_pass": ops,
                     "kernel": f"coset_full_kernel<{info.max_block_bits}>",
                     "kernel_avg_ms": kernel_avg_ms, "kernel_launches": nk},
    }


def batch_cpu_baseline(_lib, name: str, budget_s: float = 10.0) -> dict:
    """The project's C restatement of the per-target coset minimum (histogram
    of the other rows' column patterns, in-place FWHT, scan; one target at a
    time, one thread) over the workload's matrices in order; for the batch,
    matrices of up to 22 rows (larger ones take minutes each on one core)."""
    import oracle  # checker / baseline only

    ora = oracle.oracle()
    rows, cols, words, offs = batch_inputs(_lib, name)
    evals = 0.0
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 <= budget_s:   # the 20 x 128 case repeats
        for i in range(rows.size):
            r, c = int(rows[i]), int(cols[i])
            if r > 22:
                continue
            w = (c + 63) // 64
            m = words[offs[i]:offs[i] + r * w].reshape(r, w)
            ora.coset_min(m, c)
            evals += r * 2.0 ** (r - 1)
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
    dt = time.perf_counter() - t0
    return {"value": evals / dt, "unit": "subset evals/s", "cores": 1, "kind": "port",
            "sample": f"C restatement coset minimum on {done} matrix passes "
                      f"({'20 x 128' if name == 'coset' else 'rows <= 22 of the batch'}), "
                      f"per target histogram + FWHT + scan, one thread"}


WIN_SEED, WIN_ROWS, WIN_COLS, WIN_P, WIN_WIDTH = 0x5EED0003, 48, 4096, 0.02, 40


def window_inputs(_lib):
    S, A, _ = _lib.gen_planted(WIN_SEED, WIN_ROWS, WIN_COLS, WIN_P)
    wins = np.array([[j for j in range(WIN_ROWS) if j != t][:WIN_WIDTH] for t in range(WIN_ROWS)],
                    np.int32)
    return S, A, wins


def run_window(torch, _lib, steps: int, warmup: int) -> dict:
    """Per target (targets 0, 1, ... one per step): min over the 2^40 subsets
    of the 40 lowest-index other rows of wt(row_t ^ XOR(subset)), with the
    planted row weight wt(S_t) beside it."""
    S, A, wins = window_inputs(_lib)
    planted_w = [int(sum(bin(int(v)).count("1") for v in S[t])) for t in range(WIN_ROWS)]
    for i in range(warmup):  # warm-up on a 30-row window (same kernel, 2^30 masks)
        _lib.window_coset_min(A, WIN_COLS, [i % WIN_ROWS], wins[i % WIN_ROWS:i % WIN_ROWS + 1, :30])
    torch.cuda.synchronize()
    results = []
    _lib.timing_begin()
    t0 = time.perf_counter()
    for i in range(steps):
        t = i % WIN_ROWS
        w, m = _lib.window_coset_min(A, WIN_COLS, [t], wins[t:t + 1])
        results.append({"target": t, "min_weight": int(w[0]), "mask": hex(int(m[0])),
                        "planted_row_weight": planted_w[t],
                        "start_row_weight": int(sum(bin(int(v)).count("1") for v in A[t]))})
    dt = time.perf_counter() - t0
    kms, nk = _lib.timing_end()
    kernel_avg_ms = kms / max(nk, 1)
    evals = float(1 << WIN_WIDTH)
    # 4096 columns (at most 32767 active): packed-int16 blocks of 2^14 masks,
    # a 14-stage transform per block (14 add/sub per mask, two per lane op).
    ops = evals * 14
    achieved = ops / (kernel_avg_ms / 1e3) / 1e12
    return {
        "ms": 1e3 * dt / steps, "evals": evals, "results": results,
        "config": {"workload": "window_coset_planted_48x4096", "rows": WIN_ROWS,
                   "cols": WIN_COLS, "p": WIN_P, "window": WIN_WIDTH, "seed": hex(WIN_SEED),
                   "subset_evals_per_step": evals, "targets_timed": steps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PK16_PEAK_TOPS,
                     "unit": "TOPS (packed int16 VALU)", "frac": achieved / VALU_PK16_PEAK_TOPS,
                     "traffic": TRAFFIC_BYTES["window_coset_planted_48x4096"],
                     "ops_per_eval": 14, "kernel": "win_full_kernel<14>",
                     "kernel_avg_ms": kernel_avg_ms, "kernel_launches": nk},
    }


def window_cpu_baseline(_lib, budget_s: float = 20.0) -> dict:
    """Reference MitmFwhtSearch (compiled from the reference sources, serial)
    with k = 1 on 30-row windows of the planted matrix (target + its 29
    lowest-index other rows): 2^30 masks each, timed whole.  The full 2^40
    windows take hours per target on the CPU."""
    import oracle  # checker / baseline only

    _, A, wins = window_inputs(_lib)
    ref = oracle.reference()
    ora = oracle.oracle()
    evals = 0.0
    done = 0
    t0 = time.perf_counter()
    for t in range(WIN_ROWS):
        win = np.concatenate([A[t:t + 1], A[wins[t, :29]]])
        if ref is not None:
            ref.mitm_search(win, WIN_COLS, 1)
        else:
            ora.mitm_topk(win, WIN_COLS, 1)
        evals += 2.0 ** 30
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    kind = "reference" if ref is not None else "port"
    return {"value": evals / dt, "unit": "subset evals/s", "cores": 1, "kind": kind,
            "sample": f"{'reference' if ref is not None else 'C restatement'} MitmFwhtSearch "
                      f"k=1 on {done} windows of 30 rows x 4096 cols (2^30 masks each), serial"}


BASIS_SEED, BASIS_ROWS, BASIS_COLS, BASIS_P = 0x5EED0007, 24, 1024, 0.3


FWHT_HBM_BITS = 28


def run_basis(torch, steps: int, warmup: int) -> dict:
    """ExactBasisSolver on a 24 x 1024 Bernoulli(0.3) window: every one of the
    2^24 - 1 nonzero row combinations is weighed (one FWHT of 2^24 int32 in
    HBM), the (weight, Gray index) stream is selected on the GPU and the
    basis assembled on the host.  A step is one solve."""
    import bmmpy_amd as bmm
    from bmmpy_amd import _lib

    m = _lib.gen_bernoulli(BASIS_SEED, BASIS_ROWS, BASIS_COLS, BASIS_P)
    rows = ["".join("1" if (int(r[c // 64]) >> (c % 64)) & 1 else "0" for c in range(BASIS_COLS))
            for r in m]
    mat = bmm.matrix_from_rows(rows)
    win = mat.row_window(list(range(BASIS_ROWS)))
    solver = bmm.solver.ExactBasisSolver()
    for _ in range(warmup):
        solver.solve(win)
    _lib.timing_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = solver.solve(win)
    dt = (time.perf_counter() - t0) / steps
    kms, nk = _lib.timing_end()
    kernel_avg_ms = kms / max(nk, 1)
    n = 1 << BASIS_ROWS
    # The solve's 2^24-entry transform (64 MiB) stays in the 256 MiB
    # last-level cache, so it is not an HBM measurement.  The array FWHT's
    # HBM roofline is taken on 2^28 int32 (1 GiB).  Algorithmic bytes: 16 per
    # element (two out-of-LDS passes, each one read and one write of 4
    # bytes).  The kernels take exactly two passes at this size (index bits
    # 0-15 in 2^16-element register tiles, then bits 16-27), 8 bytes per
    # element each.
    big = torch.zeros(1 << FWHT_HBM_BITS, dtype=torch.int32, device="cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    fwht_bytes = 16.0 * (1 << FWHT_HBM_BITS)
    fwht_ms = _time(torch, stream,
                    lambda: _lib.call("bmm_hip_fwht_i32_dev", big.data_ptr(), big.numel(), sp),
                    10)
    achieved = fwht_bytes / (fwht_ms / 1e3) / 1e9
    # Practical ceiling: a plain device copy of the same 1 GiB (read + write).
    dst = torch.empty_like(big)
    copy_ms = _time(torch, stream, lambda: dst.copy_(big), 10)
    copy_gbps = 2.0 * 4 * (1 << FWHT_HBM_BITS) / (copy_ms / 1e3) / 1e9
    del big, dst
    return {
        "ms": 1e3 * dt, "evals": float(n - 1),
        "results": [{"rank": int(res.rank), "total_weight": int(res.total_weight),
                     "basis_weights": [int(x) for x in res.basis_weights]}],
        "config": {"workload": "exact_basis_24x1024", "rows": BASIS_ROWS, "cols": BASIS_COLS,
                   "p": BASIS_P, "seed": hex(BASIS_SEED), "subset_evals_per_step": n - 1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0,
                     "traffic": TRAFFIC_BYTES["fwht_array_2p28"],
                     "bytes_per_launch": fwht_bytes,
                     "kernel": f"array FWHT of 2^{FWHT_HBM_BITS} int32: fwht_lo16_kernel "
                               "+ fwht_hi12_kernel",
                     "kernel_avg_ms": fwht_ms,
                     "solve_fwht_2^24_avg_ms": kernel_avg_ms, "solve_fwht_launches": nk,
                     "copy_1gib_gbps": copy_gbps,
                     "solve_ms": 1e3 * dt},
    }


def basis_cpu_baseline(budget_s: float = 15.0) -> dict:
    """The reference ExactBasisSolver (its scalar build) on the same window."""
    import oracle  # checker / baseline only
    from bmmpy_amd import _lib

    m = _lib.gen_bernoulli(BASIS_SEED, BASIS_ROWS, BASIS_COLS, BASIS_P)
    ref = oracle.reference(scalar=True)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        if ref is not None:
            ref.exact_basis(m, BASIS_COLS)
        else:
            oracle.oracle().exact_basis(m, BASIS_COLS)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done * float((1 << BASIS_ROWS) - 1) / dt, "unit": "subset evals/s",
            "cores": 1, "kind": "reference" if ref is not None else "port",
            "sample": f"{'reference' if ref is not None else 'C restatement'} ExactBasisSolver "
                      f"on the 24 x 1024 window, {done} solves (two Gray enumerations of "
                      f"2^24 - 1 masks each)"}


SCORES_BITS, SCORES_N, SCORES_SEED = 28, 1023, 0x5EED0010
SA_ROWS, SA_COLS, SA_P, SA_SEED = 4096, 8192, 0.3, 0x5EED0011


def run_scores(torch, steps: int, warmup: int) -> dict:
    """calc_scores_and_order on a spectrum of 2^28 int32 entries (scores in
    [0, 1023], the range of a 1023-column window): scores, score histogram
    and the stable order of masks 1 .. m-1.  A step is one call; the roofline
    is taken on the kernels' time.  Algorithmic bytes: read h, write the
    scores and the order, 12 B per mask."""
    from bmmpy_amd import _lib

    rng = np.random.default_rng(SCORES_SEED)
    m = 1 << SCORES_BITS
    h = rng.integers(-SCORES_N, SCORES_N + 1, m, dtype=np.int32)
    for _ in range(warmup):
        _lib.calc_scores_and_order(h, SCORES_N)
    _lib.timing_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        s, order, cnt, off = _lib.calc_scores_and_order(h, SCORES_N)
    dt = (time.perf_counter() - t0) / steps
    kms, nk = _lib.timing_end()
    kernel_ms = kms / max(steps, 1)   # the kernels of one call (two timed regions)
    alg = 12.0 * m
    achieved = alg / (kernel_ms / 1e3) / 1e9
    # Moved by the kernels (10-bit scores, one-pass tile order): scores (read
    # h, write s: 8 B, plus the per-tile 16-bit score counts, 0.125 B), the
    # table scan (read the counts twice, write the 32-bit starts: 0.5 B) and
    # the tile order (read s, write the order: 8 B, plus the tile's starts):
    # about 17 B per mask.  rocprof's FETCH_SIZE and WRITE_SIZE of the
    # kernels (copy-calibrated) give the figure used.
    moved = TRAFFIC_BYTES["scores_order_2p28"]
    return {
        "ms": 1e3 * dt, "evals": float(m - 1),
        "results": [{"cnt_head": [int(x) for x in cnt[:4]], "order_head": [int(x) for x in order[:4]],
                     "off_last": int(off[-1])}],
        "config": {"workload": "scores_order_2^28", "entries": m, "n": SCORES_N,
                   "seed": hex(SCORES_SEED), "subset_evals_per_step": m - 1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": moved, "bytes_per_launch": alg,
                     "kernel": "scores_tile_kernel + col_sum/col_base/col_start_kernel "
                               "+ order_tile_kernel",
                     "kernel_avg_ms": kernel_ms, "host_call_ms": 1e3 * dt},
    }


def scores_cpu_baseline(budget_s: float = 10.0) -> dict:
    """The reference calc_scores_and_order on 2^24-entry samples."""
    import oracle  # checker / baseline only

    ref = oracle.reference()
    rng = np.random.default_rng(SCORES_SEED)
    h = rng.integers(-SCORES_N, SCORES_N + 1, 1 << 24, dtype=np.int32)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        if ref is not None:
            ref.scores_order(h, SCORES_N)
        else:
            oracle.oracle().scores_order(h, SCORES_N)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done * float((1 << 24) - 1) / dt, "unit": "subset evals/s", "cores": 1,
            "kind": "reference" if ref is not None else "port",
            "sample": f"calc_scores_and_order on 2^24 entries, {done} calls"}


def run_sa(torch, steps: int, warmup: int) -> dict:
    """SASelector's pairwise synergy table on 4096 rows x 8192 columns
    (Bernoulli(0.3)): popcount(r_i & r_j) for every pair.  Roofline: 4 int32
    VALU ops per pair and 64-bit word (AND of two halves, two bit counts),
    over the upper triangle."""
    from bmmpy_amd import _lib

    m = _lib.gen_bernoulli(SA_SEED, SA_ROWS, SA_COLS, SA_P)
    for _ in range(warmup):
        _lib.pair_overlaps(m)
    _lib.timing_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = _lib.pair_overlaps(m)
    dt = (time.perf_counter() - t0) / steps
    kms, nk = _lib.timing_end()
    kernel_ms = kms / max(nk, 1)
    pairs = SA_ROWS * (SA_ROWS + 1) / 2
    ops = pairs * (SA_COLS // 64) * 4
    achieved = ops / (kernel_ms / 1e3) / 1e12
    return {
        "ms": 1e3 * dt, "evals": pairs,
        "results": [{"overlap_0_1": int(out[0, 1]), "trace": int(np.trace(out))}],
        "config": {"workload": "pair_overlaps_4096x8192", "rows": SA_ROWS, "cols": SA_COLS,
                   "p": SA_P, "seed": hex(SA_SEED)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": 39.3216,
                     "unit": "TOPS (int32 VALU)", "frac": achieved / 39.3216, "traffic": None,
                     "kernel": "overlap_kernel", "kernel_avg_ms": kernel_ms,
                     "host_call_ms": 1e3 * dt},
    }


def sa_cpu_baseline(budget_s: float = 10.0) -> dict:
    """The reference SASelector (scalar build), one restart of one iteration
    on 1024 x 8192 samples: its time is the pairwise synergy table."""
    import oracle  # checker / baseline only
    from bmmpy_amd import _lib

    ref = oracle.reference(scalar=True)
    m = _lib.gen_bernoulli(SA_SEED, 1024, SA_COLS, SA_P)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        if ref is None:
            break
        ref.sa_select(m, SA_COLS, 16, iterations=1, restarts=1, seed=1,
                      policy="pairwise_synergy")
        done += 1
    dt = time.perf_counter() - t0
    if ref is None or done == 0:
        return None
    return {"value": done * 1024 * 1025 / 2 / dt, "unit": "row pairs/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference SASelector.select on 1024 x 8192 (iterations=1, restarts=1: "
                      f"one pairwise synergy table), {done} calls"}


COMB_N, COMB_K = 32, 16


def run_comb(torch, steps: int, warmup: int) -> dict:
    """fixed_weight_masks_u32(32, 16): the C(32, 16) = 601,080,390 masks of
    32 bits with 16 set, increasing.  A step is one call from the host; the
    roofline is the kernel's store of 4 bytes per mask."""
    from bmmpy_amd import _lib

    for _ in range(warmup):
        _lib.fixed_weight_masks(COMB_N, COMB_K)
    _lib.timing_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = _lib.fixed_weight_masks(COMB_N, COMB_K)
    dt = (time.perf_counter() - t0) / steps
    kms, nk = _lib.timing_end()
    kernel_ms = kms / max(nk, 1)
    count = out.size
    alg = 4.0 * count
    achieved = alg / (kernel_ms / 1e3) / 1e9
    return {
        "ms": 1e3 * dt, "evals": float(count),
        "results": [{"count": int(count), "first": int(out[0]), "last": int(out[-1])}],
        "config": {"workload": "fixed_weight_masks_u32_32_16", "n": COMB_N, "k": COMB_K},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": alg, "bytes_per_launch": alg,
                     "kernel": "fixed_weight_kernel<uint32_t>", "kernel_avg_ms": kernel_ms,
                     "host_call_ms": 1e3 * dt},
    }


def comb_cpu_baseline(budget_s: float = 10.0) -> dict:
    """The reference fixed_weight_masks_u32 on C(28, 14) samples."""
    import math
    import oracle  # checker / baseline only

    ref = oracle.reference()
    if ref is None:
        return None
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        ref.fixed_weight_masks(28, 14)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done * math.comb(28, 14) / dt, "unit": "masks/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference fixed_weight_masks_u32(28, 14) (40,116,600 masks), {done} "
                      f"calls, copied out as uint64"}


def _window_of(m: np.ndarray, cols: int):
    import bmmpy as bmm
    rows = ["".join("1" if (int(r[c // 64]) >> (c % 64)) & 1 else "0" for c in range(cols))
            for r in m]
    return bmm.matrix_from_rows(rows).row_window(list(range(m.shape[0])))


TOPK_L = 13   # fwht_topk_kernel: blocks of 2^13 masks, int32 transform in LDS
TOPK_PK_L = 14   # fwht_topk_pk_kernel: blocks of 2^14 masks, packed int16


def topk_packed(t: int, cols: int) -> bool:
    """Whether the top-k engine runs a t-row window over `cols` columns as
    packed-int16 blocks of 2^14 masks (at least 256 blocks, at most 32767
    active columns) rather than int32 blocks of 2^13."""
    return t >= 22 and cols <= 32767


def run_search(torch, steps: int, warmup: int) -> dict:
    """The reference-API top-k searchers (FwhtSearch engine: persistent
    workgroups over blocks of masks, exact per-workgroup top-K, a tree of
    merges).  Cases: MitmFwhtSearch and CudaMitmFwhtSearch on 28 x 96
    Bernoulli(0.5) with k = 128 (2^28 masks), MitmFwhtSearch with k = 64 on
    rows 0..39 of the planted 48 x 4096 instance (2^40 masks; the value),
    BruteforceSearch and CudaBruteforceSearch on 16 x 4096 Bernoulli(0.5)
    with k = 128.  Per case: host call time and kernel time per call (events
    around the block kernel and, for windows whose workgroups see few
    blocks, the sampled bound pass before it); roofline: the block transform's add/sub per mask, 14
    against the packed-int16 peak for windows run as packed blocks of 2^14,
    min(t, 13) against the int32 peak otherwise."""
    import bmmpy as bmm
    from bmmpy_amd import _lib

    S, A, _ = _lib.gen_planted(WIN_SEED, WIN_ROWS, WIN_COLS, WIN_P)
    cases = [
        ("mitm_28x96_k128", bmm.MitmFwhtSearch(max_candidates=128),
         _lib.gen_bernoulli(0x5EED0012, 28, 96, 0.5), 96, steps),
        ("cuda_mitm_28x96_k128", bmm.CudaMitmFwhtSearch(max_candidates=128),
         _lib.gen_bernoulli(0x5EED0012, 28, 96, 0.5), 96, steps),
        ("mitm_planted_40x4096_k64", bmm.MitmFwhtSearch(max_candidates=64), A[:40], WIN_COLS, 1),
        ("bruteforce_16x4096_k128", bmm.BruteforceSearch(max_candidates=128),
         _lib.gen_bernoulli(0x5EED0013, 16, 4096, 0.5), 4096, steps),
        ("cuda_bruteforce_16x4096_k128", bmm.CudaBruteforceSearch(max_candidates=128),
         _lib.gen_bernoulli(0x5EED0013, 16, 4096, 0.5), 4096, steps),
    ]
    out = {}
    for name, searcher, m, cols, reps in cases:
        win = _window_of(m, cols)
        for _ in range(min(warmup, 2) if reps > 1 else 0):
            searcher.search(win)
        _lib.timing_begin()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = searcher.search(win)
        dt = (time.perf_counter() - t0) / reps
        kms, nk = _lib.timing_end()
        t = m.shape[0]
        masks = float((1 << t) - 1)
        kernel_ms = kms / reps   # per call: the sampled bound pass (if any) + the block pass
        packed = topk_packed(t, cols)
        ops = float(1 << t) * (TOPK_PK_L if packed else min(t, TOPK_L))
        peak = VALU_PK16_PEAK_TOPS if packed else VALU_PEAK_TOPS
        print(f"search {name}: {1e3 * dt:.2f} ms per call, kernel {kernel_ms:.3f} ms",
              file=sys.stderr, flush=True)
        out[name] = {"rows": t, "cols": cols, "k": len(res), "masks": masks,
                     "call_ms": 1e3 * dt, "masks_per_s": masks / dt,
                     "kernel_avg_ms": kernel_ms, "kernel_launches": nk,
                     "kernel_masks_per_s": masks / (kernel_ms / 1e3),
                     "kernel": "fwht_topk_pk_kernel" if packed else "fwht_topk_kernel<13>",
                     "roofline_peak_tops": peak,
                     "roofline_frac": ops / (kernel_ms / 1e3) / 1e12 / peak,
                     "best": [int(res[0].weight), int(res[0].mask_u64())] if res else None}
    main = out["mitm_planted_40x4096_k64"]
    return {
        "ms": main["call_ms"], "evals": main["masks"], "results": [out],
        "config": {"workload": "topk_searchers", "cases": list(out),
                   "subset_evals_per_step": main["masks"]},
        "roofline": {"bound": "valu", "achieved": main["roofline_frac"] * VALU_PK16_PEAK_TOPS,
                     "peak": VALU_PK16_PEAK_TOPS, "unit": "TOPS (packed int16 VALU)",
                     "frac": main["roofline_frac"], "traffic": None,
                     "ops_per_eval": TOPK_PK_L, "kernel": "fwht_topk_pk_kernel<256, 4>",
                     "kernel_avg_ms": main["kernel_avg_ms"]},
    }


def secondary(torch, _lib, dev, stream) -> dict:
    """The other configured workloads, timed beside the headline: the FWHT
    coset minima (graph replay, autotuned launch shape), one 2^40 window of
    the planted 48 x 4096 instance, and the plain and learned island GA on
    the planted 256 x 8192 instance until 1.01 x the planted weight."""
    res = {}
    r = run_window(torch, _lib, 1, 1)
    res["window_48x4096"] = {"subset_evals": r["evals"], "ms_per_target": r["ms"],
                             "subset_evals_per_s": r["evals"] / (r["ms"] / 1e3),
                             "roofline_frac": r["roofline"]["frac"],
                             "kernel_avg_ms": r["roofline"]["kernel_avg_ms"],
                             "result": r["results"][0]}
    g = ga_compare(20.0)
    res["ga_256x8192"] = {
        "planted_weight": g["planted_weight"], "start_weight": g["start_weight"],
        "target": g["target"], "isd_trials_per_epoch": g["isd_trials_per_epoch"],
        "window_rounds_per_epoch": g["window_rounds_per_epoch"],
        "init_rounds": g["init_rounds"],
        **{f"{k}_{f}": g[k][f] for k in ("plain", "learned")
           for f in ("best_weight", "time_to_target_s", "generations", "ms_per_generation",
                     "evals_per_s", "replayed_ops_per_s")},
        "learned_over_plain_time_ratio":
            (g["learned"]["time_to_target_s"] / g["plain"]["time_to_target_s"]
             if g["plain"]["time_to_target_s"] and g["learned"]["time_to_target_s"] else None)}
    for name, key in (("coset", "coset_20x128"), ("batch", "batch_4096")):
        r = run_batch_plan(torch, _lib, dev, stream, name, 50 if name == "coset" else 5, 2)
        res[key] = {"subset_evals": r["evals"], "ms_per_pass": r["ms"],
                    "ms_per_pass_direct_launch": r["ms_eager"],
                    "subset_evals_per_s": r["evals"] / (r["ms"] / 1e3),
                    "transform_elements_per_s":
                        r["config"]["transform_elements_per_step"] / (r["ms"] / 1e3),
                    "roofline_frac": r["roofline"]["frac"],
                    "kernel_avg_ms": r["roofline"]["kernel_avg_ms"],
                    "persistent_wgs": r["config"]["persistent_wgs"],
                    "chunk": r["config"]["chunk"],
                    "max_block_bits": r["config"]["max_block_bits"]}
    return res


def mode_line(args, r: dict, cpu: dict | None) -> dict:
    out = {
        "metric": METRIC, "value": r["evals"] / (r["ms"] / 1e3), "unit": "subset evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i32",
        "data": "synthetic", "config": r["config"], "roofline": r["roofline"],
    }
    if "results" in r:
        out["results"] = r["results"]
    if "ms_eager" in r:
        out["ms_per_step_direct_launch"] = r["ms_eager"]
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def ga_compare(budget_s: float, seed: int = 0x5EED0005, ga_seed: int | None = None,
               **overrides) -> dict:
    """Planted 256 x 8192 matrix (Bernoulli(0.01) sparse S, random invertible
    256 x 256 transform M, A = M S, drawn from `seed`); 64 islands x 256
    individuals x 512 ops, ring migration every 20 generations, GA random
    streams from `ga_seed` (default: `seed`); the library's GA parameters
    with `overrides` on top.  Plain and learned proposal runs get the same
    wall-clock budget; reported: weight reached, time to reach 1.01 * wt(S)
    (null when not reached), the trajectory and the epoch settings used."""
    from bmmpy_amd import _lib

    S, A, _ = _lib.gen_planted(seed, 256, 8192, 0.01)
    ga_seed = seed if ga_seed is None else ga_seed
    planted = int(sum(bin(int(v)).count("1") for v in S.reshape(-1)))
    start_w = int(sum(bin(int(v)).count("1") for v in A.reshape(-1)))
    target = int(planted * 1.01)
    res = {"planted_weight": planted, "start_weight": start_w, "target": target}
    for name, learned in (("plain", 0), ("learned", 1)):
        p = _lib.ga_default_params(learned=learned, seed=ga_seed, **overrides)
        res["isd_trials_per_epoch"] = int(p.isd_trials)
        res["window_rounds_per_epoch"] = int(p.win_rounds)
        res["init_rounds"] = int(p.init_rounds)
        t0 = time.perf_counter()
        ga = _lib.GpuGA(A, 8192, p)
        setup = time.perf_counter() - t0
        traj = []
        reached_at = None
        elapsed = 0.0
        st = None
        while elapsed < budget_s:
            st = ga.run(20, target=target, budget_s=0.0)
            elapsed += st.elapsed_s
            traj.append([round(elapsed, 3), int(st.best_weight)])
            if len(traj) % 25 == 0:
                print(f"ga {name}: {elapsed:.1f}s gen {st.generations} best {st.best_weight}",
                      file=sys.stderr, flush=True)
            if st.reached_target:
                reached_at = elapsed
                break
        best, w = ga.best()
        ga.close()
        res[name] = {"best_weight": int(w), "ratio_to_planted": w / planted,
                     "time_to_target_s": reached_at, "elapsed_s": elapsed,
                     "setup_s": setup, "generations": int(st.generations),
                     "evaluations": int(st.evaluations),
                     "evals_per_s": st.evaluations / max(elapsed, 1e-9),
                     "replayed_ops": int(st.replayed_ops),
                     "replayed_ops_per_s": st.replayed_ops / max(elapsed, 1e-9),
                     "ms_per_generation": 1e3 * elapsed / max(int(st.generations), 1),
                     "train_loss": st.train_loss, "li_moves": int(st.li_moves),
                     "window_moves": int(st.win_moves), "isd_rows": int(st.isd_rows),
                     "trajectory": traj[-50:]}
    return res


def ga_seeds(budget_s: float, n_seeds: int, overrides: dict, learned_overrides: dict,
             marks=(40.0, 20.0, 1.01)) -> dict:
    """Plain and learned GA over `n_seeds` GA seeds on the same planted
    256 x 8192 instance: per variant the time to reach each multiple of the
    planted weight in `marks` (None when not reached within the budget) and
    the weight at equal generations; median and range over the seeds."""
    from bmmpy_amd import _lib

    S, A, _ = _lib.gen_planted(0x5EED0005, 256, 8192, 0.01)
    planted = int(sum(bin(int(v)).count("1") for v in S.reshape(-1)))
    runs = {"plain": [], "learned": []}
    for k in range(n_seeds):
        seed = 0x5EED0005 + 0x1000 * k
        for name, learned in (("plain", 0), ("learned", 1)):
            kw = dict(learned=learned, seed=seed, **overrides)
            if learned:
                kw.update(learned_overrides)
            p = _lib.ga_default_params(**kw)
            ga = _lib.GpuGA(A, 8192, p)
            hit = {m: None for m in marks}
            elapsed, st, traj = 0.0, None, []
            while elapsed < budget_s:
                st = ga.run(20, target=int(planted * min(marks)), budget_s=0.0)
                elapsed += st.elapsed_s
                w = int(st.best_weight)
                traj.append((int(st.generations), w))
                for m in marks:
                    if hit[m] is None and w <= planted * m:
                        hit[m] = elapsed
                if st.reached_target:
                    break
            ga.close()
            runs[name].append({"seed": hex(seed), "time_to": hit, "best_weight": int(st.best_weight),
                               "generations": int(st.generations), "trajectory": traj})
            print(f"ga seeds {name} {k}: {hit} best {st.best_weight} gens {st.generations}",
                  file=sys.stderr, flush=True)

    def summary(name, m):
        t = [r["time_to"][m] for r in runs[name]]
        done = sorted(x for x in t if x is not None)
        return {"reached": len(done), "of": len(t),
                "median_s": float(np.median(done)) if len(done) == len(t) else None,
                "min_s": done[0] if done else None, "max_s": done[-1] if done else None}

    p0 = _lib.ga_default_params(**overrides)
    out = {"planted_weight": planted, "isd_trials": int(p0.isd_trials),
           "window_rounds": int(p0.win_rounds), "budget_s": budget_s,
           "learned_overrides": learned_overrides, "marks": list(marks)}
    for m in marks:
        a, b = summary("plain", m), summary("learned", m)
        ratio = (b["median_s"] / a["median_s"]) if a["median_s"] and b["median_s"] else None
        out[f"x{m}"] = {"plain": a, "learned": b, "median_ratio_learned_over_plain": ratio}
    # weights at equal generations (the shorter run's last common point)
    eq = []
    for rp, rl in zip(runs["plain"], runs["learned"]):
        gp, gl = dict(rp["trajectory"]), dict(rl["trajectory"])
        common = sorted(set(gp) & set(gl))
        if common:
            g = common[-1]
            eq.append({"generations": g, "plain": gp[g], "learned": gl[g]})
    out["equal_generations"] = eq
    out["runs"] = {k: [{kk: vv for kk, vv in r.items() if kk != "trajectory"} for r in v]
                   for k, v in runs.items()}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--chunk-bits", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="gray",
                    choices=["gray", "coset", "batch", "window", "basis", "scores", "sa", "comb",
                             "ga", "search"],
                    help="gray: the Gray coset sweep (default); coset / batch / window: the "
                         "FWHT coset-minimum workloads; ga: plain vs learned GA")
    ap.add_argument("--ga-budget", type=float, default=60.0,
                    help="wall-clock seconds per GA variant in --workload ga")
    ap.add_argument("--ga-isd-trials", type=int, default=-1,
                    help="information-set trials per island per migration epoch in "
                         "--workload ga (both variants; 0 = off; default: the library's)")
    ap.add_argument("--ga-win-rounds", type=int, default=-1,
                    help="exact 20-row window rounds per migration epoch in --workload ga "
                         "(both variants; default: the library's)")
    ap.add_argument("--ga-seeds", type=int, default=0,
                    help="--workload ga: compare plain and learned over this many GA seeds "
                         "(median and range of the times to 40x, 20x and 1.01x planted)")
    ap.add_argument("--ga-win-candidates", type=int, default=0,
                    help="--workload ga: window candidates the learned proposal ranks per "
                         "target (default: the library's)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once (a plain "
                         "device copy of known size, to scale HBM byte counters)")
    ap.add_argument("--persistent-wgs", type=int, default=0,
                    help="coset / batch: fixed persistent workgroup count (default: autotuned)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the coset-minimum workloads reported beside the headline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        return 2

    import torch
    from bmmpy_amd import _lib

    if args.copy_calibration_mib > 0:
        import atexit

        def _copy_once():
            src = torch.empty(args.copy_calibration_mib << 18, dtype=torch.float32,
                              device="cuda:0").fill_(1.0)
            torch.cuda.synchronize()
            src.clone()
            torch.cuda.synchronize()

        atexit.register(_copy_once)

    ga_ov = {}
    if args.ga_isd_trials >= 0:
        ga_ov["isd_trials"] = args.ga_isd_trials
    if args.ga_win_rounds >= 0:
        ga_ov["win_rounds"] = args.ga_win_rounds
    if args.workload == "ga" and args.ga_seeds > 0:
        ov = {"win_candidates": args.ga_win_candidates} if args.ga_win_candidates else {}
        r = ga_seeds(args.ga_budget, args.ga_seeds, ga_ov, ov)
        print(json.dumps({"metric": METRIC, "value": r["x1.01"]["median_ratio_learned_over_plain"],
                          "unit": "median time ratio learned / plain to 1.01 x planted",
                          "n_gpus": 1, "steps": args.ga_seeds, "warmup": 0, "ms_per_step": None,
                          "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
                          "dtype": "u32", "data": "synthetic",
                          "config": {"workload": "ga_island_planted_256x8192_seeds",
                                     "seeds": args.ga_seeds, "isd_trials": r["isd_trials"],
                                     "window_rounds": r["window_rounds"]},
                          "ga": r}))
        return 0
    if args.workload == "ga":
        r = ga_compare(args.ga_budget, **ga_ov)
        plain, learned = r["plain"], r["learned"]
        ratio = None
        if plain["time_to_target_s"] and learned["time_to_target_s"]:
            ratio = learned["time_to_target_s"] / plain["time_to_target_s"]
        print(json.dumps({
            "metric": METRIC, "value": learned["best_weight"], "unit": "total weight reached",
            "n_gpus": 1, "steps": learned["generations"], "warmup": 0,
            "ms_per_step": learned["ms_per_generation"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": "ga_island_planted_256x8192", "islands": 64,
                       "population": 256, "max_ops": 512, "migrate_every": 20,
                       "window_rows": 20,
                       "window_rounds_per_epoch": r["window_rounds_per_epoch"],
                       "isd_trials_per_epoch": r["isd_trials_per_epoch"],
                       "init_rounds": r["init_rounds"],
                       "budget_s": args.ga_budget},
            "learned_over_plain_time_ratio": ratio, "ga": r}))
        return 0

    dev = torch.device("cuda:0")
    if args.workload == "search":
        r = run_search(torch, args.steps, args.warmup)
        print(json.dumps(mode_line(args, r, None)))
        return 0
    if args.workload in ("coset", "batch", "window", "basis", "scores", "sa", "comb"):
        stream = torch.cuda.current_stream(dev)
        if args.workload == "comb":
            r = run_comb(torch, args.steps, args.warmup)
            cpu = None if args.no_cpu_baseline else comb_cpu_baseline()
            line = mode_line(args, r, cpu)
            line["value"], line["unit"] = r["evals"] / (r["ms"] / 1e3), "masks/s"
            print(json.dumps(line))
            return 0
        if args.workload == "scores":
            r = run_scores(torch, args.steps, args.warmup)
            cpu = None if args.no_cpu_baseline else scores_cpu_baseline()
        elif args.workload == "sa":
            r = run_sa(torch, args.steps, args.warmup)
            cpu = None if args.no_cpu_baseline else sa_cpu_baseline()
            line = mode_line(args, r, cpu)
            line["value"], line["unit"] = r["evals"] / (r["ms"] / 1e3), "row pairs/s"
            print(json.dumps(line))
            return 0
        elif args.workload == "basis":
            r = run_basis(torch, args.steps, args.warmup)
            cpu = None if args.no_cpu_baseline else basis_cpu_baseline()
        elif args.workload == "window":
            r = run_window(torch, _lib, args.steps, args.warmup)
            cpu = None if args.no_cpu_baseline else window_cpu_baseline(_lib)
        else:
            r = run_batch_plan(torch, _lib, dev, stream, args.workload, args.steps, args.warmup,
                               args.persistent_wgs)
            cpu = None if args.no_cpu_baseline else batch_cpu_baseline(_lib, args.workload)
        print(json.dumps(mode_line(args, r, cpu)))
        return 0

    matrix = _lib.gen_bernoulli(SEED_GRAY_32X1024, ROWS, COLS, P)
    d_rows = torch.from_numpy(matrix.view(np.int64).copy()).to(dev)
    targets = np.arange(ROWS, dtype=np.int32)
    d_first = torch.empty(ROWS, dtype=torch.int64, device=dev)
    d_mask = torch.empty(ROWS, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def step():
        _lib.call("bmm_hip_gray_coset_min_dev", d_rows.data_ptr(), ROWS, COLS // 64, BLOCK,
                  targets, ROWS, args.chunk_bits, d_first.data_ptr(), d_mask.data_ptr(), sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    _lib.timing_begin()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    kernel_ms, launches = _lib.timing_end()
    ms = start.elapsed_time(end) / args.steps
    kernel_avg_ms = kernel_ms / max(launches, 1)

    # Results of the last pass (checked against the oracle by the tests).
    keys = d_first.cpu().numpy().view(np.uint64)
    evals = gray_sweep_evals()
    value = evals / (ms / 1e3)
    achieved_tops = evals * GRAY_OPS_PER_EVAL / (kernel_avg_ms / 1e3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "subset evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "gray_coset_sweep", "rows": ROWS, "cols": COLS, "p": P,
                   "pivot_block": BLOCK, "targets": ROWS, "seed": hex(SEED_GRAY_32X1024),
                   "subset_evals_per_step": evals},
        "roofline": {"bound": "valu", "achieved": achieved_tops, "peak": VALU_PEAK_TOPS,
                     "unit": "TOPS (int32 VALU)", "frac": achieved_tops / VALU_PEAK_TOPS,
                     "traffic": TRAFFIC_BYTES["gray_coset_sweep"],
                     "ops_per_eval": GRAY_OPS_PER_EVAL,
                     "kernel": "gray_coset_kernel<32,3>", "kernel_avg_ms": kernel_avg_ms,
                     "kernel_launches": launches},
        "result_min_weights": [int(k >> 32) for k in keys],
    }
    if not args.no_secondary:
        out["secondary"] = secondary(torch, _lib, dev, stream)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(matrix)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

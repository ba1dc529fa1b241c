"""GPU parity: Gray coset sweep, coset minimum (single and batched) and the
stand-alone FWHT, against the oracle on identical seeded inputs."""
import os
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED_COSET_20X128, SEED_GRAY_32X1024, SEED_BATCH_4096 = 0x5EED0001, 0x5EED0002, 0x5EED0004


def _say(*a):
    print(*a, flush=True)


def test_fwht_kats(lib):
    _say("fwht kats")
    assert lib.fwht_i32([1, 2, 3, 4]).tolist() == [10, -2, -4, 0]
    assert lib.fwht_i16([32767, 1]).tolist() == [-32768, 32766]


@pytest.mark.parametrize("n_bits", [1, 5, 12, 13, 14, 15, 18, 19, 21, 22, 24, 25, 26, 27, 28])
def test_fwht_i32_matches_oracle(lib, ora, n_bits):
    """Dense random int32 arrays (wrapping sums).  2^28 takes the two-pass
    register-tile path (16 + 12 index bits), the other sizes the LDS pass and
    register passes."""
    _say("fwht_i32", n_bits)
    rng = np.random.default_rng(n_bits)
    x = rng.integers(-2**31, 2**31, size=1 << n_bits, dtype=np.int64).astype(np.int32)
    assert np.array_equal(lib.fwht_i32(x), ora.fwht_i32(x))


@pytest.mark.parametrize("n_bits", [3, 11, 12, 14, 16, 19, 21, 24])
def test_fwht_i16_wraps_like_oracle(lib, ora, n_bits):
    _say("fwht_i16", n_bits)
    rng = np.random.default_rng(100 + n_bits)
    x = rng.integers(-2**15, 2**15, size=1 << n_bits, dtype=np.int64).astype(np.int16)
    assert np.array_equal(lib.fwht_i16(x), ora.fwht_i16(x))


@pytest.mark.parametrize("n_bits", [12, 20, 28])
def test_fwht_dev_unaligned_array(lib, ora, n_bits):
    """A device array 4 bytes off a 16-byte boundary takes the element-wise
    load path of the first pass (at 2^28 instead of the two-pass path, whose
    16-byte accesses need an aligned array)."""
    import torch
    _say("fwht unaligned", n_bits)
    rng = np.random.default_rng(50 + n_bits)
    x = rng.integers(-2**31, 2**31, size=1 << n_bits, dtype=np.int64).astype(np.int32)
    buf = torch.zeros((1 << n_bits) + 1, dtype=torch.int32, device="cuda:0")
    view = buf[1:]
    view.copy_(torch.from_numpy(x))
    torch.cuda.synchronize()
    assert view.data_ptr() % 16 != 0
    lib.call("bmm_hip_fwht_i32_dev", view.data_ptr(), x.size, None)
    torch.cuda.synchronize()
    assert np.array_equal(view.cpu().numpy(), ora.fwht_i32(x))


def test_gray_c2_full_config(lib, ora):
    """32x1024 Bernoulli(0.3); every target against rows 0..23 minus itself."""
    _say("gray c2 full")
    m = lib.gen_bernoulli(SEED_GRAY_32X1024, 32, 1024, 0.3)
    targets = np.arange(32, dtype=np.int32)
    w, first, low = lib.gray_coset_min(m, 24, targets)
    for t in range(32):
        ew, es, em = ora.gray_min(m, 24, t)
        assert (int(w[t]), int(first[t]), int(low[t])) == (ew, es, em), t


@pytest.mark.parametrize("rows,cols,p,block", [(20, 128, 0.5, 16), (18, 512, 0.2, 18),
                                               (40, 2048, 0.1, 24)])
def test_gray_other_widths(lib, ora, rows, cols, p, block):
    _say("gray", rows, cols, block)
    m = lib.gen_bernoulli(rows * 1000 + cols, rows, cols, p)
    targets = np.array([0, block - 1, rows - 1], dtype=np.int32)
    w, first, low = lib.gray_coset_min(m, block, targets)
    for i, t in enumerate(targets):
        assert (int(w[i]), int(first[i]), int(low[i])) == ora.gray_min(m, block, int(t))


def test_coset_c1_full_config(lib, ora):
    """20x128 Bernoulli(0.5): per-target min and argmin over the other 19 rows."""
    _say("coset c1")
    m = lib.gen_bernoulli(SEED_COSET_20X128, 20, 128, 0.5)
    w, idx = lib.coset_min(m, 128)
    ew, eidx = ora.coset_min(m, 128)
    assert np.array_equal(w, ew)
    assert np.array_equal(idx, eidx)


@pytest.mark.parametrize("rows,cols", [(2, 1), (3, 5), (8, 64), (13, 100), (14, 777),
                                       (16, 2048), (22, 300)])
def test_coset_edge_shapes(lib, ora, rows, cols):
    _say("coset", rows, cols)
    m = lib.gen_bernoulli(rows * 7 + cols, rows, cols, 0.4)
    w, idx = lib.coset_min(m, cols)
    ew, eidx = ora.coset_min(m, cols)
    assert np.array_equal(w, ew) and np.array_equal(idx, eidx)


def test_coset_zero_rows_and_duplicates(lib, ora):
    _say("coset degenerate")
    m = lib.gen_bernoulli(5, 10, 200, 0.3)
    m[3] = 0
    m[7] = m[2]
    w, idx = lib.coset_min(m, 200)
    ew, eidx = ora.coset_min(m, 200)
    assert np.array_equal(w, ew) and np.array_equal(idx, eidx)
    assert w[2] == 0 and w[7] == 0


def test_coset_batch_matches_per_matrix(lib, ora):
    """Batched uneven workload (the 4096-matrix generator), matrices with
    <= 20 rows checked one by one against the oracle."""
    _say("coset batch")
    rows, cols, dens, words, offs = lib.gen_batch(SEED_BATCH_4096, 256)
    w, idx = lib.coset_min_batch(words, offs, rows, cols)
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]])
    checked = 0
    for m in range(256):
        if rows[m] > 20:
            continue
        W = cols[m] // 64
        mat = words[offs[m]: offs[m] + rows[m] * W].reshape(rows[m], W)
        ew, eidx = ora.coset_min(mat, int(cols[m]))
        o = out_off[m]
        assert np.array_equal(w[o:o + rows[m]], ew), m
        assert np.array_equal(idx[o:o + rows[m]], eidx), m
        checked += 1
    assert checked > 100


def test_batch_plan_graph_and_autotune(lib, ora):
    """Plan with autotuned launch shape, replayed as a captured hipGraph three
    times; every replay equals the one-shot result and the oracle."""
    import ctypes as C
    import torch
    _say("batch plan graph")
    rows, cols, dens, words, offs = lib.gen_batch(0x5EED0004, 96)
    w1, i1 = lib.coset_min_batch(words, offs, rows, cols)
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    n = int(rows.sum())
    dev = torch.device("cuda:0")
    d_words = torch.from_numpy(words.view(np.int64).copy()).to(dev)
    d_w = torch.empty(n, dtype=torch.int32, device=dev)
    d_i = torch.empty(n, dtype=torch.int64, device=dev)
    plan = C.c_void_p()
    lib.call("bmm_hip_batch_plan_create", d_words.data_ptr(), offs, rows, cols, rows.size,
             out_off, d_w.data_ptr(), d_i.data_ptr(), 0, 1, C.byref(plan))
    try:
        info = lib.BatchInfo()
        lib.call("bmm_hip_batch_plan_info", plan, C.byref(info))
        assert info.n_matrices == 96 and info.n_slots == n
        assert info.subset_evals == float(sum(int(r) * 2 ** (int(r) - 1) for r in rows))
        for rep in range(3):
            d_w.fill_(-1)
            torch.cuda.synchronize()
            lib.call("bmm_hip_batch_plan_run", plan, 1, None)
            torch.cuda.synchronize()
            assert np.array_equal(d_w.cpu().numpy().view(np.uint32), w1), rep
            assert np.array_equal(d_i.cpu().numpy().view(np.uint64), i1), rep
    finally:
        lib.call("bmm_hip_batch_plan_destroy", plan)
    for m in range(0, 96, 7):
        if rows[m] > 18:
            continue
        W = cols[m] // 64
        mat = words[offs[m]: offs[m] + rows[m] * W].reshape(rows[m], W)
        ew, ei = ora.coset_min(mat, int(cols[m]))
        o = out_off[m]
        assert np.array_equal(w1[o:o + rows[m]], ew)
        assert np.array_equal(i1[o:o + rows[m]], ei)


def _plan_run(lib, words, offs, rows, cols, bits, graph):
    """Plan with a fixed block size; returns (weights, indices)."""
    import ctypes as C
    import torch
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    n = int(rows.sum())
    dev = torch.device("cuda:0")
    d_words = torch.from_numpy(words.view(np.int64).copy()).to(dev)
    d_w = torch.full((n,), -1, dtype=torch.int32, device=dev)
    d_i = torch.full((n,), -1, dtype=torch.int64, device=dev)
    plan = C.c_void_p()
    lib.call("bmm_hip_batch_plan_create", d_words.data_ptr(), offs, rows, cols, rows.size,
             out_off, d_w.data_ptr(), d_i.data_ptr(), 0, 0, C.byref(plan))
    try:
        lib.call("bmm_hip_batch_plan_set_shape", plan, bits, 0, 0)
        info = lib.BatchInfo()
        lib.call("bmm_hip_batch_plan_info", plan, C.byref(info))
        assert info.max_block_bits == bits
        for _ in range(2 if graph else 1):
            lib.call("bmm_hip_batch_plan_run", plan, int(graph), None)
        torch.cuda.synchronize()
        return d_w.cpu().numpy().view(np.uint32), d_i.cpu().numpy().view(np.uint64)
    finally:
        lib.call("bmm_hip_batch_plan_destroy", plan)


def _check_small(ora, words, offs, rows, cols, w, idx, max_rows):
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]])
    checked = 0
    for m in range(rows.size):
        if rows[m] > max_rows:
            continue
        W = (int(cols[m]) + 63) // 64
        mat = words[offs[m]: offs[m] + rows[m] * W].reshape(rows[m], W)
        ew, ei = ora.coset_min(mat, int(cols[m]))
        o = out_off[m]
        assert np.array_equal(w[o:o + rows[m]], ew), (m, rows[m], cols[m])
        assert np.array_equal(idx[o:o + rows[m]], ei), (m, rows[m], cols[m])
        checked += 1
    return checked


@pytest.mark.parametrize("bits", [12, 13, 14])
def test_batch_plan_block_sizes(lib, ora, bits):
    """Every block size (int32 2^12 / 2^13 blocks, packed-int16 2^14 blocks)
    against the oracle on the matrices of <= 20 rows, and against the 2^13
    result on all matrices (up to 28 rows), eager and graph replay."""
    _say("batch block size", bits)
    rows, cols, dens, words, offs = lib.gen_batch(SEED_BATCH_4096, 160)
    w, i = _plan_run(lib, words, offs, rows, cols, bits, graph=bits == 14)
    assert _check_small(ora, words, offs, rows, cols, w, i, 20) > 60
    w13, i13 = _plan_run(lib, words, offs, rows, cols, 13, graph=False)
    assert np.array_equal(w, w13) and np.array_equal(i, i13)


@pytest.mark.parametrize("bits", [13, 14])
def test_batch_wide_matrices_read_patterns_from_memory(lib, ora, bits):
    """Matrices wider than 2048 columns (column patterns read from memory
    instead of the per-thread cache), including a tail word and a matrix
    exactly at the block size."""
    _say("batch wide", bits)
    shapes = [(16, 4096, 0.3), (15, 2561, 0.5), (14, 3000, 0.1), (17, 2049, 0.45), (9, 5000, 0.2)]
    mats = [lib.gen_bernoulli(0x5EED0100 + k, r, c, p) for k, (r, c, p) in enumerate(shapes)]
    rows = np.array([s[0] for s in shapes], np.int32)
    cols = np.array([s[1] for s in shapes], np.int32)
    offs = np.concatenate([[0], np.cumsum([m.size for m in mats])[:-1]]).astype(np.int64)
    words = np.concatenate([m.reshape(-1) for m in mats])
    w, i = _plan_run(lib, words, offs, rows, cols, bits, graph=False)
    assert _check_small(ora, words, offs, rows, cols, w, i, 17) == len(shapes)


def test_batch_int16_blocks_refuse_wide_matrices(lib):
    import ctypes as C
    import torch
    rows = np.array([14], np.int32)
    cols = np.array([32768], np.int32)
    words = lib.gen_bernoulli(7, 14, 32768, 0.01).reshape(-1)
    d_words = torch.from_numpy(words.view(np.int64).copy()).to("cuda:0")
    d_w = torch.empty(14, dtype=torch.int32, device="cuda:0")
    d_i = torch.empty(14, dtype=torch.int64, device="cuda:0")
    plan = C.c_void_p()
    lib.call("bmm_hip_batch_plan_create", d_words.data_ptr(), np.array([0], np.int64), rows,
             cols, 1, np.array([0], np.int64), d_w.data_ptr(), d_i.data_ptr(), 0, 0,
             C.byref(plan))
    try:
        with pytest.raises(ValueError, match="32767"):
            lib.call("bmm_hip_batch_plan_set_shape", plan, 14, 0, 0)
        with pytest.raises(ValueError):
            lib.call("bmm_hip_batch_plan_set_shape", plan, 11, 0, 0)
    finally:
        lib.call("bmm_hip_batch_plan_destroy", plan)


@pytest.mark.parametrize("n_bits", [29, 30])
def test_fwht_large_sparse_closed_form(lib, n_bits):
    """2 GiB (buffer-load passes) and 4 GiB (global-pointer passes) int32
    arrays: a few spikes x_p = v transform to H[m] = sum_p v (-1)^|p & m|,
    checked on 65536 random indices."""
    _say("fwht large", n_bits)
    rng = np.random.default_rng(900 + n_bits)
    n = 1 << n_bits
    pos = rng.choice(n, 5, replace=False).astype(np.int64)
    val = rng.integers(-1000, 1000, 5).astype(np.int64)
    x = np.zeros(n, np.int32)
    x[pos] = val
    h = lib.fwht_i32(x)
    del x
    m = rng.integers(0, n, 65536, dtype=np.int64)
    exp = np.zeros(m.size, np.int64)
    for p, v in zip(pos, val):
        par = np.bitwise_count((m & p).astype(np.uint64)) & 1
        exp += np.where(par == 1, -v, v)
    assert np.array_equal(h[m].astype(np.int64), exp)
    assert int(h[0]) == int(val.sum())


@pytest.mark.parametrize("bits,graph", [(13, False), (14, True), (0, True)])
def test_batch_4096_large_matrices_vs_golden(lib, bits, graph):
    """The whole 4096-matrix workload in one plan (block size 2^13 int32,
    2^14 packed int16, or the autotuned shape), compared on the first matrix
    of each row count 21 .. 28 (where nearly all the transform work is) with
    golden minima and indices computed by the oracle's coset minimum
    (tests/golden/make_golden_batch_large.py)."""
    import ctypes as C
    import torch
    _say("batch 4096 large matrices vs golden", bits, graph)
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "golden_batch_large.npz"))
    rows, cols, dens, words, offs = lib.gen_batch(SEED_BATCH_4096, 4096)
    assert np.array_equal(rows[g["matrices"]], g["rows"])
    assert np.array_equal(cols[g["matrices"]], g["cols"])
    if bits:
        w, i = _plan_run(lib, words, offs, rows, cols, bits, graph=graph)
    else:
        out_off = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
        n = int(rows.sum())
        dev = torch.device("cuda:0")
        d_words = torch.from_numpy(words.view(np.int64).copy()).to(dev)
        d_w = torch.full((n,), -1, dtype=torch.int32, device=dev)
        d_i = torch.full((n,), -1, dtype=torch.int64, device=dev)
        plan = C.c_void_p()
        lib.call("bmm_hip_batch_plan_create", d_words.data_ptr(), offs, rows, cols, rows.size,
                 out_off, d_w.data_ptr(), d_i.data_ptr(), 0, 1, C.byref(plan))
        try:
            lib.call("bmm_hip_batch_plan_run", plan, 1, None)
            lib.call("bmm_hip_batch_plan_run", plan, 1, None)
            torch.cuda.synchronize()
            w, i = d_w.cpu().numpy().view(np.uint32), d_i.cpu().numpy().view(np.uint64)
        finally:
            lib.call("bmm_hip_batch_plan_destroy", plan)
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]])
    k = 0
    for m, r in zip(g["matrices"], g["rows"]):
        o = out_off[m]
        assert np.array_equal(w[o:o + r], g["weights"][k:k + r]), (m, r)
        assert np.array_equal(i[o:o + r], g["indices"][k:k + r]), (m, r)
        k += r


def test_small_batch_single_launch_vs_oracle_and_staged(lib, ora):
    """Small batches run as one launch (every workgroup one block, exact
    per-target lowest mask, the last workgroup writes the outputs): against
    the oracle, against the staged three-launch plan (set_shape), and over
    repeated graph replays (the kernel clears its own state)."""
    import ctypes as C
    import torch
    _say("small batch single launch")
    rng = np.random.default_rng(2024)
    shapes = [(2, 64), (3, 1), (8, 100), (13, 2048), (14, 700), (17, 130), (20, 128),
              (21, 64), (9, 0), (12, 1999)]
    rows = np.array([r for r, _ in shapes], np.int32)
    cols = np.array([c for _, c in shapes], np.int32)
    mats, offs, o = [], [], 0
    for r, c in shapes:
        W = (c + 63) // 64
        m = rng.integers(0, 2 ** 63, size=(r, max(W, 1)), dtype=np.uint64)[:, :W]
        if c % 64:
            m[:, -1] &= np.uint64((1 << (c % 64)) - 1)
        if r == 17:
            m[5] = m[2] ^ m[9]          # a zero-weight coset
        mats.append(np.ascontiguousarray(m))
        offs.append(o)
        o += r * W
    words = np.concatenate([m.reshape(-1) for m in mats]).astype(np.uint64)
    offs = np.array(offs, np.int64)
    out_off = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    n = int(rows.sum())
    dev = torch.device("cuda:0")
    d_words = torch.from_numpy(words.view(np.int64).copy()).to(dev)
    d_w = torch.full((n,), -1, dtype=torch.int32, device=dev)
    d_i = torch.full((n,), -1, dtype=torch.int64, device=dev)
    plan = C.c_void_p()
    lib.call("bmm_hip_batch_plan_create", d_words.data_ptr(), offs, rows, cols, rows.size,
             out_off, d_w.data_ptr(), d_i.data_ptr(), 0, 1, C.byref(plan))
    try:
        info = lib.BatchInfo()
        lib.call("bmm_hip_batch_plan_info", plan, C.byref(info))
        assert info.persistent_wgs == info.total_blocks <= 1024
        results = []
        for rep in range(4):
            d_w.fill_(-1)
            torch.cuda.synchronize()
            lib.call("bmm_hip_batch_plan_run", plan, int(rep > 0), None)
            torch.cuda.synchronize()
            results.append((d_w.cpu().numpy().view(np.uint32).copy(),
                            d_i.cpu().numpy().view(np.uint64).copy()))
        lib.call("bmm_hip_batch_plan_set_shape", plan, 13, 0, 0)
        lib.call("bmm_hip_batch_plan_run", plan, 0, None)
        torch.cuda.synchronize()
        staged = (d_w.cpu().numpy().view(np.uint32).copy(),
                  d_i.cpu().numpy().view(np.uint64).copy())
    finally:
        lib.call("bmm_hip_batch_plan_destroy", plan)
    for w, i in results[1:] + [staged]:
        assert np.array_equal(w, results[0][0]) and np.array_equal(i, results[0][1])
    w, i = results[0]
    for k, (r, c) in enumerate(shapes):
        ew, ei = ora.coset_min(mats[k].reshape(r, -1) if c else np.zeros((r, 1), np.uint64),
                               c)
        o = out_off[k]
        assert np.array_equal(w[o:o + r], ew), (r, c)
        assert np.array_equal(i[o:o + r], ei), (r, c)
    assert w[out_off[5] + 5] == 0

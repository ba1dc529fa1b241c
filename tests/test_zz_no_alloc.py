"""The allocation-free Gray coset sweep (bmm_hip_gray_coset_min_noalloc)
against the allocating device entry point (bmm_hip_gray_coset_min_dev) on
two calls in a row, the second smaller, through the C ABI, with the keys
checked against the oracle too; and its limits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _keys(lib, torch, name, rows, block, targets, chunk=0):
    dev = torch.device("cuda:0")
    d_rows = torch.from_numpy(rows.view(np.int64).copy()).to(dev)
    t = np.ascontiguousarray(targets, np.int32)
    k1 = torch.zeros(t.size, dtype=torch.int64, device=dev)
    k2 = torch.zeros(t.size, dtype=torch.int64, device=dev)
    lib.call(name, d_rows.data_ptr(), rows.shape[0], rows.shape[1], block, t, t.size, chunk,
             k1.data_ptr(), k2.data_ptr(), None)
    torch.cuda.synchronize()
    return k1.cpu().numpy().view(np.uint64), k2.cpu().numpy().view(np.uint64)


def test_noalloc_gray_equals_allocating_entry(lib, ora):
    import torch
    print("noalloc gray: 32 x 1024 block 24 all targets, then 20 x 320 block 18", flush=True)
    big = lib.gen_bernoulli(0x5EED0002, 32, 1024, 0.3)
    small = lib.gen_bernoulli(77, 20, 320, 0.3)      # 5 words: padded to 8 in scratch
    cases = [(big, 24, np.arange(32)), (small, 18, np.array([0, 3, 17, 18, 19]))]
    for rows, block, targets in cases:
        a = _keys(lib, torch, "bmm_hip_gray_coset_min_noalloc", rows, block, targets)
        b = _keys(lib, torch, "bmm_hip_gray_coset_min_dev", rows, block, targets)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        for i in (0, len(targets) - 1):
            w, first, low = ora.gray_min(rows, block, int(targets[i]))
            assert (int(a[0][i]) >> 32, int(a[0][i]) & 0xffffffff) == (w, first)
            assert int(a[1][i]) & 0xffffffff == low
    # the small call after the big one must not see the big call's tables
    again = _keys(lib, torch, "bmm_hip_gray_coset_min_noalloc", small, 18, [19, 0])
    ref = _keys(lib, torch, "bmm_hip_gray_coset_min_dev", small, 18, [19, 0])
    assert np.array_equal(again[0], ref[0]) and np.array_equal(again[1], ref[1])


def test_noalloc_gray_limits(lib):
    import torch
    rows = lib.gen_bernoulli(5, 80, 128, 0.3)
    with pytest.raises(ValueError, match="at most 64 targets"):
        _keys(lib, torch, "bmm_hip_gray_coset_min_noalloc", rows, 16, np.arange(65))
    with pytest.raises(ValueError, match="61440 workgroups"):
        _keys(lib, torch, "bmm_hip_gray_coset_min_noalloc", rows, 40, np.arange(48, 52), chunk=4)
    with pytest.raises(IndexError):
        _keys(lib, torch, "bmm_hip_gray_coset_min_noalloc", rows, 16, [80])

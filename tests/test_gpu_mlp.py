"""GPU proposal MLP (MFMA bf16 products, fp32 accumulation) against the
float64 restatement with the same bf16 rounding points.  Tolerances:
forward max |diff| <= 1e-3 * (1 + |y|); gradients relative L2 error <= 2e-3
(fp32 accumulation order; ReLU masks at exactly zero may flip).  Batches of
1024 and more take the weight-gradient products cut along the batch into
K slices (2080: slices of unequal length)."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def model(lib):
    m = lib.GpuMlp(seed=11)
    yield m
    m.close()


def test_param_roundtrip(lib, model):
    p = mlp.init_params(5).astype(np.float32)
    model.set_params(p)
    assert np.array_equal(model.get_params(), p)


@pytest.mark.parametrize("batch", [1, 31, 32, 200])
def test_forward_matches_oracle(lib, model, batch):
    _say("mlp forward", batch)
    rng = np.random.default_rng(batch)
    p = mlp.init_params(7).astype(np.float32)
    model.set_params(p)
    X = rng.choice([-1.0, 1.0], size=(batch, 256)).astype(np.float32)
    y = model.forward(X)
    ref, _ = mlp.forward(p.astype(np.float64), X, bf16=True)
    assert np.all(np.abs(y - ref) <= 1e-3 * (1 + np.abs(ref))), np.max(np.abs(y - ref))


@pytest.mark.parametrize("batch", [32, 96, 256, 1024, 2080, 4096])
def test_gradients_match_oracle(lib, model, batch):
    _say("mlp grads", batch)
    rng = np.random.default_rng(100 + batch)
    p = mlp.init_params(9).astype(np.float32)
    model.set_params(p)
    X = rng.choice([-1.0, 1.0], size=(batch, 256)).astype(np.float32)
    t = rng.normal(size=batch).astype(np.float32)
    loss, g = model.gradients(X, t)
    rl, rg = mlp.loss_and_grads(p.astype(np.float64), X, t, bf16=True)
    assert abs(loss - rl) <= 1e-3 * (1 + abs(rl))
    names = [n for n, _ in mlp.SIZES]
    G, R = mlp.unpack(g), mlp.unpack(rg)
    for n in names:
        err = np.linalg.norm(G[n] - R[n]) / (np.linalg.norm(R[n]) + 1e-12)
        assert err <= 2e-3, (n, err)


def test_training_fits_pair_deltas(lib):
    """Online training on exact row-pair deltas: the loss falls, and the pairs
    the trained network ranks best have a lower mean exact delta than random
    pairs (the gain of the learned proposal over uniform proposals)."""
    _say("mlp training gain")
    rng = np.random.default_rng(3)
    S, A, M = lib.gen_planted(0x5EED0005, 128, 2048, 0.02)
    rows = A
    words = rows.shape[1]
    pop = np.array([[bin(int(v)).count("1") for v in r] for r in rows])   # per-word popcounts

    def sketch(r):
        per = pop[r]                     # 32 words -> 32 buckets; pad to 128
        tot = per.sum()
        bits = (per * 32 > tot).astype(np.float32) * 2 - 1
        return np.concatenate([bits, -np.ones(128 - words, np.float32)])

    sk = np.stack([sketch(r) for r in range(rows.shape[0])])
    wt = pop.sum(1)

    def delta(s, d):
        return sum(bin(int(a ^ b)).count("1") for a, b in zip(rows[s], rows[d])) - wt[d]

    def batch(n):
        s = rng.integers(0, 128, n)
        d = (s + rng.integers(1, 128, n)) % 128
        X = np.concatenate([sk[s], sk[d]], 1)
        t = np.array([delta(a, b) for a, b in zip(s, d)], np.float32) / 64.0
        return X, t

    m = lib.GpuMlp(seed=21)
    Xv, tv = batch(2048)
    l0 = float(np.mean((m.forward(Xv) - tv) ** 2))
    for it in range(300):
        X, t = batch(256)
        m.train_step(X, t, lr=1e-3)
    l1 = float(np.mean((m.forward(Xv) - tv) ** 2))
    assert l1 < 0.5 * l0, (l0, l1)
    pred = m.forward(Xv)
    top = np.argsort(pred)[:128]
    assert tv[top].mean() < tv.mean() - 0.5 * tv.std(), (tv[top].mean(), tv.mean(), tv.std())
    m.close()


@pytest.mark.parametrize("n", [1, 63, 1000])
def test_fused_pair_scoring_matches_forward(lib, model, n):
    """The GA's fused scorer (features from 128-bit row sketches built in LDS,
    both layers on MFMA in one kernel) against the float64 oracle on the
    same +-1 features and against the generic forward path."""
    _say("fused scoring", n)
    p = mlp.init_params(21).astype(np.float32)
    model.set_params(p)
    rng = np.random.default_rng(n)
    rows = 40
    sketch = rng.integers(0, 2**63, size=(rows, 2), dtype=np.uint64) ^ \
        (rng.integers(0, 2, size=(rows, 2), dtype=np.uint64) << np.uint64(63))
    src = rng.integers(0, rows, size=n)
    dst = rng.integers(0, rows, size=n)
    pairs = (src | (dst << 8)).astype(np.uint16)
    bits = np.unpackbits(sketch.view(np.uint8), bitorder="little").reshape(rows, 128)
    x = np.concatenate([bits[src], bits[dst]], axis=1).astype(np.float32) * 2 - 1
    y = model.score_pairs(sketch, pairs)
    ref, _ = mlp.forward(p.astype(np.float64), x.astype(np.float64), bf16=True)
    assert np.all(np.abs(y - ref) <= 1e-3 * (1 + np.abs(ref)))
    assert np.allclose(y, model.forward(x), rtol=1e-4, atol=1e-4)

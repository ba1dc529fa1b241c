"""GPU island GA operators checked against restatements and invariants:
one breeding step at a time against oracle/ga_breed.py (ranked elitism,
tournaments, crossover with num_parents / num_offspring, both mutation forms,
catastrophe and restarts), every individual's fitness against an oracle
replay of its op sequence, ring migration with export / import flags, a
per-island best that never worsens, and the exported best op sequence
replayed by the oracle."""
import numpy as np
import pytest

from oracle.ga_breed import breed, params_of

pytestmark = pytest.mark.gpu

COLS = 2048


def _say(*a):
    print(*a, flush=True)


def _weight(m):
    return int(np.unpackbits(np.ascontiguousarray(m).view(np.uint8)).sum())


def _seqs(ops, lens):
    return [[int(x) for x in ops[i, :lens[i]]] for i in range(len(lens))]


def _pairs(seq):
    return np.array([[o & 0xff, o >> 8] for o in seq], np.uint16).reshape(-1, 2)


CONFIGS = {
    # GPU mutation draws (append / replace / delete), two parents
    "gpu_mutations": dict(mutations=3, elite=2),
    # reference mutation with rate adaptation, three parents, two offspring
    # per parent set, restarts after one stale generation
    "reference_ops": dict(mutations=0, elite=3, mutation_rate=0.2, num_parents=3,
                          num_offspring=2, restart_after=1, restart_p=0.3),
    # catastrophe every 2 stale generations, one parent (copy + mutation)
    "catastrophe": dict(mutations=0, elite=1, mutation_rate=0.1, num_parents=1,
                        catastrophe_threshold=2, catastrophe_survival=0.25),
    # learned proposal: ranked pairs in mutations and in the crossover cuts
    "learned": dict(mutations=3, elite=2, learned=1, n_candidates=64, train_steps=2,
                    p_learned=0.7, num_parents=3),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_breeding_matches_restatement(lib, ora, name):
    """Generation by generation on a sparse 32 x 2048 matrix (random moves
    rarely help there, so stale counters, restarts and catastrophes occur):
    the evaluated fitness of every individual equals the oracle's replay of
    its ops, and the bred population, random streams and counters equal the
    restated breeding step."""
    _say("ga breeding restatement", name)
    S, A, M = lib.gen_planted(0x5EED0005, 32, COLS, 0.02)
    kw = dict(islands=2, population=48, max_ops=40, tournament=3, migrate_every=1000,
              check_every=1, learned=0, li_sweeps=0, win_rows=0, isd_trials=0)
    if name == "learned":
        S, _, _ = lib.gen_planted(0x5EED0006, 32, COLS, 0.05)   # room for improving moves
    kw.update(CONFIGS[name])
    p = lib.ga_default_params(**kw)
    ga = lib.GpuGA(S, COLS, p)
    try:
        ga.configure_islands(seeds=[11, 12])
        P = params_of(p, 32)
        modes, restarts, ops_seen = set(), 0, 0
        for gen in range(12):
            pre = [ga.population(i, 48, 40) for i in range(2)]
            st_run = ga.run(1)
            # every evaluated individual's ops are counted once
            ops_seen += sum(int(p["len_next"].sum()) for p in pre)
            assert st_run.replayed_ops == ops_seen
            for isl in range(2):
                a, b = pre[isl], ga.population(isl, 48, 40)
                assert np.array_equal(b["len_prev"], a["len_next"])
                seqs = _seqs(a["ops_next"], a["len_next"])
                base, _ = ga.best_ops(isl)
                for i in range(0, 48, 7):
                    _, w = ora.replay(base, COLS, _pairs(seqs[i]))
                    assert int(b["fit_prev"][i]) == w, (gen, isl, i)
                st0 = {k: a[k] for k in ("best_weight", "best_len", "no_improvement",
                                         "catastrophe", "mutation_rate")}
                cand = ga.proposals(isl)
                assert cand.size == (64 if name == "learned" else 0)
                kids, rng, st, _, mode = breed(P, seqs, b["fit_prev"], a["rng"], st0, cand)
                modes.add(mode)
                got = _seqs(b["ops_next"], b["len_next"])
                assert got == kids, (name, gen, isl)
                assert [int(x) for x in b["rng"]] == rng
                for k in ("best_weight", "no_improvement", "catastrophe"):
                    assert int(b[k]) == int(st[k]), (k, gen, isl)
                assert np.float32(b["mutation_rate"]) == np.float32(st["mutation_rate"])
                if name == "reference_ops" and st0["no_improvement"] >= 1:
                    restarts += 1
        if name == "catastrophe":
            assert 1 in modes            # a catastrophe happened
        if name == "reference_ops":
            assert restarts > 0          # generations past the restart limit were bred
    finally:
        ga.close()


def _migration_ga(lib, A, flags=None):
    p = lib.ga_default_params(islands=3, population=8, max_ops=8, migrate_every=1,
                              check_every=1, mutations=0, learned=0, li_sweeps=0, win_rows=0,
                              isd_trials=0)
    ga = lib.GpuGA(A, COLS, p)
    if flags is not None:
        ga.configure_islands(migrate=flags)
    return ga


def test_ring_migration_adopts_only_a_lighter_predecessor(lib):
    """Island 0 holds the light planted rows, islands 1 and 2 the mixed ones:
    after one epoch island 1 takes island 0's base (lighter predecessor),
    island 2 keeps its own (its predecessor was as heavy before the epoch),
    island 0 keeps its own (island 2 is heavier)."""
    _say("ga ring migration")
    S, A, M = lib.gen_planted(0x5EED0006, 32, COLS, 0.02)
    ga = _migration_ga(lib, A)
    try:
        ga.set_island_base(0, S)
        ga.run(1)
        b = [ga.island_best(i) for i in range(3)]
        assert np.array_equal(b[0][0], S) and b[0][1] == _weight(S)
        assert np.array_equal(b[1][0], S) and b[1][1] == _weight(S)
        assert np.array_equal(b[2][0], A) and b[2][1] == _weight(A)
        # next epoch: island 2 now sees a lighter predecessor
        ga.run(1)
        assert np.array_equal(ga.island_best(2)[0], S)
    finally:
        ga.close()


@pytest.mark.parametrize("flags", [[2, 3, 3], [3, 1, 3]])
def test_ring_migration_respects_export_and_import(lib, flags):
    """Island 0 not exporting, or island 1 not importing: island 1 keeps A."""
    _say("ga migration flags", flags)
    S, A, M = lib.gen_planted(0x5EED0006, 32, COLS, 0.02)
    ga = _migration_ga(lib, A, flags)
    try:
        ga.set_island_base(0, S)
        ga.run(1)
        assert np.array_equal(ga.island_best(1)[0], A)
        assert ga.population(1, 8, 8)["migrate"] == flags[1]
    finally:
        ga.close()


@pytest.mark.parametrize("learned", [0, 1])
def test_island_best_never_worsens_and_replays(lib, ora, learned):
    """Across generations and epochs (fold, migration, pairwise, window and
    information-set improvement) no island's best gets heavier, and each
    island's exported (base, best ops) replayed by the oracle gives the
    reported best matrix and weight, which keeps the row space."""
    _say("ga best monotone", learned)
    S, A, M = lib.gen_planted(0x5EED0005, 64, COLS, 0.02)
    p = lib.ga_default_params(islands=4, population=32, max_ops=32, migrate_every=4,
                              check_every=1, learned=learned, n_candidates=32, train_steps=2,
                              win_rows=16, win_rounds=1, win_candidates=32)
    ga = lib.GpuGA(A, COLS, p)
    try:
        last = [_weight(A)] * 4
        for step in range(10):
            ga.run(1 if step % 3 else 3)
            for i in range(4):
                w = ga.population(i, 32, 32)["best_weight"]
                assert w <= last[i], (step, i)
                last[i] = w
                base, ops = ga.best_ops(i)
                rows, wr = ora.replay(base, COLS, ops)
                best, wb = ga.island_best(i)
                assert np.array_equal(rows, best) and wr == wb == w
        assert min(last) < _weight(A)
        assert ora.rank(np.concatenate([A, best]), COLS) == ora.rank(A, COLS) == 64
    finally:
        ga.close()


def test_island_configuration_errors(lib):
    S, A, M = lib.gen_planted(0x5EED0006, 16, COLS, 0.02)
    ga = _migration_ga(lib, A)
    try:
        with pytest.raises(ValueError):
            ga.configure_islands(seeds=[1, 2])          # three islands
        with pytest.raises(ValueError):
            ga.configure_islands(migrate=[0, 4, 1])
        with pytest.raises(IndexError):
            ga.best_ops(3)
        with pytest.raises(IndexError):
            ga.population(-1, 8, 8)
    finally:
        ga.close()
    with pytest.raises(ValueError):
        lib.GpuGA(A, COLS, lib.ga_default_params(num_parents=5))
    with pytest.raises(ValueError):
        lib.GpuGA(A, COLS, lib.ga_default_params(num_offspring=0))


def test_fitness_exact_in_both_eval_tiers(lib, ora):
    """Fitness of every individual equals the oracle's replay of its ops on a
    128-row window whose op lists grow from a few ops to well over 64
    touched rows: individuals touching at most 64 rows are evaluated by the
    small-LDS tier, the others by the large tier; both occur."""
    _say("ga fitness, both evaluation tiers")
    S, A, M = lib.gen_planted(0x5EED000A, 128, COLS, 0.03)
    p = lib.ga_default_params(islands=2, population=32, max_ops=200, tournament=3, elite=2,
                              migrate_every=1000, check_every=1, learned=0, li_sweeps=0,
                              win_rows=0, isd_trials=0, mutations=12, p_append=1.0,
                              p_replace=0.0, p_delete=0.0)
    ga = lib.GpuGA(A, COLS, p)
    small = large = 0
    try:
        for gen in range(10):
            pre = [ga.population(i, 32, 200) for i in range(2)]
            ga.run(1)
            for isl in range(2):
                a, b = pre[isl], ga.population(isl, 32, 200)
                seqs = _seqs(a["ops_next"], a["len_next"])
                base, _ = ga.best_ops(isl)
                for i in range(32):
                    pr = _pairs(seqs[i])
                    moved = pr[pr[:, 0] != pr[:, 1]] if len(pr) else pr
                    touched = len(set(moved.reshape(-1).tolist()))
                    if touched <= 64:
                        small += 1
                    else:
                        large += 1
                    _, w = ora.replay(base, COLS, pr)
                    assert int(b["fit_prev"][i]) == w, (gen, isl, i, touched)
    finally:
        ga.close()
    print(f"  evaluations: {small} small-tier, {large} large-tier", flush=True)
    assert small > 0 and large > 0

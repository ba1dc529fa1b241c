"""One breeding step of the GPU island GA, restated on the CPU (test
infrastructure: a checker used only by tests/).

The GPU GA keeps the reference's GeneticAlgorithm::step()
(src/ga/genetic_algorithm_lifecycle.cpp:112-200) with op-sequence individuals
(a child is a list of row XORs (src -> dst) over the island's base matrix).
Per island, from the population the last generation evaluated and its exact
fitness:

1. rank by (fitness, index); a best lighter than the island's best is kept and
   resets the stale counters (lifecycle.cpp:176-186), else they grow;
2. catastrophe when its counter reaches the threshold
   (genetic_algorithm_operators.cpp:15-37): the floor(pop * survival) best
   survive, the others are random individuals;
3. else the mutation rate adapts (genetic_algorithm_core.cpp:33-41), the
   elite_count best are copied (lifecycle.cpp:133-134), and the others come
   in groups of num_offspring that share num_parents tournament winners
   (operators.cpp:39-56) and their cut points; after restart_after stale
   generations a child is a random individual with probability restart_p
   (lifecycle.cpp:118-142);
4. each child is mutated on its own random stream: with mutation_rate > 0,
   max(1, floor(rows * rate)) row XORs r_i ^= r_j with i, j uniform and
   i == j skipped (operators.cpp:100-129); otherwise `mutations` draws of
   append / replace / delete of one op.

With the learned proposal, mutations take one of the island's ranked pairs
with probability p_learned, and a later parent's segment starts at the best
of 4 uniform cut draws by the number of its next 8 ops that are ranked pairs
(first best on ties).

A random individual is the identity followed by 3 * rows random row XORs
(genetic_algorithm_population.cpp:20-36), at most max_ops of them.

Random numbers: one XorShift64 stream per child (13, 7, 17 shifts), a unit
draw is (x >> 40) * 2^-24 in float32, an index draw is x mod n.  Float
arithmetic follows the kernel's float32 rounding.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
F = np.float32


class _Rng:
    def __init__(self, s: int) -> None:
        self.s = int(s)

    def next(self) -> int:
        s = self.s
        s ^= (s << 13) & M64
        s ^= s >> 7
        s ^= (s << 17) & M64
        self.s = s
        return s

    def unit(self) -> np.float32:
        return F(self.next() >> 40) * F(1.0 / 16777216.0)

    def index(self, n: int) -> int:
        return self.next() % n


def _op(src: int, dst: int) -> int:
    return src | (dst << 8)


def random_sequence(rows: int, max_ops: int, rng: _Rng) -> list[int]:
    out = []
    for _ in range(3 * rows):
        if len(out) >= max_ops:
            break
        i = rng.index(rows)
        j = rng.index(rows)
        if i != j:
            out.append(_op(j, i))
    return out


def _tournament(size: int, pop: int, rng: _Rng, fit) -> int:
    best = rng.index(pop)
    for _ in range(1, size):
        c = rng.index(pop)
        if fit[c] < fit[best]:
            best = c
    return best


def _propose(p: dict, rng: _Rng, cand) -> int:
    if p["learned"] and p["n_cand"] > 0 and rng.unit() < F(p["p_learned"]):
        return int(cand[rng.index(p["n_cand"])])
    r = p["rows"]
    s = rng.index(r)
    d = rng.index(r - 1)
    if d >= s:
        d += 1
    return _op(s, d)


def breed(p: dict, seqs: list[list[int]], fit, rng_states, state: dict, cand=None):
    """One breeding step of one island.

    p: rows, pop, max_ops, tournament, elite, mutations, p_append,
    p_replace, p_delete, learned, n_cand, p_learned, num_parents,
    num_offspring, mutation_rate, catastrophe_threshold, catastrophe_survival,
    restart_after, restart_p.  seqs: the evaluated population (op lists),
    fit: its fitness, rng_states: one stream state per child, state:
    best_weight, best_len, no_improvement, catastrophe, mutation_rate.
    cand: the island's ranked pairs (learned proposal).  Returns (children,
    new stream states, new state, best index or None, 1 for a catastrophe
    step else 0)."""
    pop = p["pop"]
    order = sorted(range(pop), key=lambda i: (int(fit[i]), i))
    arg = order[0]
    st = dict(state)
    best_idx = None
    if int(fit[arg]) < st["best_weight"]:
        st["best_weight"] = int(fit[arg])
        st["best_len"] = len(seqs[arg])
        st["no_improvement"] = 0
        st["catastrophe"] = 0
        best_idx = arg
    else:
        st["no_improvement"] += 1
        st["catastrophe"] += 1
    mode = 0
    if p["catastrophe_threshold"] > 0 and st["catastrophe"] == p["catastrophe_threshold"]:
        mode = 1
        st["catastrophe"] = 0
    rate = F(st["mutation_rate"])
    if p["mutation_rate"] > 0:
        if st["no_improvement"] > 10:
            rate = min(F(0.8), F(rate * F(1.5)))
        elif st["no_improvement"] > 5:
            rate = min(F(0.6), F(rate * F(1.3)))
        else:
            rate = max(F(0.3), F(rate * F(0.95)))
    st["mutation_rate"] = rate
    rngs = [_Rng(s) for s in rng_states]
    kinds = [0] * pop
    for t in range(pop):
        if mode == 1:
            survivors = int(F(pop) * F(p["catastrophe_survival"]))
            kinds[t] = 0 if t < survivors else 1
        elif t < p["elite"]:
            kinds[t] = 0
        elif (p["restart_after"] > 0 and st["no_improvement"] > p["restart_after"]
              and rngs[t].unit() < F(p["restart_p"])):
            kinds[t] = 1
        else:
            kinds[t] = 2
    # Group leaders draw parents and cut points.
    good = set(int(x) for x in cand) if (p["learned"] and p["n_cand"] > 0) else None
    groups = {}
    for t in range(pop):
        c = t - p["elite"]
        if mode == 0 and c >= 0 and c % p["num_offspring"] == 0:
            sel = []
            for q in range(p["num_parents"]):
                par = _tournament(p["tournament"], pop, rngs[t], fit)
                lp = len(seqs[par])
                s0, e0 = 0, lp
                if p["num_parents"] == 1:
                    pass
                elif q == 0:
                    e0 = rngs[t].index(lp + 1)
                else:
                    if good is not None:
                        best_sc = -1
                        for _ in range(4):
                            cs = rngs[t].index(lp + 1)
                            sc = sum(1 for k in range(cs, min(cs + 8, lp)) if seqs[par][k] in good)
                            if sc > best_sc:
                                best_sc, s0 = sc, cs
                    else:
                        s0 = rngs[t].index(lp + 1)
                    if q + 1 < p["num_parents"]:
                        e0 = s0 + rngs[t].index(lp - s0 + 1)
                sel.append((par, s0, e0))
            groups[c // p["num_offspring"]] = sel
    children = []
    for t in range(pop):
        rng = rngs[t]
        if kinds[t] == 0:
            child = list(seqs[order[t]])
        elif kinds[t] == 1:
            child = random_sequence(p["rows"], p["max_ops"], rng)
        else:
            child = []
            for par, s0, e0 in groups[(t - p["elite"]) // p["num_offspring"]]:
                for k in range(s0, e0):
                    if len(child) >= p["max_ops"]:
                        break
                    child.append(seqs[par][k])
            if p["mutation_rate"] > 0:
                count = max(1, int(F(p["rows"]) * rate))
                for _ in range(count):
                    if p["learned"] and p["n_cand"] > 0 and rng.unit() < F(p["p_learned"]):
                        op = int(cand[rng.index(p["n_cand"])])
                    else:
                        i = rng.index(p["rows"])
                        j = rng.index(p["rows"])
                        if i == j:
                            continue
                        op = _op(j, i)
                    if len(child) < p["max_ops"]:
                        child.append(op)
                    else:
                        child[rng.index(len(child))] = op
            else:
                for _ in range(p["mutations"]):
                    u = rng.unit()
                    if u < F(p["p_append"]):
                        if len(child) < p["max_ops"]:
                            child.append(_propose(p, rng, cand))
                    elif u < F(p["p_append"]) + F(p["p_replace"]):
                        if child:
                            child[rng.index(len(child))] = _propose(p, rng, cand)
                    elif u < F(p["p_append"]) + F(p["p_replace"]) + F(p["p_delete"]):
                        if child:
                            del child[rng.index(len(child))]
        children.append(child)
    return children, [r.s for r in rngs], st, best_idx, mode


def params_of(gp, rows: int) -> dict:
    """The restatement's parameter dict from a bmm_ga_params structure."""
    return {"rows": rows, "pop": gp.population, "max_ops": gp.max_ops,
            "tournament": gp.tournament, "elite": gp.elite, "mutations": gp.mutations,
            "p_append": gp.p_append, "p_replace": gp.p_replace, "p_delete": gp.p_delete,
            "learned": gp.learned, "n_cand": gp.n_candidates if gp.learned else 0,
            "p_learned": gp.p_learned, "num_parents": gp.num_parents,
            "num_offspring": gp.num_offspring, "mutation_rate": gp.mutation_rate,
            "catastrophe_threshold": gp.catastrophe_threshold,
            "catastrophe_survival": gp.catastrophe_survival,
            "restart_after": gp.restart_after, "restart_p": gp.restart_p}

"""Known answers for the oracle's GA decisions on scripted random streams.

The oracle's Rng replays a script of draws (unit values, below(n) indices,
standard normal deviates) in place of its generator, so every decision of
Population::evolve (population.rs:74-158, ga.rs:7-39, chromosome.rs:56-77),
the early stop of the generation loop (run.rs:401-425) and model selection
(run.rs:694-733) can be derived by hand.  The expected populations below are
written out from the scripts, step by step in the comments.

Each test names the plausible mistake it must catch; the oracle is also built
with that mistake planted (oracle/mutants, ORACLE_MUTANT=k), and the test is
run against that build too, where it must fail."""
import math

import numpy as np
import pytest

import oracle.oracle as orc

TWO_PI = 2.0 * math.pi


def rem_euclid(a, b):  # Rust f64::rem_euclid
    r = math.fmod(a, b)
    return r + abs(b) if r < 0.0 else r


# ---- the evolve scenario: 8 individuals --------------------------------------
# Fitness with a three-way tie at the best value, so that the tie-breaks of
# the sorts and the tournament reach the children that survive.
FIT = [5.0, 1.0, 1.0, 7.0, 3.0, 9.0, 1.0, 2.0]
P = [
    [1.00, 6.20, 0.20, -40.0, 45.0, -49.0],
    [1.50, 5.50, 0.50, -38.0, 45.5, -46.0],
    [2.00, 4.80, 0.80, -36.0, 46.0, -43.0],
    [2.50, 4.10, 1.10, -34.0, 46.5, -40.0],
    [3.00, 3.40, 1.40, -32.0, 47.0, -37.0],
    [3.50, 2.70, 1.70, -30.0, 47.5, -34.0],
    [0.05, 2.00, 2.00, -28.0, 48.0, -31.0],
    [4.50, 6.25, 2.30, -26.0, 48.5, -28.0],
]
GENES = np.array(P)

# Tournament (population.rs:124-158): 8 rounds of 3 below(8) draws; the first
# minimum wins.  Winners: P0 (5,7,9); P1 (1,1,1: the first); P7 (3,3,2);
# P6 (1,1,1: the first); P7 (9,7,2); P0 (7,5,7); P2 (1,9,9); P1 (2,1,3).
TOURNAMENT = [0, 3, 5, 1, 2, 6, 4, 4, 7, 6, 2, 1, 5, 3, 7, 3, 0, 3, 2, 5, 5, 7, 1, 4]
# Fisher-Yates from the end (population.rs:88, SliceRandom::shuffle): swap
# [i-1] with [below(i)] for i = 8..2.  [P0 P1 P7 P6 P7 P0 P2 P1] ->
# i=8 j=2: [P0 P1 P1 P6 P7 P0 P2 P7]; i=7 j=6, i=6 j=0, i=5 j=4: no change;
# i=4 j=1: [P0 P6 P1 P1 P7 P0 P2 P7]; i=3 j=1: [P0 P1 P6 P1 P7 P0 P2 P7];
# i=2 j=0: [P1 P0 P6 P1 P7 P0 P2 P7].
SHUFFLE = [2, 6, 0, 4, 1, 1, 0]
# Crossover (ga.rs:7-39) of the pairs A (P1,P0), B (P6,P1), C (P7,P0),
# D (P2,P7): a gene swaps when its unit draw is <= 0.6.
CROSS = [0.9, 0.9, 0.9, 0.9, 0.9, 0.9,        # A: none
         0.9, 0.61, 0.9, 0.9, 0.9, 0.9,       # B: none (0.61 does not swap)
         0.1, 0.1, 0.1, 0.1, 0.1, 0.1,        # C: all
         0.6, 0.7, 0.5, 0.7, 0.5, 0.59999]    # D: genes 0 (0.6 itself swaps), 2, 4, 5
# Mutation (chromosome.rs:56-77) of the children A1 A2 B1 B2 C1 C2 D1 D2: per
# gene a unit draw, and when it is <= 0.1 a normal deviate (angles: + 0.174 z,
# rem_euclid 2 pi; shifts: + z, clamped to +-50).
MUT = ([0.5] * 6                                        # A1
       + [0.5, 0.5, 0.5, 0.0, 0.5, 0.5, 0.5]           # A2: gene 3, z = 0.5
       + [0.1, -1.0, 0.5, 0.5, 0.5, 0.5, 0.5]          # B1: gene 0 (0.1 itself mutates), z = -1
       + [0.5, 0.5, 0.5, 0.5, 0.05, 5.0, 0.5]          # B2: gene 4, z = 5
       + [0.5] * 12                                    # C1 C2
       + [0.5, 0.02, 9.0, 0.5, 0.5, 0.5, 0.1, -25.0]   # D1: gene 1 (z = 9), gene 5 (z = -25)
       + [0.5] * 6)                                    # D2
SCRIPT = TOURNAMENT + SHUFFLE + CROSS + MUT


def expected_population():
    """Elites (a stable ascending sort of the population): P1(1) P2(1) P6(1)
    P7(2) P4(3).  Children and their stale fitness (the first or second
    parent's): A1 = P1 (1), A2 = P0 (5), B1 = P6 (1), B2 = P1 (1), C1 (2),
    C2 (5), D1 (1), D2 (2).  Stable descending: A2 C2 C1 D2 A1 B1 B2 D1; the
    elites replace the first five, the worst, so B1, B2, D1 remain."""
    b1 = list(P[6])
    b1[0] = rem_euclid(P[6][0] + 0.174 * -1.0, TWO_PI)  # -0.124 wraps to 2 pi - 0.124
    b2 = list(P[1])
    b2[4] = max(min(P[1][4] + 1.0 * 5.0, 50.0), -50.0)  # 50.5 clamps to 50
    d1 = [P[7][0], P[2][1], P[7][2], P[2][3], P[7][4], P[7][5]]  # genes 0, 2, 4, 5 from P7
    d1[1] = rem_euclid(P[2][1] + 0.174 * 9.0, TWO_PI)  # 6.366 wraps past 2 pi
    d1[5] = max(min(P[7][5] + 1.0 * -25.0, 50.0), -50.0)  # -53 clamps to -50
    genes = [P[1], P[2], P[6], P[7], P[4], b1, b2, d1]
    fitness = [1.0, 1.0, 1.0, 2.0, 3.0, 1.0, 1.0, 1.0]
    return np.array(genes), np.array(fitness)


def expected_draws():
    d = [("below", 8)] * 24 + [("below", n) for n in range(8, 1, -1)] + [("unit", 0)] * 24
    mut_genes = {1: {3}, 2: {0}, 3: {4}, 6: {1, 5}}  # child -> mutated genes
    for child in range(8):
        for g in range(6):
            d.append(("unit", 0))
            if g in mut_genes.get(child, ()):
                d.append(("normal", 0))
    return d


def _lib(mutant):
    if mutant is None:
        return orc.lib()
    path = orc.MUTANTS / f"libgdock_oracle_m{mutant}.so"
    if not path.exists():
        orc.build()
    return orc.load(path)


def evolve_matches(L):
    genes, fit, draws, ok = orc.evolve_scripted(GENES, FIT, SCRIPT, L)
    eg, ef = expected_population()
    assert ok, "the script was not consumed exactly"
    assert draws == expected_draws()
    assert np.array_equal(fit, ef), fit
    assert np.array_equal(genes, eg), genes


# Each case: (the reference behaviour, the mutant that breaks it, a check).
def tournament_first_minimum(L):  # population.rs:144-151: min_by keeps the first of equal minima
    evolve_matches(L)


def elites_replace_the_worst_children(L):  # population.rs:104-118
    genes, fit, _, _ = orc.evolve_scripted(GENES, FIT, SCRIPT, L)
    eg, ef = expected_population()
    assert np.array_equal(genes[5:], eg[5:]) and np.array_equal(fit[5:], ef[5:])


def elite_sort_is_stable(L):  # population.rs:79-83: sort_by on fitness, stable
    genes, _, _, _ = orc.evolve_scripted(GENES, FIT, SCRIPT, L)
    eg, _ = expected_population()
    assert np.array_equal(genes[:5], eg[:5])


def crossover_rate_inclusive(L):  # ga.rs:21: gen_range(0.0..1.0) <= crossover_rate
    evolve_matches(L)


def mutation_rate_inclusive(L):  # chromosome.rs:62: gen_range(0.0..1.0) <= mutation_rate
    evolve_matches(L)


def mutation_wraps_negative_angles(L):  # chromosome.rs:66: rem_euclid, not %
    genes, _, _, _ = orc.evolve_scripted(GENES, FIT, SCRIPT, L)
    eg, _ = expected_population()
    assert genes[5, 0] == eg[5, 0] and 0.0 <= genes[5, 0] < TWO_PI


def shuffle_from_the_end(L):  # rand's SliceRandom::shuffle: Fisher-Yates from the last element
    evolve_matches(L)


def early_stop_strict_threshold(L):
    """run.rs:401-425: a generation stalls when (last - best) / |last| <
    0.001, counted from generation 1 (generation 0 has no previous best).
    -1000 then -1001: an improvement of exactly 0.001 does not stall; the
    flat generations 2..11 stall 1..10, so the loop stops after generation
    11 (12 generations evaluated)."""
    assert orc.ga_fitness_sequence([-1000.0, -1001.0], L=L) == (12, 11)


def early_stop_relative_to_magnitude(L):
    """The improvement divides by |last|: with negative fitness, a better
    (lower) best is a positive improvement.  -100 for 6 generations (stalls
    1..5), -200 (improvement 1: reset), then flat: stops after generation 16.
    Worse fitness is a negative improvement and stalls: 5 then 6 stops after
    generation 10.  Steady 1% improvements run to max_gen with no stop."""
    seq = [-100.0] * 6 + [-200.0]
    assert orc.ga_fitness_sequence(seq, L=L) == (17, 16)
    assert orc.ga_fitness_sequence([-1000.0, -1001.0], L=L) == (12, 11)
    assert orc.ga_fitness_sequence([5.0, 6.0], L=L) == (11, 10)
    steady = [-(1.01 ** g) for g in range(40)]
    assert orc.ga_fitness_sequence(steady, max_gen=40, L=L) == (40, -1)


def models_filled_by_fitness(L):
    """run.rs:694-733.  Hall of Fame fitness [-5, -2, -7, -9, -1, -8] and
    clusters (centre, size) [(4, 6), (0, 5), (3, 4)]: centres by fitness
    3 (-9), 0 (-5), 4 (-1); then the best unchosen entries 5 (-8), 2 (-7),
    each with cluster size 1."""
    assert orc.select_models([-5, -2, -7, -9, -1, -8], [(4, 6), (0, 5), (3, 4)], L) == \
        [(3, 4), (0, 5), (4, 6), (5, 1), (2, 1)]
    # no clusters: the five best, equal fitness in Hall of Fame order
    assert orc.select_models([-3, -7, -3, -7, 0, -1], [], L) == [(1, 1), (3, 1), (0, 1), (2, 1), (5, 1)]


CASES = [
    (tournament_first_minimum, 5),
    (elites_replace_the_worst_children, 1),
    (elite_sort_is_stable, 2),
    (crossover_rate_inclusive, 3),
    (mutation_rate_inclusive, 4),
    (mutation_wraps_negative_angles, 8),
    (shuffle_from_the_end, 10),
    (early_stop_strict_threshold, 6),
    (early_stop_relative_to_magnitude, 9),
    (models_filled_by_fitness, 7),
]


@pytest.mark.parametrize("case,mutant", CASES, ids=[c.__name__ for c, _ in CASES])
def test_oracle_ga_known_answer(case, mutant):
    case(_lib(None))


@pytest.mark.parametrize("case,mutant", CASES, ids=[f"{c.__name__}-kills-mutant-{m}" for c, m in CASES])
def test_known_answer_fails_on_its_mutant(case, mutant):
    with pytest.raises(AssertionError):
        case(_lib(mutant))


def test_models_more_clusters_than_models():
    """Seven clusters: the five whose centres score best, equal fitness in
    cluster order; fewer than five Hall of Fame entries: all of them."""
    hof = [-1.0, -6.0, -3.0, -6.0, -2.0, -5.0, -4.0, 0.0]
    clusters = [(0, 9), (1, 8), (2, 7), (3, 6), (4, 5), (5, 4), (6, 4)]
    assert orc.select_models(hof, clusters) == [(1, 8), (3, 6), (5, 4), (6, 4), (2, 7)]
    assert orc.select_models([-1.0, -2.0, -3.0], []) == [(2, 1), (1, 1), (0, 1)]


def test_hall_of_fame_from_a_population():
    """run.rs:194-216: candidates in stable ascending fitness order, an entry
    similar to a kept one (every angle within 0.2 rad, with wrap-around, and
    every shift within 2 A) is skipped, at most 10 are kept per generation.
    Individual 2 is similar to 1, 5 to 0, 7 to 3 across the 2 pi wrap; 13 is
    the eleventh acceptable one.  Kept: 0 1 3 4 6 8 9 10 11 12."""
    base = [[0.3 * i + 0.02, 0.4 * i + 0.1, 0.5 * i + 0.3, 3.0 * i, -2.5 * i, 4.0 * i - 20.0] for i in range(14)]
    base[2] = [base[1][0] + 0.1, base[1][1], base[1][2] - 0.15, base[1][3] + 1.0, base[1][4], base[1][5] - 1.5]
    base[5] = [base[0][0] - 0.05, base[0][1] + 0.15, base[0][2], base[0][3] - 2.0, base[0][4] + 2.0, base[0][5]]
    base[3][0] = 0.02
    base[7] = [6.25, base[3][1] + 0.1, base[3][2], base[3][3] + 0.5, base[3][4], base[3][5] - 0.5]
    fitness = [-10.0, -9.0, -9.0, -8.0, -7.0, -6.0, -5.0, -4.0, -3.0, -2.0, -1.0, 0.0, 1.0, 2.0]
    h = orc.Hof()
    h.add_population(np.array(base), fitness, 0)
    kept = [0, 1, 3, 4, 6, 8, 9, 10, 11, 12]
    assert np.array_equal(h.fitness(), [fitness[i] for i in kept])
    assert np.array_equal(h.genes(), np.array([base[i] for i in kept]))

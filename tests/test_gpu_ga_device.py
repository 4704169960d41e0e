"""The GA generation on the device (gdock_ga_run_device) against the host
driver (gdock_ga_run) on the same job: bit-identical populations, fitness,
best indices, Hall of Fame, clusters, selected models and final population;
and, like the host driver, the oracle's GA decisions (run.rs:361-733,
population.rs:74-158, ga.rs:7-39, chromosome.rs:27-77)."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import ga, structure, synth
from gdock_amd.device import Complex
from test_gpu_ga import _compare

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)


def _cfg0():
    c = synth.make_complex(0, rec_res=60, lig_res=76, n_restraints=6)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    return rec, lig, c.restraints, Complex(rec, lig, c.restraints)


def _identical(a, b):
    assert a.generations == b.generations and a.converged_at == b.converged_at
    for f in ("genes", "fitness", "best_idx", "hof_genes", "hof_fitness", "hof_generation", "cluster_center",
              "cluster_size", "final_genes"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), f
    assert a.selected == b.selected


@pytest.mark.parametrize("seed", [42, 7])
def test_device_ga_equals_host_driver_cfg0(seed):
    rec, lig, pairs, cx = _cfg0()
    host = ga.run(cx, seed=seed, keep_trace=1)
    dev = ga.run_device(cx, seed=seed, keep_trace=1)
    _identical(host, dev)
    ora = orc.GaRun(rec.atoms, lig.atoms, pairs, W, seed=seed, pop_size=150, max_gen=250, nthreads=16)
    outcome, compared = _compare(dev, ora, 150)
    print(f"seed {seed}: {dev.generations} generations, {len(dev.hof_fitness)} HoF, vs oracle: {outcome}", flush=True)
    assert compared >= min(10, dev.generations), outcome


def test_device_ga_odd_population_and_final_breed():
    """An odd population breeds one fewer child; without early stopping the
    loop breeds once more after its last generation."""
    c = synth.make_complex(5, rec_res=30, lig_res=25, n_restraints=4)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints)
    kw = dict(population_size=41, max_generations=30, early_stopping=0, keep_trace=1, elitism=3, tournament_size=4)
    host = ga.run(cx, seed=11, **kw)
    dev = ga.run_device(cx, seed=11, **kw)
    _identical(host, dev)
    assert dev.generations == 30 and dev.final_genes.shape == (40, 6)


def test_device_ga_hall_of_fame_overflow():
    """Enough distinct poses that the Hall of Fame passes 500 entries and is
    trimmed (first by a full sort, then by sorted insertion) many times."""
    rec, lig, pairs, cx = _cfg0()
    kw = dict(max_generations=90, early_stopping=0, mutation_rate=0.6, keep_trace=0, cluster=0)
    host = ga.run(cx, seed=3, **kw)
    dev = ga.run_device(cx, seed=3, **kw)
    _identical(host, dev)
    assert len(dev.hof_fitness) == 500
    assert np.all(np.diff(dev.hof_fitness) >= 0)


def test_device_ga_one_lane_draws_equal_host_driver():
    """Tournaments of 140 draw more indices (60 x 140 + 59) than the lanes'
    index buffer holds: one lane takes the generation's draws in order."""
    c = synth.make_complex(5, rec_res=30, lig_res=25, n_restraints=4)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints)
    kw = dict(population_size=60, max_generations=12, early_stopping=0, keep_trace=1, tournament_size=140)
    _identical(ga.run(cx, seed=5, **kw), ga.run_device(cx, seed=5, **kw))


@pytest.mark.parametrize("pop", [1024, 3])
def test_device_ga_population_extremes(pop):
    """The largest device population (each lane takes 112 consecutive draws)
    and a population of 3 (one crossed pair, one individual left over)."""
    rec, lig, pairs, cx = _cfg0()
    kw = dict(population_size=pop, max_generations=6, early_stopping=0, keep_trace=1, elitism=min(5, pop - 1))
    _identical(ga.run(cx, seed=9, **kw), ga.run_device(cx, seed=9, **kw))

"""The benchmark workloads at their full sizes.

- Uneven batch: 64 complexes (synth.uneven_batch(64, 100): receptor 500-8000,
  ligand 300-3000 atoms), 150 GA-initial poses each, one replay of the
  captured persistent launch, as the uneven_batch benchmark runs it.  Two
  replays are bit-identical; every complex's rows equal its single-complex
  rows within the energy bar; 2 poses of every complex and 4 of the 4
  largest match the FP64 oracle.
- Served stream: 16 GA docking jobs (synth.uneven_batch(16, 1000), population
  150, early stop after 10 flat generations, FCC of the Hall of Fame)
  submitted as a Poisson process of 20 jobs/s.  Each job's result equals the
  same job served alone, bit for bit, and takes the decisions of the plain
  path (gdock_ga_run, one scoring launch per generation) except after a
  near-tie of the two summation orders (population.rs:48-61, run.rs:361-733).
"""
import time

import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import ga, structure, synth
from gdock_amd.device import Batch, Complex, default_context
from gdock_amd.serve import Server
from helpers import assert_energies_close
from test_gpu_ga import _compare
from test_gpu_serve import _same

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)
PER = 150


@pytest.fixture(scope="module")
def uneven64():
    ctx = default_context(0)
    mols, cxs, genes = [], [], []
    for i, c in enumerate(synth.uneven_batch(64, 100)):
        rec = structure.read_pdb_text(c.receptor_pdb)[0]
        lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
        mols.append((rec, lig, c.restraints))
        cxs.append(Complex(rec, lig, c.restraints, ctx))
        genes.append(synth.ga_genes(100 + i, PER))
    return mols, cxs, genes


def test_uneven_batch_64_complexes(uneven64):
    mols, cxs, genes = uneven64
    b = Batch(cxs, [PER] * len(cxs))
    g = np.concatenate(genes)
    t0 = time.perf_counter()
    out1, ns1 = b.score(g)
    t1 = time.perf_counter()
    out2, ns2 = b.score(g)
    assert np.array_equal(out1, out2) and np.array_equal(ns1, ns2), "replays differ"
    assert np.isfinite(out1).all()
    sizes = [len(r) * len(l) for r, l, _ in mols]
    largest = set(np.argsort(sizes)[-4:].tolist())
    worst_elec = 0.0
    for i, (rec, lig, pairs) in enumerate(mols):
        a, z = b.offsets[i], b.offsets[i + 1]
        single, sns = cxs[i].score_poses(genes[i])
        assert_energies_close(out1[a:z], single, f"complex {i}: batch vs single")
        assert np.array_equal(ns1[a:z], sns)
        # the poses nearest the receptor (most pairs within the cutoffs) and the first pose
        near = np.argsort(np.abs(genes[i][:, 3:]).max(axis=1))
        pick = sorted({0, *near[:3 if i in largest else 1].tolist()})
        ref, rns = orc.score_poses(rec.atoms, lig.atoms, pairs, genes[i][pick], W, nthreads=16)
        assert_energies_close(out1[a:z][pick], ref, f"complex {i} ({len(rec)} x {len(lig)}) vs oracle")
        assert np.array_equal(ns1[a:z][pick], rns)
        worst_elec = max(worst_elec, float(np.abs(out1[a:z][pick, 1] - ref[:, 1]).max()))
    print(f"64 complexes x {PER} poses: first replay {1e3 * (t1 - t0):.1f} ms; "
          f"largest elec error on oracle-checked poses {worst_elec:.2e} kcal/mol", flush=True)
    assert worst_elec <= 5e-4


def test_served_poisson_stream_16_jobs():
    jobs = []
    for c in synth.uneven_batch(16, 1000):
        rec = structure.read_pdb_text(c.receptor_pdb)[0]
        lig = structure.read_pdb_text(c.ligand_pdb)[0]  # raw: the server positions it
        jobs.append((rec, lig, c.restraints))
    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / 20.0, size=len(jobs)))
    arrivals -= arrivals[0]
    params = dict(population_size=150, early_stopping=1, convergence_window=10, cluster=1, keep_trace=1)
    srv = Server(max_jobs=64, max_receptor_atoms=8192)
    try:
        t0 = time.perf_counter()
        ids = []
        for k, ((rec, lig, pairs), a) in enumerate(zip(jobs, arrivals)):
            d = a - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids.append(srv.submit(rec, lig, pairs, seed=42 + k, **params))
        mixed = [srv.wait(j, 600.0) for j in ids]
        peak = srv.stats()["peak_active"]
        alone = [srv.wait(srv.submit(rec, lig, pairs, seed=42 + k, **params), 600.0)[0]
                 for k, (rec, lig, pairs) in enumerate(jobs)]
    finally:
        srv.close()
    lat = np.array([t.latency_s for _, t in mixed])
    print(f"16 jobs at 20 jobs/s: latency p50 {np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s, "
          f"peak active {peak}", flush=True)
    assert peak >= 2
    outcomes = []
    for k, (rec, lig, pairs) in enumerate(jobs):
        res = mixed[k][0]
        assert res.generations >= 1 and len(res.hof_fitness) > 0
        _same(alone[k], res)
        cx = Complex(rec, structure.position_ligand(rec, lig), pairs)
        plain = ga.run(cx, seed=42 + k, **params)
        cx.close()
        outcome, compared = _compare(res, plain, 150)
        outcomes.append(outcome)
        assert compared >= min(5, res.generations), (k, outcome)
    print("served vs plain:", outcomes, flush=True)

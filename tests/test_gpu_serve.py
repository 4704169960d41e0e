"""Docking server: continuous batching of GA jobs (gdock_server_*).

A job's results must be the same alone and in a mix of other jobs: every
pose is scored by the same kernel code with the same per-shape lane choice,
so the whole GA (populations, fitness, Hall of Fame, clusters, models) is
bit-identical.  A served job is also checked against the oracle's GA."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import _lib, structure, synth
from gdock_amd.serve import Server
from test_gpu_ga import _compare

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)
SMALL = dict(population_size=60, max_generations=40, keep_trace=1)


def _job(seed, rec_res, lig_res, n_restraints=5):
    c = synth.make_complex(seed, rec_res=rec_res, lig_res=lig_res, n_restraints=n_restraints)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.read_pdb_text(c.ligand_pdb)[0]  # raw: the server positions it
    return rec, lig, c.restraints


def _same(a, b):
    assert a.generations == b.generations and a.converged_at == b.converged_at
    assert np.array_equal(a.genes, b.genes) and np.array_equal(a.fitness, b.fitness)
    assert np.array_equal(a.hof_genes, b.hof_genes) and np.array_equal(a.hof_fitness, b.hof_fitness)
    assert a.cluster_center.tolist() == b.cluster_center.tolist()
    assert a.cluster_size.tolist() == b.cluster_size.tolist()
    assert a.selected == b.selected


def test_alone_equals_in_mix():
    jobs = [_job(60 + i, 20 + 7 * i, 18 + 5 * i) for i in range(5)]
    srv = Server(max_jobs=8)
    try:
        alone, _ = srv.wait(srv.submit(*jobs[0], seed=3, **SMALL), 180.0)
        order = [1, 2, 0, 3, 4]
        ids = [srv.submit(*jobs[i], seed=3 if i == 0 else 10 + i, **SMALL) for i in order]
        got = {i: srv.wait(j, 180.0) for i, j in zip(order, ids)}
        _same(alone, got[0][0])
        assert srv.stats()["peak_active"] >= 2
        for i in order:
            assert got[i][1].latency_s > 0 and got[i][1].generations == got[i][0].generations
    finally:
        srv.close()


def test_split_job_alone_equals_in_mix():
    """A job large enough for split poses (5000 x 2000 atoms: a pose near the
    receptor over several workgroups) is bit-identical alone and among small
    jobs (a pose's split depends only on the pose and its complex), and takes
    the same decisions as the plain one-launch-per-generation path (whose
    poses are unsplit: sums in another order, so near-ties may differ)."""
    from gdock_amd import ga
    from gdock_amd.device import Complex
    c = synth.make_complex(91, rec_atoms=5000, lig_atoms=2000, n_restraints=6)
    big = (structure.read_pdb_text(c.receptor_pdb)[0], structure.read_pdb_text(c.ligand_pdb)[0], c.restraints)
    small = [_job(92 + i, 30, 25) for i in range(3)]
    params = dict(population_size=40, max_generations=10, keep_trace=1)
    srv = Server(max_jobs=6)
    try:
        alone, _ = srv.wait(srv.submit(*big, seed=8, **params), 180.0)
        ids = [srv.submit(*j, seed=20 + k, **params) for k, j in enumerate(small[:2])]
        ids.append(srv.submit(*big, seed=8, **params))
        ids.append(srv.submit(*small[2], seed=30, **params))
        got = [srv.wait(j, 180.0)[0] for j in ids]
    finally:
        srv.close()
    _same(alone, got[2])
    rec, lig, pairs = big
    plain = ga.run(Complex(rec, structure.position_ligand(rec, lig), pairs), seed=8, **params)
    outcome, compared = _compare(alone, plain, 40)
    print("split served vs plain:", outcome)
    assert compared >= min(5, alone.generations), outcome


def test_served_job_matches_oracle_ga():
    rec, lig, pairs = _job(0, 60, 76, 6)
    srv = Server(max_jobs=4)
    try:
        jid = srv.submit(rec, lig, pairs, seed=42, keep_trace=1)
        res, t = srv.wait(jid, 180.0)
    finally:
        srv.close()
    lig_pos = structure.position_ligand(rec, lig)
    ora = orc.GaRun(rec.atoms, lig_pos.atoms, pairs, W, seed=42, pop_size=150, max_gen=250, nthreads=16)
    outcome, compared = _compare(res, ora, 150)
    print("served vs oracle:", outcome, f"latency {t.latency_s:.2f} s")
    assert compared >= min(10, res.generations), outcome


def test_memory_admission_limits_active_jobs():
    jobs = [_job(80 + i, 25, 20) for i in range(5)]
    est = max(256 * (len(r) + len(l)) + (1 << 20) for r, l, _ in jobs)
    srv = Server(max_jobs=8, memory_budget=2 * est + 1000)
    try:
        ids = [srv.submit(*j, seed=5 + k, population_size=40, max_generations=12) for k, j in enumerate(jobs)]
        for j in ids:
            res, t = srv.wait(j, 180.0)
            assert res.generations >= 1 and t.admit_s >= t.submit_s
        st = srv.stats()
        assert st["jobs_finished"] == 5 and st["peak_active"] <= 2 and st["bytes_in_use"] == 0
        # the scoring, preparing and finishing threads each account their time
        assert st["busy_s"] > 0 and st["admit_s"] > 0 and st["ga_s"] > 0 and st["finish_s"] >= 0
    finally:
        srv.close()


def test_submit_rejects_bad_parameters():
    rec, lig, pairs = _job(90, 20, 15)
    srv = Server(max_jobs=2, max_population=50)
    try:
        with pytest.raises(_lib.GdockError) as e:
            srv.submit(rec, lig, pairs, population_size=80)
        assert e.value.status == _lib.INVALID_ARG
        with pytest.raises(_lib.GdockError):
            srv.wait(12345)
    finally:
        srv.close()


def test_job_failing_in_preparation_leaves_the_server_working():
    """A receptor larger than the server's limit is accepted at submission and
    fails while the job is prepared (on the preparing thread): its wait
    reports the error, its memory is released, and the next job runs."""
    big = _job(93, 60, 20)
    small = _job(94, 20, 15)
    srv = Server(max_jobs=2, max_receptor_atoms=200)
    try:
        bad = srv.submit(*big, seed=1, population_size=20, max_generations=3)
        good = srv.submit(*small, seed=2, population_size=20, max_generations=3)
        with pytest.raises(_lib.GdockError) as e:
            srv.wait(bad, 120.0)
        assert e.value.status == _lib.INVALID_ARG
        res, _ = srv.wait(good, 120.0)
        assert res.generations >= 1
        st = srv.stats()
        assert st["jobs_finished"] == 2 and st["bytes_in_use"] == 0
    finally:
        srv.close()


def test_destroy_ends_pending_waits():
    """Destroying the server while another thread waits on a job that has not
    ended: the job ends as failed, the wait returns NOT_FOUND, and destroy
    returns once the waiter has left.  One job slot and three jobs of 5000
    generations each (seconds apiece): the third is still queued when destroy
    begins."""
    import threading
    import time
    jobs = [_job(95 + i, 120, 90) for i in range(3)]
    srv = Server(max_jobs=1)
    ids = [srv.submit(*j, seed=1 + k, max_generations=5000, early_stopping=0) for k, j in enumerate(jobs)]
    got = {}

    def waiter():
        try:
            srv.wait(ids[2], 300.0)
            got["status"] = _lib.OK
        except _lib.GdockError as e:
            got["status"] = e.status

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.3)
    srv.close()
    t.join(60.0)
    assert not t.is_alive() and got["status"] == _lib.NOT_FOUND


def test_device_ga_alone_equals_in_mix():
    """With ga_on_device the populations stay in HBM and advance in one GA
    launch per step (one workgroup per job, each job its own stream): a job is
    bit-identical alone and in a mix, trace, Hall of Fame and models too."""
    jobs = [_job(60 + i, 20 + 7 * i, 18 + 5 * i) for i in range(4)]
    srv = Server(max_jobs=8, ga_on_device=True)
    try:
        alone, _ = srv.wait(srv.submit(*jobs[0], seed=3, **SMALL), 180.0)
        order = [1, 0, 2, 3]
        ids = [srv.submit(*jobs[i], seed=3 if i == 0 else 10 + i, **SMALL) for i in order]
        got = {i: srv.wait(j, 180.0) for i, j in zip(order, ids)}
        _same(alone, got[0][0])
        st = srv.stats()
        assert st["peak_active"] >= 2 and st["ga_s"] > 0
        for i in order:
            assert got[i][0].generations >= 1 and got[i][1].generations == got[i][0].generations
    finally:
        srv.close()


def test_device_ga_served_job_matches_oracle_ga():
    """The 60 x 76-residue job served with the GA on the device takes the
    oracle GA's decisions on the same stream (near-ties of the two scorers
    aside), and equals the same job served with the host GA wherever their
    scores agree."""
    rec, lig, pairs = _job(0, 60, 76, 6)
    srv = Server(max_jobs=4, ga_on_device=True)
    try:
        res, t = srv.wait(srv.submit(rec, lig, pairs, seed=42, keep_trace=1), 180.0)
    finally:
        srv.close()
    lig_pos = structure.position_ligand(rec, lig)
    ora = orc.GaRun(rec.atoms, lig_pos.atoms, pairs, W, seed=42, pop_size=150, max_gen=250, nthreads=16)
    outcome, compared = _compare(res, ora, 150)
    print("device-GA served vs oracle:", outcome, f"latency {t.latency_s:.2f} s")
    assert compared >= min(10, res.generations), outcome


def test_device_ga_population_limit():
    with pytest.raises(_lib.GdockError):
        Server(max_jobs=2, max_population=2048, ga_on_device=True)
    srv = Server(max_jobs=2, max_population=2048)  # the host GA has no such limit
    srv.close()

"""GA docking job on the device against the oracle's GA (run.rs restated),
same seed, same random stream.

The device scores every generation within the energy tolerance of the FP64
oracle, so the two runs take the same decisions (tournaments, elite, Hall of
Fame, early stop) unless two fitness values in a generation are ordered
differently by the two scorers, which can only happen when they lie within
the tolerance of each other (a near-tie).  The test walks the generations:
every order flip must be a near-tie, and the populations must stay
bit-identical until a generation after a near-tie (a flip between two
individuals that never meet in a decision changes nothing, so the walk goes
on past it).  With no divergence the whole run (generations, Hall of Fame,
clusters, selected models) must match."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import ga, structure, synth
from gdock_amd.device import Complex
from helpers import RESTRAINTS_2OOB, read_text

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)


def _fitness_tol(f):
    # components within max(1e-5 |E|, 1e-3): the weighted sum within the weighted tolerances
    return 2e-5 * np.abs(f) + (0.4 + 0.05 + 3.4 + 1.0) * 1e-3


def _order_flips(fg, fo):
    """Pairs (i, j) ordered differently by the two fitness vectors."""
    sg = np.sign(fg[:, None] - fg[None, :])
    so = np.sign(fo[:, None] - fo[None, :])
    return np.argwhere(sg != so)


def _compare(res, ora, pop, threshold=0.001):
    """Walks both traces; returns (how the comparison ended, generations compared)."""
    G = min(res.generations, ora.generations)
    tie = None  # first generation with a near-tie order flip
    for g in range(G):
        if not np.array_equal(res.genes[g], ora.genes[g]):
            assert tie is not None, f"generation {g}: populations differ with equal decisions"
            return f"diverged at generation {g} after a near-tie at generation {tie}", g
        fg, fo = res.fitness[g], ora.fitness[g]
        tol = _fitness_tol(fo)
        assert (np.abs(fg - fo) <= tol).all(), f"generation {g}: fitness off by {np.abs(fg - fo).max()}"
        for i, j in _order_flips(fg, fo):  # the scorers order two individuals differently: must be a near-tie
            assert abs(fo[i] - fo[j]) <= tol[i] + tol[j], f"generation {g}: order flip of a non-tie {fo[i]} {fo[j]}"
            tie = g if tie is None else tie
        if g > 0:  # run.rs early stop: relative improvement of the best against the threshold
            bg0, bg1 = res.fitness[g - 1].min(), fg.min()
            bo0, bo1 = ora.fitness[g - 1].min(), fo.min()
            ig, io = (bg0 - bg1) / abs(bg0), (bo0 - bo1) / abs(bo0)
            if (ig < threshold) != (io < threshold):
                assert abs(io - threshold) <= abs(ig - io) + 1e-12, f"generation {g}: early-stop decision differs"
                return f"early-stop near-tie at generation {g}", g + 1
    assert res.generations == ora.generations and res.converged_at == ora.converged_at
    if tie is not None:  # same populations throughout; the Hall of Fame may order the tied pair differently
        return f"same populations, near-tie at generation {tie}", G
    assert np.array_equal(res.best_idx, ora.best_idx)
    assert np.array_equal(res.hof_genes, ora.hof_genes)
    assert np.array_equal(res.hof_generation, ora.hof_generation)
    assert (np.abs(res.hof_fitness - ora.hof_fitness) <= _fitness_tol(ora.hof_fitness)).all()
    assert res.cluster_center.tolist() == ora.cluster_center.tolist()
    assert res.cluster_size.tolist() == ora.cluster_size.tolist()
    assert res.selected == ora.selected
    return "identical", G


@pytest.mark.parametrize("seed", [42, 7])
def test_ga_synthetic_60x76_matches_oracle(seed):
    c = synth.make_complex(0, rec_res=60, lig_res=76, n_restraints=6)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints)
    res = ga.run(cx, seed=seed, keep_trace=1)
    ora = orc.GaRun(rec.atoms, lig.atoms, c.restraints, W, seed=seed, pop_size=150, max_gen=250, nthreads=16)
    outcome, compared = _compare(res, ora, 150)
    print(f"seed {seed}: {res.generations} generations, {len(res.hof_fitness)} HoF, outcome: {outcome}")
    assert compared >= min(10, res.generations), outcome


def test_ga_2oob_short_run_matches_oracle():
    rec = structure.read_pdb_text(read_text("2oob_A.pdb"))[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(read_text("2oob_B.pdb"))[0])
    cx = Complex(rec, lig, RESTRAINTS_2OOB)
    res = ga.run(cx, seed=3, keep_trace=1, max_generations=25)
    ora = orc.GaRun(rec.atoms, lig.atoms, RESTRAINTS_2OOB, W, seed=3, pop_size=150, max_gen=25, nthreads=16)
    outcome, compared = _compare(res, ora, 150)
    print(f"2oob: {res.generations} generations, outcome: {outcome}")
    assert compared >= min(10, res.generations), outcome


def test_ga_small_population_parameters():
    c = synth.make_complex(5, rec_res=30, lig_res=25, n_restraints=4)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints)
    res = ga.run(cx, seed=11, keep_trace=1, population_size=40, max_generations=30)
    ora = orc.GaRun(rec.atoms, lig.atoms, c.restraints, W, seed=11, pop_size=40, max_gen=30, nthreads=16)
    outcome, compared = _compare(res, ora, 40)
    assert compared >= min(10, res.generations), outcome
    assert res.generations <= 30 and len(res.selected) <= 5


def test_run_outputs_written_as_the_reference(tmp_path):
    """run.rs:738-880 after a GA run: model_<i>.pdb for the selected models,
    ranked_<i>.pdb for the five best Hall of Fame entries, metrics.tsv; each
    PDB is the receptor then the ligand moved by the entry's genes (equal to
    the oracle's apply_genes up to the three written decimals).  With an
    evaluator the metrics carry DockQ, L-RMSD, iRMSD and fnat."""
    from gdock_amd.evaluator import Evaluator
    rec = structure.read_pdb_text(read_text("2oob_A.pdb"))[0]
    native = structure.read_pdb_text(read_text("2oob_B.pdb"))[0]
    lig = structure.position_ligand(rec, native)
    cx = Complex(rec, lig, RESTRAINTS_2OOB)
    res = ga.run(cx, seed=9, max_generations=20)
    assert len(res.selected) >= 1
    order = np.argsort(res.hof_fitness, kind="stable")[:5]
    for ev, sub in ((None, "plain"), (Evaluator(cx, native), "dockq")):
        out = tmp_path / sub
        ga.write_outputs(cx, res, out, evaluator=ev)
        rows = (out / "metrics.tsv").read_text().splitlines()
        head = "model\tcluster_size\tscore" + ("\tdockq\trmsd\tirmsd\tfnat" if ev else "")
        assert rows[0] == head
        models = [(f"model_{i + 1}", h, str(s)) for i, (h, s) in enumerate(res.selected)]
        models += [(f"ranked_{r + 1}", int(h), "-") for r, h in enumerate(order)]
        assert len(rows) == 1 + len(models)
        met = ev.calc_metrics(res.hof_genes[[h for _, h, _ in models]]) if ev else None
        for k, ((name, h, size), row) in enumerate(zip(models, rows[1:])):
            want = [name, size, f"{res.hof_fitness[h]:.4f}"]
            if ev:
                rmsd, irmsd, fnat, dockq = met[k]
                want += [f"{dockq:.4f}", f"{rmsd:.4f}", f"{irmsd:.4f}", f"{fnat:.4f}"]
            assert row.split("\t") == want
            text = (out / f"{name}.pdb").read_text()
            moved = structure.apply_genes(lig, res.hof_genes[h])
            assert text == structure.to_pdb_string(structure.combine_molecules(rec, moved))
            back = structure.read_pdb_text(text)[0]
            ref = orc.apply_genes(lig.atoms, res.hof_genes[h])
            assert len(back) == len(rec) + len(lig)
            for f in ("x", "y", "z"):
                assert np.abs(back.atoms[f][len(rec):] - ref[f]).max() <= 5.0001e-4
                assert np.array_equal(back.atoms[f][:len(rec)], rec.atoms[f])


def test_no_clust_outputs_from_the_final_population(tmp_path):
    """`--no-clust` (run.rs:549-637): the population the generation loop
    leaves behind (after the last generation the loop has evolved once more)
    is scored; best_by_score.pdb is its first pose of least fitness, and with
    an evaluator best_by_dockq.pdb is its last pose of greatest DockQ, with
    both in metrics.tsv.  The final population equals the oracle GA's."""
    from gdock_amd.evaluator import Evaluator
    rec = structure.read_pdb_text(read_text("2oob_A.pdb"))[0]
    native = structure.read_pdb_text(read_text("2oob_B.pdb"))[0]
    lig = structure.position_ligand(rec, native)
    cx = Complex(rec, lig, RESTRAINTS_2OOB)
    res = ga.run(cx, seed=3, keep_trace=1, max_generations=25, cluster=0)
    ora = orc.GaRun(rec.atoms, lig.atoms, RESTRAINTS_2OOB, W, seed=3, pop_size=150, max_gen=25, cluster=False,
                    nthreads=16)
    outcome, compared = _compare(res, ora, 150)
    assert compared >= min(10, res.generations), outcome
    if outcome == "identical" and res.converged_at < 0:
        assert res.generations == 25
        assert np.array_equal(res.final_genes, ora.final_genes)
        assert not np.array_equal(res.final_genes, res.genes[-1])  # bred once more after generation 25
    ev = Evaluator(cx, native)
    out = tmp_path / "noclust"
    ga.write_best_outputs(cx, res, out, evaluator=ev)
    fit = cx.score_poses(res.final_genes)[0][:, 5]
    met = ev.calc_metrics(res.final_genes)
    best = int(np.argmin(fit))                                        # first minimum
    top = len(met) - 1 - int(np.argmax(met[::-1, 3]))                 # last maximum
    for name, i in (("best_by_score", best), ("best_by_dockq", top)):
        moved = structure.apply_genes(lig, res.final_genes[i])
        assert (out / f"{name}.pdb").read_text() == structure.to_pdb_string(structure.combine_molecules(rec, moved))
    rows = (out / "metrics.tsv").read_text().splitlines()
    assert rows[0] == "model\tdockq\trmsd\tirmsd\tfnat\tscore"
    for row, (name, i) in zip(rows[1:], (("best_by_score", best), ("best_by_dockq", top))):
        rmsd, irmsd, fnat, dockq = met[i]
        assert row.split("\t") == [name, f"{dockq:.4f}", f"{rmsd:.4f}", f"{irmsd:.4f}", f"{fnat:.4f}", f"{fit[i]:.4f}"]
    plain = tmp_path / "plain"
    ga.write_best_outputs(cx, res, plain)
    assert sorted(p.name for p in plain.iterdir()) == ["best_by_score.pdb"]

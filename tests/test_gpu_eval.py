"""DockQ evaluator on the device (gdock_calc_metrics) against the oracle's
restatement of evaluator.rs.

Bar: fnat exact (integer contact counts over FP64 distance decisions);
rmsd, irmsd and dockq within 1e-12 relative (the device sums the per-atom
distances in a different order than the reference's sequential sum)."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import structure, synth
from gdock_amd.device import Complex
from gdock_amd.evaluator import Evaluator
from helpers import RESTRAINTS_2OOB, read_text

pytestmark = pytest.mark.gpu


def _check(rec, lig_pos, ref, genes, what):
    cx = Complex(rec, lig_pos)
    ev = Evaluator(cx, ref)
    got = ev.calc_metrics(genes)
    oev = orc.Evaluator(rec.atoms, ref.atoms)
    assert ev.n_native == oev.n_native
    assert ev.n_receptor_interface_res == len(oev.receptor_interface)
    assert ev.n_reference_interface_res == len(oev.reference_interface)
    exp = oev.calc_metrics([orc.apply_genes(lig_pos.atoms, g) for g in genes])
    assert np.array_equal(got[:, 2], exp[:, 2]), what  # fnat
    for c in (0, 1, 3):
        assert np.allclose(got[:, c], exp[:, c], rtol=1e-12, atol=1e-14), (what, c)
    return got, exp


def _native_genes(rec, lig):
    """Genes that carry the positioned ligand back onto the input (native) one."""
    cr, cl = np.array(orc.center_of_mass(rec.atoms)), np.array(orc.center_of_mass(lig.atoms))
    return np.array([[0.0, 0.0, 0.0, *(cl - cr)]])


def test_2oob_native_and_perturbed_poses():
    rec = structure.read_pdb_text(read_text("2oob_A.pdb"))[0]
    native = structure.read_pdb_text(read_text("2oob_B.pdb"))[0]
    lig = structure.position_ligand(rec, native)
    g0 = _native_genes(rec, native)
    rng = np.random.default_rng(3)
    near = g0 + np.concatenate([rng.normal(0, 0.08, (24, 3)), rng.normal(0, 1.5, (24, 3))], axis=1)
    genes = np.concatenate([g0, near, synth.ga_genes(4, 16)])
    got, _ = _check(rec, lig, native, genes, "2oob")
    assert got[0, 2] == 1.0 and got[0, 3] > 0.999  # the native pose
    assert (got[:, 2] <= 1.0).all() and (got[:, 3] <= 1.0 + 1e-12).all()


def test_synthetic_reference_and_edges():
    c = synth.make_complex(17, rec_res=70, lig_res=50)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    raw = structure.read_pdb_text(c.ligand_pdb)[0]
    lig = structure.position_ligand(rec, raw)
    # a reference placed in contact: the ligand moved onto the receptor surface
    ref = structure.Molecule(orc.apply_genes(lig.atoms, [0.4, 1.1, 2.0, 9.0, 3.0, -4.0]))
    genes = np.concatenate([[[0.4, 1.1, 2.0, 9.0, 3.0, -4.0]], synth.random_genes(17, 40, 15.0)])
    got, _ = _check(rec, lig, ref, genes, "synthetic")
    assert got[0, 0] < 1e-6 and got[0, 2] == 1.0
    # a reference far from the receptor: no interface, fnat 0, irmsd 0
    far = structure.Molecule(orc.apply_genes(lig.atoms, [0.0, 0.0, 0.0, 80.0, 0.0, 0.0]))
    got, _ = _check(rec, lig, far, genes[:5], "no interface")
    assert (got[:, 2] == 0.0).all() and (got[:, 1] == 0.0).all()


def test_ga_debug_mode_dockq_matches_oracle():
    """run.rs debug mode: the GA minimises -DockQ against the reference.  The
    device and the oracle GA take the same decisions (DockQ differs only in
    the last bits from the summation order), so the runs are identical."""
    from gdock_amd import ga
    from test_gpu_ga import _compare
    rec = structure.read_pdb_text(read_text("2oob_A.pdb"))[0]
    native = structure.read_pdb_text(read_text("2oob_B.pdb"))[0]
    lig = structure.position_ligand(rec, native)
    cx = Complex(rec, lig)
    ev = Evaluator(cx, native)
    res = ga.run_dockq(cx, ev, seed=5, keep_trace=1, population_size=60, max_generations=30)
    ora = orc.GaRun(rec.atoms, lig.atoms, [], seed=5, pop_size=60, max_gen=30, reference=native.atoms)
    assert np.allclose(res.fitness, ora.fitness[:len(res.fitness)], rtol=1e-12, atol=1e-14)
    outcome, compared = _compare(res, ora, 60)
    assert compared >= min(10, res.generations), outcome
    assert (res.fitness <= 0).all() and res.fitness.min() >= -1.0 - 1e-12
    print("debug-mode GA vs oracle:", outcome, res.generations, "generations, best DockQ", -res.fitness.min())


def test_debug_no_clust_outputs_rank_by_dockq(tmp_path):
    """`--debug --no-clust` (run.rs:549-637 with the debug fitness of
    chromosome.rs:97-108): the final population's fitness is -DockQ, so
    best_by_score.pdb is its first pose of greatest DockQ, best_by_dockq.pdb
    its last such pose, and metrics.tsv carries -DockQ as the score."""
    from gdock_amd import ga
    rec = structure.read_pdb_text(read_text("2oob_A.pdb"))[0]
    native = structure.read_pdb_text(read_text("2oob_B.pdb"))[0]
    lig = structure.position_ligand(rec, native)
    cx = Complex(rec, lig)
    ev = Evaluator(cx, native)
    res = ga.run_dockq(cx, ev, seed=5, population_size=60, max_generations=12, cluster=0)
    fin = res.final_genes
    # two copies of the best pose: the first and the last of greatest DockQ differ
    met0 = ev.calc_metrics(fin)
    k = int(np.argmax(met0[:, 3]))
    fin = np.concatenate([fin, fin[k:k + 1]])
    res.final_genes = fin
    out = tmp_path / "debug"
    ga.write_best_outputs(cx, res, out, evaluator=ev, debug=True)
    met = ev.calc_metrics(fin)
    ref = orc.Evaluator(rec.atoms, native.atoms).calc_metrics([orc.apply_genes(lig.atoms, g) for g in fin])
    assert np.allclose(met, ref, rtol=1e-12, atol=1e-12)
    fit = -met[:, 3]
    best = int(np.argmin(fit))                                   # first minimum of -DockQ
    top = len(met) - 1 - int(np.argmax(met[::-1, 3]))            # last maximum of DockQ
    assert best == k and top == len(fin) - 1
    for name, i in (("best_by_score", best), ("best_by_dockq", top)):
        moved = structure.apply_genes(lig, fin[i])
        assert (out / f"{name}.pdb").read_text() == structure.to_pdb_string(structure.combine_molecules(rec, moved))
    rows = (out / "metrics.tsv").read_text().splitlines()
    assert rows[0] == "model\tdockq\trmsd\tirmsd\tfnat\tscore"
    for row, (name, i) in zip(rows[1:], (("best_by_score", best), ("best_by_dockq", top))):
        rmsd, irmsd, fnat, dockq = met[i]
        assert row.split("\t") == [name, f"{dockq:.4f}", f"{rmsd:.4f}", f"{irmsd:.4f}", f"{fnat:.4f}", f"{-dockq:.4f}"]
    with pytest.raises(ValueError):
        ga.write_best_outputs(cx, res, tmp_path / "x", debug=True)

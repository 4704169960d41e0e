"""GPU contacts and FCC clusters for every chain layout (clustering.rs:62-115
keys residues by (chain, resseq) over the whole complex): bit-exact against
the oracle with two receptor chains, two ligand chains, one chain ID shared
by receptor and ligand (no candidate pair: empty sets, no clusters, models
from the Hall of Fame as run.rs:713-733), and residues that hold receptor and
ligand atoms; ligand atom pairs at exactly the cutoff, re-decided per pose."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import ga, structure, synth
from gdock_amd.clustering import ClusteringConfig, cluster_structures
from gdock_amd.device import Complex
from gdock_amd.structure import make_atom, molecule
from helpers import TOPOLOGIES, read_text, relabel
from test_gpu_fcc import docked_genes

pytestmark = pytest.mark.gpu


def _topology(name):
    rr, lr = TOPOLOGIES[name]
    rec = structure.Molecule(relabel(structure.read_pdb_text(read_text("2oob_A.pdb"))[0].atoms, rr))
    lig0 = structure.Molecule(relabel(structure.read_pdb_text(read_text("2oob_B.pdb"))[0].atoms, lr))
    lig = structure.position_ligand(rec, lig0)
    return rec, lig, Complex(rec, lig)


def _oracle_sets(rec, lig, genes, cutoff):
    return [orc.contacts(np.concatenate([rec.atoms, orc.apply_genes(lig.atoms, g)]), cutoff) for g in genes]


def _genes():
    return np.concatenate([docked_genes(1, 4, 8), synth.random_genes(2, 16, 12.0), synth.ga_genes(3, 8)])


@pytest.mark.parametrize("name", sorted(TOPOLOGIES))
def test_contacts_chain_layouts_bit_exact(name):
    rec, lig, cx = _topology(name)
    genes = _genes()
    print(f"layout {name}: {len(genes)} poses", flush=True)
    for cutoff in (5.0, 3.5):
        bits, sizes = cx.contacts(genes, cutoff)
        ref = _oracle_sets(rec, lig, genes, cutoff)
        for i in range(len(genes)):
            assert cx.contact_set(bits[i], cutoff) == ref[i], (name, cutoff, i)
            assert sizes[i] == len(ref[i])
        if name == "shared_chain_id":
            assert cx.contact_words(cutoff) == 0 and not sizes.any()
        else:
            assert sizes.max() > 0


@pytest.mark.parametrize("name", sorted(TOPOLOGIES))
def test_cluster_chain_layouts_exact(name):
    rec, lig, cx = _topology(name)
    genes = docked_genes(5, 6, 10)
    for cfg in (ClusteringConfig(), ClusteringConfig(4.0, 0.75, 1.0, 2)):
        got = cluster_structures(cx, genes, cfg)
        ref, _ = orc.cluster_poses(rec.atoms, lig.atoms, genes,
                                   (cfg.contact_distance, cfg.fcc_cutoff, cfg.strictness, cfg.min_cluster_size))
        assert [(r.center_idx, r.members) for r in got] == ref, (name, cfg)
        if name == "shared_chain_id":
            assert ref == []


@pytest.mark.parametrize("name", ["receptor_two_chains", "shared_chain_id"])
def test_ga_run_completes_on_chain_layouts(name):
    """A short GA job ends normally: clusters of its Hall of Fame equal the
    oracle's, and the selected models follow run.rs:694-733 (clusters first,
    then the Hall of Fame by fitness; only the latter without clusters)."""
    rec, lig, cx = _topology(name)
    res = ga.run(cx, seed=3, max_generations=12)
    assert res.generations >= 1 and len(res.hof_fitness) > 0
    ref, _ = orc.cluster_poses(rec.atoms, lig.atoms, res.hof_genes)
    assert res.cluster_center.tolist() == [c for c, _ in ref]
    assert res.cluster_size.tolist() == [len(m) for _, m in ref]
    assert res.selected == orc.select_models(res.hof_fitness, [(c, len(m)) for c, m in ref])
    if name == "shared_chain_id":
        assert len(ref) == 0 and len(res.selected) == min(5, len(res.hof_fitness))
        assert all(size == 1 for _, size in res.selected)


def window_complex():
    """A small receptor (chain A) and a ligand of two chains (B, D) whose
    residue pairs (B k, D k) each hold one atom pair exactly 5 A apart and no
    other pair within 6 A: whether such a pair is a contact turns on the last
    bits of each pose's FP64 transform."""
    rec = [make_atom(20.0 + 1.4 * i, 0.5 * i, -3.0, serial=i + 1, chainid="A", resseq=i // 3 + 1) for i in range(12)]
    pairs = [((0, 0, 0), (5, 0, 0)), ((0, 10, 0), (3, 14, 0)), ((0, 0, 10), (0, 4, 13)), ((10, 10, 10), (10, 13, 14))]
    atoms, s = [], 1
    for k, (a, b) in enumerate(pairs):
        away_a = tuple(x - 1.5 * (y - x) / 5.0 for x, y in zip(a, b))
        away_b = tuple(y + 1.5 * (y - x) / 5.0 for x, y in zip(a, b))
        for p in (a, away_a):
            atoms.append(make_atom(*p, serial=s, chainid="B", resseq=k + 1))
            s += 1
        for p in (b, away_b):
            atoms.append(make_atom(*p, serial=s, chainid="D", resseq=k + 1))
            s += 1
    rec = molecule(*rec)
    return rec, structure.position_ligand(rec, molecule(*atoms))


def test_ligand_pairs_at_the_cutoff_per_pose():
    rec, lig = window_complex()
    cx = Complex(rec, lig)
    genes = synth.random_genes(11, 200, 4.0)
    bits, sizes = cx.contacts(genes, 5.0)
    ref = _oracle_sets(rec, lig, genes, 5.0)
    hits = {}
    for i in range(len(genes)):
        assert cx.contact_set(bits[i]) == ref[i], i
        for c in ref[i]:
            hits[c] = hits.get(c, 0) + 1
    for k in range(1, 5):  # each pair is in some poses' sets and not in others'
        assert 0 < hits.get(f"B {k} D {k}", 0) < len(genes)
    got = cx.cluster(genes, ClusteringConfig(5.0, 0.6, 0.75, 2))
    exp, _ = orc.cluster_poses(rec.atoms, lig.atoms, genes, (5.0, 0.6, 0.75, 2))
    assert [(r.center_idx, r.members) for r in got] == exp


def test_ligand_pairs_reject_far_poses():
    """Past the translation within which a ligand-ligand decision taken once
    is exact, a multi-chain ligand's poses are rejected, not guessed."""
    rec, lig = window_complex()
    cx = Complex(rec, lig)
    g = synth.random_genes(12, 2, 4.0)
    g[1, 3] = 1e9
    with pytest.raises(Exception, match="translation"):
        cx.contacts(g, 5.0)

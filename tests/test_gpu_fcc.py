"""GPU FCC clustering: contact sets and clusters bit-exact against the oracle."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import structure, synth
from gdock_amd.clustering import ClusteringConfig, calculate_contacts, cluster_structures
from gdock_amd.device import Complex
from helpers import read_text

pytestmark = pytest.mark.gpu


def _complex(text_r, text_l):
    rec = structure.read_pdb_text(text_r)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(text_l)[0])
    return rec, lig, Complex(rec, lig)


def docked_genes(seed, n_base, per_base, rot=0.06, shift=0.7, spread=14.0):
    """Poses in a few tight families near the receptor, so FCC finds clusters."""
    rng = np.random.default_rng(seed)
    base = synth.random_genes(seed, n_base, radius=spread)
    out = []
    for b in base:
        for _ in range(per_base):
            g = b.copy()
            g[:3] = (g[:3] + rng.normal(0, rot, 3)) % (2 * np.pi)
            g[3:] += rng.normal(0, shift, 3)
            out.append(g)
    out = np.array(out)
    return out[rng.permutation(len(out))]


def _oracle_sets(rec, lig, genes, cutoff=5.0):
    sets = []
    for g in genes:
        moved = orc.apply_genes(lig.atoms, g)
        sets.append(orc.contacts(np.concatenate([rec.atoms, moved]), cutoff))
    return sets


def test_contacts_bit_exact_2oob():
    rec, lig, cx = _complex(read_text("2oob_A.pdb"), read_text("2oob_B.pdb"))
    genes = np.concatenate([docked_genes(1, 4, 8), synth.random_genes(2, 16, 12.0), synth.ga_genes(3, 8)])
    bits, sizes = cx.contacts(genes, 5.0)
    ref = _oracle_sets(rec, lig, genes)
    for i, g in enumerate(genes):
        assert cx.contact_set(bits[i]) == ref[i], i
        assert sizes[i] == len(ref[i])
    assert max(sizes) > 0


@pytest.mark.parametrize("cutoff", [3.5, 5.0, 7.25])
def test_contacts_cutoffs(cutoff):
    c = synth.make_complex(31, rec_res=80, lig_res=60)
    rec, lig, cx = _complex(c.receptor_pdb, c.ligand_pdb)
    genes = synth.random_genes(31, 24, 16.0)
    assert calculate_contacts(cx, genes, cutoff) == _oracle_sets(rec, lig, genes, cutoff)


@pytest.mark.parametrize("seed,nb,per", [(5, 6, 10), (6, 12, 6), (7, 3, 30)])
def test_cluster_structures_exact(seed, nb, per):
    rec, lig, cx = _complex(read_text("2oob_A.pdb"), read_text("2oob_B.pdb"))
    genes = docked_genes(seed, nb, per)
    got = cluster_structures(cx, genes)
    ref, _ = orc.cluster_poses(rec.atoms, lig.atoms, genes)
    assert [(r.center_idx, r.members) for r in got] == ref
    assert all(r.size == len(r.members) for r in got)
    assert len(ref) > 0


def test_cluster_configs_and_edges():
    rec, lig, cx = _complex(read_text("2oob_A.pdb"), read_text("2oob_B.pdb"))
    genes = np.concatenate([docked_genes(9, 5, 8), synth.ga_genes(9, 10)])  # GA spread: many empty sets
    for cfg in [ClusteringConfig(), ClusteringConfig(5.0, 0.3, 0.5, 2), ClusteringConfig(4.0, 0.75, 1.0, 1),
                ClusteringConfig(6.0, 0.6, 0.75, 10)]:
        got = cluster_structures(cx, genes, cfg)
        ref, _ = orc.cluster_poses(rec.atoms, lig.atoms, genes,
                                   (cfg.contact_distance, cfg.fcc_cutoff, cfg.strictness, cfg.min_cluster_size))
        assert [(r.center_idx, r.members) for r in got] == ref, cfg
    assert cluster_structures(cx, np.zeros((0, 6))) == []


def test_cluster_medium_synthetic():
    c = synth.make_complex(2, rec_res=120, lig_res=90)
    rec, lig, cx = _complex(c.receptor_pdb, c.ligand_pdb)
    genes = np.concatenate([docked_genes(4, 20, 12, spread=18.0), synth.random_genes(5, 160, 18.0)])
    got = cluster_structures(cx, genes)
    ref, sizes = orc.cluster_poses(rec.atoms, lig.atoms, genes)
    assert [(r.center_idx, r.members) for r in got] == ref
    bits, gsz = cx.contacts(genes)
    assert np.array_equal(gsz, sizes)

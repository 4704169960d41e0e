"""Precision and size limits of the scoring path against the FP64 oracle.

- Coulomb energy under cancellation: a pose whose terms sum to zero in
  magnitude-weighted terms must still be held to the 1e-3 kcal/mol floor
  (fitness.rs:284-302).
- Receptors larger than one workgroup's LDS table: scored over receptor
  ranges (several workgroups per pose) by the plain path and by batches
  (fitness.rs:86-184 has no size limit).
- The ligand-atoms-per-lane choice is a function of the complex: two fresh
  processes give bit-identical batch results.

Bar: every component within max(1e-5 |ref|, 1e-3 kcal/mol); satisfied
restraint counts exact."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import structure, synth
from gdock_amd.device import Batch, Complex
from helpers import assert_energies_close

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)
K_COULOMB = 332.0636


def _elec_terms(rec, lig_posed):
    """Per-pair Coulomb terms (fitness.rs:284-302) as the reference forms
    them, in FP64: rows are ligand atoms."""
    rx = np.stack([rec["x"], rec["y"], rec["z"]], 1)
    lx = np.stack([lig_posed["x"], lig_posed["y"], lig_posed["z"]], 1)
    D = rx[None, :, :] - lx[:, None, :]
    d2 = (D[..., 0] * D[..., 0] + D[..., 1] * D[..., 1]) + D[..., 2] * D[..., 2]
    d = np.sqrt(d2)
    mask = (d > 1.0) & (d < 15.0)
    return np.where(mask, K_COULOMB * (rec["charge"][None, :] * lig_posed["charge"][:, None]) / (d * d), 0.0)


def test_elec_cancellation_held_to_the_absolute_floor():
    """The 2000 x 1500-atom benchmark complex with the ligand's charges
    rescaled per pose so that the pose's Coulomb terms cancel: the atoms
    whose terms are positive get their charges multiplied by -B/A (A their
    sum, B the rest).  The net energy is then ~0 while the terms sum to
    >= 1e4 kcal/mol in magnitude, so the tolerance is the 1e-3 floor."""
    cfg = synth.CONFIGS["batched_scoring"]
    c = synth.make_complex(cfg["seed"], rec_atoms=cfg["rec_atoms"], lig_atoms=cfg["lig_atoms"], n_restraints=6)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    genes = synth.random_genes(77, 4, 16.0)
    worst = 0.0
    for i, g in enumerate(genes):
        posed = orc.apply_genes(lig.atoms, g)
        t = _elec_terms(rec.atoms, posed)
        per_atom = t.sum(axis=1)
        pos = per_atom > 0
        A, B = per_atom[pos].sum(), per_atom[~pos].sum()
        assert A > 0 and B < 0
        atoms = lig.atoms.copy()
        atoms["charge"] = np.where(pos, atoms["charge"] * (-B / A), atoms["charge"])
        bal = structure.Molecule(atoms)
        t2 = _elec_terms(rec.atoms, orc.apply_genes(bal.atoms, g))
        assert np.abs(t2).sum() >= 1e4, np.abs(t2).sum()
        cx = Complex(rec, bal, c.restraints)
        got, ns = cx.score_poses(g[None])
        ref, rns = orc.score_poses(rec.atoms, bal.atoms, c.restraints, g[None], W, nthreads=16)
        assert abs(ref[0, 1]) <= 50.0, ref[0, 1]
        assert_energies_close(got, ref, f"cancelling pose {i}")
        assert np.array_equal(ns, rns)
        err = abs(got[0, 1] - ref[0, 1])
        worst = max(worst, err)
        print(f"pose {i}: sum |terms| {np.abs(t2).sum():.3e}, net elec {ref[0, 1]:+.3e}, |error| {err:.2e}")
    assert worst <= 1e-3


def _big(seed, n_rec, n_lig, n_restraints=6):
    c = synth.make_complex(seed, rec_atoms=n_rec, lig_atoms=n_lig, n_restraints=n_restraints)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    return rec, lig, c.restraints


def test_receptor_larger_than_one_workgroup():
    """A 20,000-atom receptor (more slots than one workgroup's LDS table):
    the plain path splits every pose over receptor ranges, the batched path
    over at least three; both match the oracle, near and far poses alike,
    and the no-allocation entry point refuses it with a clear error."""
    rec, lig, pairs = _big(21, 20000, 700)
    cx = Complex(rec, lig, pairs)
    genes = np.concatenate([np.zeros((1, 6)), synth.random_genes(22, 3, 10.0), synth.ga_genes(23, 2)])
    got, ns = cx.score_poses(genes)
    ref, rns = orc.score_poses(rec.atoms, lig.atoms, pairs, genes, W, nthreads=16)
    assert_energies_close(got, ref, "20000-atom receptor, plain path")
    assert np.array_equal(ns, rns)
    b = Batch([cx], [len(genes)])
    assert b.plan()["parts"][0] >= 3, b.plan()
    out, bns = b.score(genes)
    assert_energies_close(out, ref, "20000-atom receptor, batched")
    assert np.array_equal(bns, rns)
    out2, _ = b.score(genes)
    assert np.array_equal(out, out2)
    import ctypes as C
    import torch
    from gdock_amd import _lib
    buf = torch.zeros(64, dtype=torch.float64, device="cuda")
    w = _lib.Weights(*W)
    st = _lib.lib().gdock_score_poses_device_noalloc(cx.ctx.handle, cx.handle, C.c_void_p(buf.data_ptr()), 1,
                                                     C.byref(w), C.c_void_p(buf.data_ptr()), None, None)
    assert st == _lib.INVALID_ARG
    assert b"split" in _lib.lib().gdock_last_error(cx.ctx.handle)


def _tiled_receptor(rec, copies, step):
    """copies of the receptor side by side along x, step A apart, serials
    renumbered from 1 (PDB text cannot hold this many atoms)."""
    parts = []
    for c in range(copies):
        a = rec.atoms.copy()
        a["x"] = a["x"] + c * step
        parts.append(a)
    atoms = np.concatenate(parts)
    atoms["serial"] = np.arange(1, len(atoms) + 1, dtype=atoms["serial"].dtype)
    return structure.Molecule(atoms)


def test_maximum_sizes():
    """A receptor of 100,000 atoms (five copies of a 20,000-atom chain 30 A
    apart along x, so a pose sees several; seven receptor ranges per pose,
    the limit is 122,624 atoms) and a ligand of about 6,000 atoms (three
    chunks of eight atoms per lane): the plain path against the oracle on a
    pose at the centre, two random and one GA-initial pose; one receptor atom
    past the limit is refused with a clear error."""
    from gdock_amd import _lib
    base = synth.make_complex(31, rec_atoms=20000, lig_atoms=6000, n_restraints=6)
    rec = _tiled_receptor(structure.read_pdb_text(base.receptor_pdb)[0], 5, 30.0)
    lig0 = structure.read_pdb_text(base.ligand_pdb)[0]
    atoms = lig0.atoms.copy()
    atoms["serial"] = np.arange(len(rec) + 1, len(rec) + 1 + len(atoms), dtype=atoms["serial"].dtype)
    lig = structure.position_ligand(rec, structure.Molecule(atoms))
    pairs = base.restraints
    assert len(rec) == 100000 and len(lig) > 5000, (len(rec), len(lig))
    cx = Complex(rec, lig, pairs)
    genes = np.concatenate([np.zeros((1, 6)), synth.random_genes(32, 2, 40.0), synth.ga_genes(33, 1)])
    got, ns = cx.score_poses(genes)
    ref, rns = orc.score_poses(rec.atoms, lig.atoms, pairs, genes, W, nthreads=16)
    assert_energies_close(got, ref, f"{len(rec)} x {len(lig)} atoms")
    assert np.array_equal(ns, rns)
    assert (np.abs(ref[:, 1]) > 0).all()
    n_over = 122624 + 1 - len(rec)
    too_big = structure.Molecule(np.concatenate([rec.atoms, rec.atoms[:n_over]]))
    with pytest.raises(_lib.GdockError) as e:
        Complex(too_big, lig, pairs)
    assert "122624" in str(e.value)


_CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, {root!r})
from gdock_amd import structure, synth
from gdock_amd.device import Batch, Complex
cs = synth.uneven_batch(6, 300)
cxs, genes = [], []
for i, c in enumerate(cs):
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cxs.append(Complex(rec, lig, c.restraints))
    genes.append(synth.ga_genes(300 + i, 40))
order = {order}
b = Batch([cxs[i] for i in order], [40] * len(order))
out, ns = b.score(np.concatenate([genes[i] for i in order]))
rows = {{str(i): out[40 * k:40 * (k + 1)].tobytes().hex() for k, i in enumerate(order)}}
print(json.dumps({{"rows": rows, "lanes": dict(zip(map(str, order), b.plan()["lanes_k"]))}}))
"""


def test_batch_results_identical_in_two_fresh_processes():
    """Each process builds the batch from scratch (its own lane choices, its
    own complexes, the complexes in another order): every complex's rows are
    bit-identical, since the lanes per complex are a function of its shape."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    runs = []
    for order in ([0, 1, 2, 3, 4, 5], [5, 3, 1, 0, 2, 4]):
        r = subprocess.run([sys.executable, "-c", _CHILD.format(root=root, order=order)], capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    assert runs[0]["lanes"] == runs[1]["lanes"]
    assert runs[0]["rows"] == runs[1]["rows"]

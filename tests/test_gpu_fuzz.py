"""Randomised parity sweep of population scoring against the oracle.

Small random complexes (0-90 receptor and 0-70 ligand atoms) with random
parameters and elements (ASP and radius follow the element; H, P and other
elements carry none), colliding and distinct serial numbers, CA atoms for
restraints, and coordinates either continuous or on a 0.5 A lattice so that
many pairs sit exactly on the 1, 8, 12 and 15 A cutoffs and the 7 A
restraint bound.  Poses: identity, small moves, GA-spread; lattice cases
are also scored as given.  Every component within the scoring tolerance,
satisfied-restraint counts exact."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import structure, synth
from gdock_amd._lib import ATOM_DTYPE
from gdock_amd.device import Complex
from helpers import assert_energies_close

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)
RADIUS = {"H": 1.2, "C": 1.7, "N": 1.55, "O": 1.52}


def _atoms(rng, n, chain, serials, center, lattice):
    a = np.zeros(n, ATOM_DTYPE)
    xyz = rng.uniform(-9.0, 9.0, (n, 3)) + center
    if lattice:
        xyz = np.round(xyz * 2.0) / 2.0
    a["x"], a["y"], a["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    el = rng.choice(["C", "N", "O", "S", "H", "P"], n, p=[0.4, 0.2, 0.2, 0.05, 0.1, 0.05])
    a["element"] = [e.encode() for e in el]
    a["vdw_radius"] = [RADIUS.get(e, 1.0) for e in el]
    a["epsilon"] = -rng.uniform(0.0, 0.3, n) * (rng.random(n) < 0.95)
    a["rmin2"] = rng.uniform(0.0, 2.2, n) * (rng.random(n) < 0.95)
    a["charge"] = np.round(rng.normal(0.0, 0.4, n), 3) * (rng.random(n) < 0.8)
    a["serial"] = serials
    a["resseq"] = 1 + np.arange(n) // 4
    a["name"] = [b"CA" if i % 4 == 1 else b"CB" for i in range(n)]
    a["resname"], a["chainid"], a["altloc"], a["icode"] = b"ALA", chain.encode(), b" ", b" "
    return a


def _case(rng):
    nr, nl = int(rng.integers(0, 90)), int(rng.integers(0, 70))
    lattice = rng.random() < 0.4
    rec = _atoms(rng, nr, "A", 1 + np.arange(nr), np.zeros(3), lattice)
    if rng.random() < 0.3 and nr > 0:  # ligand serials collide with receptor ones (files numbered from 1)
        lser = rng.integers(1, nr + 1, nl)
    else:
        lser = 1000 + np.arange(nl)
    lig = _atoms(rng, nl, "B", lser, np.array([rng.uniform(0, 14), 0.0, 0.0]), lattice)
    rres = sorted(set(rec["resseq"].tolist()))
    lres = sorted(set(lig["resseq"].tolist()))
    pairs = []
    for _ in range(int(rng.integers(0, 6))):
        if rres and lres:
            pairs.append((int(rng.choice(rres)), int(rng.choice(lres))))
    pairs.append((999, 1))  # never resolves: dropped as in restraints.rs:68-73
    return rec, lig, pairs, lattice


@pytest.mark.parametrize("block", range(12))
def test_random_complexes_match_oracle(block):
    rng = np.random.default_rng(4242 + block)
    for case in range(10):
        rec, lig, pairs, lattice = _case(rng)
        R, L = structure.Molecule(rec), structure.Molecule(lig)
        cx = Complex(R, L, pairs)
        genes = np.concatenate([np.zeros((1, 6)), rng.normal(0.0, 0.3, (4, 6)), synth.ga_genes(int(rng.integers(1_000_000)), 3)])
        got, ns = cx.score_poses(genes)
        ref, rns = orc.score_poses(rec, lig, pairs, genes, W)
        what = f"block {block} case {case} ({len(rec)}x{len(lig)}, lattice={lattice})"
        assert_energies_close(got, ref, what)
        assert np.array_equal(ns, rns), what
        if lattice and len(lig):  # as given: pairs exactly on the cutoffs
            xyz = np.stack([lig["x"], lig["y"], lig["z"]], 1)[None]
            g1, n1 = cx.score_coords(xyz)
            r1, m1 = orc.score_poses(rec, lig, pairs, np.zeros((1, 6)), W, transform=False)
            assert_energies_close(g1, r1, what + " as given")
            assert np.array_equal(n1, m1), what

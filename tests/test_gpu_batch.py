"""Batched scoring of uneven complexes (one persistent launch replayed from a
HIP graph) against the FP64 oracle and against the single-complex path.

Bar: every component within max(1e-5 |ref|, 1e-3 kcal/mol); satisfied-restraint
counts exact; graph replays bit-identical to each other."""
import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import constants, structure, synth
from gdock_amd.device import Batch, Complex
from helpers import assert_energies_close, read_text

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)


def _load(c):
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    return rec, lig


@pytest.fixture(scope="module")
def uneven():
    """Six complexes spanning 1-8 ligand atoms per lane and chunked ligands,
    plus the 1h590 pair whose serial numbers collide."""
    specs = [(31, 500, 300), (32, 3000, 2600), (33, 1200, 900), (34, 800, 2100), (35, 4000, 150), (36, 650, 1700)]
    out = []
    for seed, nr, nl in specs:
        c = synth.make_complex(seed, rec_atoms=nr, lig_atoms=nl, n_restraints=6)
        rec, lig = _load(c)
        out.append((rec, lig, c.restraints))
    rec = structure.read_pdb_text(read_text("1h590A_0A.pdb"))[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(read_text("1h590B_0A_poses.pdb.gz"))[0])
    out.append((rec, lig, [(10, 5), (20, 30)]))
    return out


def _genes(i, n):
    return np.concatenate([synth.random_genes(500 + i, n - n // 3, 18.0), synth.ga_genes(600 + i, n // 3)])


def test_batch_matches_oracle_and_replays(uneven):
    counts = [7, 5, 0, 9, 6, 8, 7]
    cxs = [Complex(r, l, p) for r, l, p in uneven]
    b = Batch(cxs, counts)
    plan = b.plan()
    assert all(1 <= k <= 8 for k in plan["lanes_k"]) and plan["grid"] > 0
    genes = np.concatenate([_genes(i, n) if n else np.zeros((0, 6)) for i, n in enumerate(counts)])
    out1, ns1 = b.score(genes)
    # eager single-complex work between two replays of the graph
    for i, cx in enumerate(cxs):
        if counts[i]:
            cx.score_poses(_genes(i, counts[i]))
    out2, ns2 = b.score(genes)
    assert np.array_equal(out1, out2) and np.array_equal(ns1, ns2)
    for i, (r, l, p) in enumerate(uneven):
        a, z = b.offsets[i], b.offsets[i + 1]
        if z == a:
            continue
        ref, rns = orc.score_poses(r.atoms, l.atoms, p, genes[a:z], W, nthreads=8)
        assert_energies_close(out1[a:z], ref, f"complex {i} in batch")
        assert np.array_equal(ns1[a:z], rns)
        single, sns = cxs[i].score_poses(genes[a:z])
        assert_energies_close(out1[a:z], single, f"complex {i}: batch vs single")
        assert np.array_equal(ns1[a:z], sns)


def test_results_do_not_depend_on_the_item_order(uneven):
    """Each launch orders its items by the poses' estimated cost (and within a
    cost bucket in no fixed order): the same pose gives the same bits whatever
    the order, its row or the other poses of the launch."""
    counts = [6, 6, 0, 6, 6, 6, 6]
    b = Batch([Complex(r, l, p) for r, l, p in uneven], counts)
    genes = np.concatenate([_genes(i, n) if n else np.zeros((0, 6)) for i, n in enumerate(counts)])
    out1, ns1 = b.score(genes)
    rev = genes.copy()
    for i in range(len(counts)):
        a, z = b.offsets[i], b.offsets[i + 1]
        rev[a:z] = genes[a:z][::-1]
    out_r, ns_r = b.score(rev)
    other = np.concatenate([_genes(40 + i, n) if n else np.zeros((0, 6)) for i, n in enumerate(counts)])
    b.score(other)
    out3, ns3 = b.score(genes)
    for i in range(len(counts)):
        a, z = b.offsets[i], b.offsets[i + 1]
        assert np.array_equal(out_r[a:z][::-1], out1[a:z]) and np.array_equal(ns_r[a:z][::-1], ns1[a:z]), i
    assert np.array_equal(out3, out1) and np.array_equal(ns3, ns1)


def _lattice(mol):
    """The molecule with coordinates on a 0.5 A lattice: many pairs sit on the
    cutoffs, so the settle buffers fill and overflow."""
    a = mol.atoms.copy()
    for c in "xyz":
        a[c] = np.round(a[c] * 2.0) / 2.0
    return structure.Molecule(a)


def test_split_poses_match_oracle():
    """Expensive poses of large complexes are scored by several workgroups
    over receptor ranges (each launch estimates every pose's cost); the parts'
    sums, the shared ligand burial counts and the overflow rescoring of the
    whole pose agree with the oracle and the one-workgroup path, and repeated
    replays (which reuse the cleared split state) agree."""
    specs = [(51, 6000, 3000, False), (52, 2600, 3300, False), (53, 5200, 2400, True)]
    cases = []
    for seed, nr, nl, lattice in specs:
        c = synth.make_complex(seed, rec_atoms=nr, lig_atoms=nl, n_restraints=5)
        rec, lig = _load(c)
        if lattice:
            rec, lig = _lattice(rec), _lattice(lig)
        cases.append((rec, lig, c.restraints))
    counts = [6, 4, 5]
    cxs = [Complex(r, l, p) for r, l, p in cases]
    b = Batch(cxs, counts)
    plan = b.plan()
    print("split plan:", plan)
    assert all(q > 1 for q in plan["parts"]), plan
    genes = []
    for i, n in enumerate(counts):
        g = np.concatenate([np.zeros((1, 6)), synth.ga_genes(700 + i, n - 1)])
        genes.append(g)
    genes = np.concatenate(genes)
    out1, ns1 = b.score(genes)
    items = b.last_items()
    print("work items:", items, "poses:", len(genes))
    assert items > len(genes)  # the near poses (identity genes) were split, the far ones need not be
    out2, ns2 = b.score(genes)
    assert np.array_equal(out1, out2) and np.array_equal(ns1, ns2)
    for i, (r, l, p) in enumerate(cases):
        a, z = b.offsets[i], b.offsets[i + 1]
        ref, rns = orc.score_poses(r.atoms, l.atoms, p, genes[a:z], W, nthreads=8)
        assert_energies_close(out1[a:z], ref, f"split complex {i}")
        assert np.array_equal(ns1[a:z], rns)
        single, sns = cxs[i].score_poses(genes[a:z])
        assert_energies_close(out1[a:z], single, f"split complex {i}: batch vs single")
        assert np.array_equal(ns1[a:z], sns)


def test_split_pose_overflow_rescored():
    """A receptor and a ligand on one integer lattice, large enough for split
    poses: every part's settle buffer overflows on pairs exactly at the 8, 12
    and 15 A cutoffs, so each pose is rescored whole by the exact kernel.
    Poses: the identity and quarter turns (near-lattice in FP64)."""
    from helpers import mol, test_atom as A
    rec = structure.Molecule(mol(*[A(float(x), float(y), float(z), charge=0.5 if (x + y) % 2 else -0.5,
                                     element="O" if z % 2 else "C", serial=i + 1, resseq=i // 8 + 1)
                                   for i, (x, y, z) in enumerate((x, y, z) for x in range(16) for y in range(16)
                                                                  for z in range(16))]))
    lig = structure.Molecule(mol(*[A(float(x), float(y), float(z) + 17.0, charge=-0.25 if x % 3 else 0.4,
                                     element="N" if y % 4 else "C", serial=9000 + i, chainid="B",
                                     resseq=i // 8 + 1)
                                   for i, (x, y, z) in enumerate((x, y, z) for x in range(16) for y in range(16)
                                                                  for z in range(8))]))
    cx = Complex(rec, lig)
    b = Batch([cx], [4])
    assert b.plan()["parts"][0] > 1, b.plan()
    q = np.pi / 2
    genes = np.array([[0, 0, 0, 0, 0, 0], [q, 0, 0, 0, 0, 0], [0, q, 0, 0, 0, 0], [0, 0, q, 0, 0, 1.0]])
    out, ns = b.score(genes)
    assert b.last_items() > len(genes)  # every pose here sits in the receptor: split
    ref, rns = orc.score_poses(rec.atoms, lig.atoms, [], genes, W, nthreads=8)
    assert_energies_close(out, ref, "split lattice storm")
    single, _ = cx.score_poses(genes)
    assert_energies_close(out, single, "split lattice storm: batch vs single")
    out2, _ = b.score(genes)
    assert np.array_equal(out, out2)


def test_batch_device_launch_and_new_weights(uneven):
    import torch
    counts = [4] * len(uneven)
    cxs = [Complex(r, l, p) for r, l, p in uneven]
    b = Batch(cxs, counts)
    genes = np.concatenate([_genes(i, 4) for i in range(len(uneven))])
    ref, _ = b.score(genes)
    w2 = constants.EnergyWeights(vdw=1.0, elec=0.5, desolv=0.1, air=0.01)
    out2, _ = b.score(genes, w2)
    comp = ref[:, :5]
    assert np.array_equal(out2[:, :5], comp)  # components do not depend on the weights
    fit = 1.0 * comp[:, 0] + 0.5 * comp[:, 1] + 0.1 * comp[:, 2] + 0.01 * comp[:, 3]
    assert np.allclose(out2[:, 5], fit, rtol=1e-12, atol=1e-12)
    # device-resident replays on a torch stream, with eager torch work between
    p, o, _ = b.device_buffers()
    stream = torch.cuda.Stream()
    scratch = torch.ones(1 << 20, dtype=torch.float32, device="cuda")
    for rep in range(3):
        b.launch(None, stream.cuda_stream)
        stream.synchronize()
        assert np.array_equal(_read_device(o, b.n_poses), ref), rep
        scratch.mul_(2.0).sum().item()


def _read_device(ptr_value, n):
    """Copies n x 6 doubles from a device address (synchronous hipMemcpy)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    host = np.zeros((n, 6))
    assert hip.hipMemcpy(host.ctypes.data, ptr_value, host.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return host


def test_empty_batch():
    b = Batch([Complex(*_load(synth.make_complex(40, rec_atoms=300, lig_atoms=200))[:2])], [0])
    out, ns = b.score(np.zeros((0, 6)))
    assert out.shape == (0, 6)

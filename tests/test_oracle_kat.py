"""Pins the FP64 oracle: the reference's own known-answer tests, closed forms,
exact threshold forms, invariants and brute force on tiny inputs.
CPU only."""
import math
import re
from pathlib import Path

import numpy as np
import pytest

import oracle.oracle as orc
from helpers import RESTRAINTS_2OOB, TOPOLOGIES, contacts_brute_force, mol, read_text, relabel, test_atom as A

PI = math.pi


# ------------------------------------------------------------ fitness.rs:357-607
def test_vdw_far_atoms_zero():  # fitness.rs:358-367
    assert orc.vdw_energy(mol(A(0, 0, 0)), mol(A(20, 0, 0))) == 0.0


def test_vdw_close_and_very_close():  # fitness.rs:370-393
    assert math.isfinite(orc.vdw_energy(mol(A(0, 0, 0)), mol(A(5, 0, 0, serial=2))))
    e = orc.vdw_energy(mol(A(0, 0, 0)), mol(A(1, 0, 0, serial=2)))
    assert math.isfinite(e) and e <= 500.0


@pytest.mark.parametrize("d,ref", [(5.0, -0.04098164762097422), (3.0, -0.07744006564546031),
                                   (1.0, -0.0001951219047441999)])
def test_vdw_closed_form(d, ref):
    # eps_ij = sqrt(0.1*0.1), R = 4: 0.1 * (rho^2 - 2 rho), rho = 4^6 / (d^6 + 0.5 * 4^6)
    assert orc.vdw_energy(mol(A(0, 0, 0)), mol(A(d, 0, 0, serial=2))) == pytest.approx(ref, rel=1e-15)


def test_vdw_serial_skip():  # fitness.rs:92: equal serials never interact
    assert orc.vdw_energy(mol(A(0, 0, 0, serial=7)), mol(A(4, 0, 0, serial=7))) == 0.0
    assert orc.vdw_energy(mol(A(0, 0, 0, serial=7)), mol(A(4, 0, 0, serial=8))) < 0.0


def test_elec_signs_and_closed_form():  # fitness.rs:396-436
    assert orc.elec_energy(mol(A(0, 0, 0)), mol(A(5, 0, 0))) == 0.0
    assert orc.elec_energy(mol(A(0, 0, 0, charge=1)), mol(A(5, 0, 0, charge=-1))) == pytest.approx(-13.282544, rel=1e-15)
    assert orc.elec_energy(mol(A(0, 0, 0, charge=1)), mol(A(5, 0, 0, charge=1))) > 0.0
    # outside (1, 15): no interaction at 15 exactly, nor at 1 exactly
    assert orc.elec_energy(mol(A(0, 0, 0, charge=1)), mol(A(15, 0, 0, charge=1))) == 0.0
    assert orc.elec_energy(mol(A(0, 0, 0, charge=1)), mol(A(1, 0, 0, charge=1))) == 0.0
    assert orc.elec_energy(mol(A(0, 0, 0, charge=1)), mol(A(14.99, 0, 0, charge=1))) > 0.0


def _ring(n, r=3.0, el="C"):
    return mol(*[A(r * math.cos(i), r * math.sin(i), 0.0, element=el, serial=i + 2) for i in range(n)])


def test_desolv_signs_and_closed_form():  # fitness.rs:439-486
    e_c = orc.desolv_energy(mol(A(0, 0, 0, element="C")), _ring(15))
    e_o = orc.desolv_energy(mol(A(0, 0, 0, element="O")), _ring(15))
    assert e_c < 0.0 and e_o > 0.0
    # receptor atom fully buried (15 >= 10 partners), each ligand atom sees 1 partner
    full = 0.012 * 1.0 * 1.7 * 1.7 * 4.0 * PI
    part = 0.012 * 0.1 * 1.7 * 1.7 * 4.0 * PI
    assert e_c == pytest.approx(-(full + 15 * part), rel=1e-14)
    # burial saturates: 10 and 15 partners give the same receptor term
    assert orc.desolv_energy(mol(A(0, 0, 0, element="C")), _ring(10, el="H")) == pytest.approx(-full, rel=1e-14)


def test_desolv_sulfur_radius_one():  # structure.rs:250-256: S gets radius 1.0 from the parser
    text = ("ATOM      1  SG  CYS A   1       0.000   0.000   0.000  1.00  0.00           S\n")
    s = orc.read_pdb_text(text)[0]
    assert s["vdw_radius"][0] == 1.0


def test_air_and_satisfaction():  # fitness.rs:489-607
    r1, l1 = A(0, 0, 0, resseq=1), A(5, 0, 0, resseq=10, chainid="B")
    assert orc.air_energy([(1, 10)], mol(r1), mol(l1)) == 0.0
    l_far = A(15, 0, 0, resseq=10, chainid="B")
    assert orc.air_energy([(1, 10)], mol(r1), mol(l_far)) == 640.0  # 10 * (15 - 7)^2
    assert orc.satisfaction_ratio([(1, 10)], mol(r1), mol(l1)) == 1.0
    assert orc.satisfaction_ratio([(1, 10)], mol(r1), mol(l_far)) == 0.0
    assert orc.satisfaction_ratio([], mol(), mol()) == 0.0
    rec = mol(A(0, 0, 0, resseq=1), A(0, 0, 0, resseq=2))
    lig = mol(A(5, 0, 0, resseq=10), A(15, 0, 0, resseq=11))
    assert orc.satisfaction_ratio([(1, 10), (2, 11)], rec, lig) == 0.5


def test_restraint_boundary():  # restraints.rs:173-203: 7.0 inclusive, 10 not satisfied
    assert orc.satisfaction_ratio([(1, 10)], mol(A(0, 0, 0)), mol(A(7.0, 0, 0, resseq=10))) == 1.0
    assert orc.satisfaction_ratio([(1, 10)], mol(A(0, 0, 0)), mol(A(10.0, 0, 0, resseq=10))) == 0.0
    assert orc.air_energy([(1, 10)], mol(A(0, 0, 0)), mol(A(7.0, 0, 0, resseq=10))) == 0.0


def test_restraints_unresolved_dropped():  # restraints.rs:64-73
    assert orc.restraints_resolved([(1, 10), (99, 10)], mol(A(0, 0, 0)), mol(A(5, 0, 0, resseq=10))) == 1
    assert orc.restraints_resolved([(1, 10)], mol(A(0, 0, 0, name="CB")), mol(A(5, 0, 0, resseq=10))) == 0


# ------------------------------------------------------------ exact d2 thresholds
def _next(x, k=1):
    for _ in range(abs(k)):
        x = math.nextafter(x, math.inf if k > 0 else -math.inf)
    return x


def _common_h():
    src = (Path(__file__).resolve().parent.parent / "gdock_amd/csrc/common.h").read_text()
    return {m.group(1): float(m.group(2)) for m in re.finditer(r"constexpr double (kD2\w+) = ([0-9.e+-]+);", src)}


@pytest.mark.parametrize("name,cut,pred", [
    ("kD2ElecMinExcl", 1.0, lambda d, c: d > c),      # d2 > T  <=>  dist > 1
    ("kD2DesolvIncl", 8.0, lambda d, c: d < c),       # d2 <= T <=>  dist < 8
    ("kD2VdwIncl", 12.0, lambda d, c: d < c),
    ("kD2ElecMaxIncl", 15.0, lambda d, c: d < c),
    ("kD2RestraintIncl", 7.0, lambda d, c: d <= c),
    ("kD2ContactIncl", 5.0, lambda d, c: d <= c),
])
def test_squared_distance_thresholds(name, cut, pred):
    """The FP64 d2 forms used on the device equal the reference's sqrt-based
    predicates for every double within 4096 ulps of the cutoff (math.sqrt is
    correctly rounded, as Rust's f64::sqrt)."""
    T = _common_h()[name]
    x = _next(cut * cut, -4096)
    for _ in range(8193):
        ref = pred(math.sqrt(x), cut)
        got = (x > T) if name == "kD2ElecMinExcl" else (x <= T)
        assert ref == got, (name, x)
        x = _next(x)


# ------------------------------------------------------------ structure.rs
def test_center_of_mass_and_rotation_identity():  # structure.rs:373-460
    m = mol(A(0, 0, 0), A(1, 1, 1))
    assert orc.center_of_mass(m) == pytest.approx((0.5, 0.5, 0.5), abs=1e-10)
    r = orc.apply_genes(mol(A(1, 0, 0), A(-1, 0, 0)), [0, 0, 0, 0, 0, 0])
    assert r["x"][0] == pytest.approx(1.0, abs=1e-10) and abs(r["y"][0]) < 1e-10
    r = orc.apply_genes(mol(A(1, 0, 0), A(-1, 0, 0)), [0, 0, PI / 2, 0, 0, 0])
    assert np.allclose(orc.center_of_mass(r), 0.0, atol=1e-10)


def test_apply_genes_translation_only():  # chromosome.rs:256-279
    r = orc.apply_genes(mol(A(0, 0, 0)), [0, 0, 0, 5, 10, 15])
    assert (r["x"][0], r["y"][0], r["z"][0]) == pytest.approx((5, 10, 15), abs=1e-6)


def test_rotation_is_rigid_and_keeps_com():
    rng = np.random.default_rng(3)
    lig = mol(*[A(*rng.normal(size=3) * 5, serial=i) for i in range(30)])
    g = [1.1, 2.3, 4.9, 0, 0, 0]
    r = orc.apply_genes(lig, g)
    p0 = np.stack([lig["x"], lig["y"], lig["z"]], 1)
    p1 = np.stack([r["x"], r["y"], r["z"]], 1)
    d0 = np.linalg.norm(p0[:, None] - p0[None], axis=-1)
    d1 = np.linalg.norm(p1[:, None] - p1[None], axis=-1)
    assert np.allclose(d0, d1, atol=1e-12)
    assert np.allclose(orc.center_of_mass(lig), orc.center_of_mass(r), atol=1e-12)
    # and the matrix is Rz(a) Ry(b) Rx(g)
    a, b, c = g[:3]
    Rz = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    Ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    Rx = np.array([[1, 0, 0], [0, math.cos(c), -math.sin(c)], [0, math.sin(c), math.cos(c)]])
    com = np.array(orc.center_of_mass(lig))
    assert np.allclose((p0 - com) @ (Rz @ Ry @ Rx).T + com, p1, atol=1e-12)


def test_read_pdb_models_and_inference():  # structure.rs:521-617
    assert len(orc.read_pdb_text(read_text("1h590A_0A.pdb"))) == 1
    models = orc.read_pdb_text(read_text("1h590B_0A_poses.pdb.gz"))
    assert len(models) == 100 and all(len(m) == 341 for m in models)
    m0 = models[0]
    # no element column: element = first character of the atom name
    assert m0["element"][0] == b"N" and m0["element"][1] == b"C"


def test_toppar_lookups():  # toppar.rs:665-880 through the parser
    text = "\n".join([
        "ATOM      1  N   ALA A   1       0.000   0.000   0.000  1.00  0.00           N",
        "ATOM      2  CA  GLY A   2       0.000   0.000   0.000  1.00  0.00           C",
        "ATOM      3  CD  PRO A   3       0.000   0.000   0.000  1.00  0.00           C",
        "ATOM      4  NZ  LYS A   4       0.000   0.000   0.000  1.00  0.00           N",
        "ATOM      5  CD1 ILE A   5       0.000   0.000   0.000  1.00  0.00           C",
        "ATOM      6  OXT ALA A   6       0.000   0.000   0.000  1.00  0.00           O",
        "ATOM      7  SG  CYS A   7       0.000   0.000   0.000  1.00  0.00           S",
    ])
    m = orc.read_pdb_text(text)[0]
    assert (m["epsilon"][0], m["rmin2"][0], m["charge"][0]) == (-0.17, 1.8240008, -0.5)   # N238
    assert (m["epsilon"][1], m["charge"][1]) == (-0.066, 0.08)                          # C223
    assert (m["epsilon"][2], m["charge"][2]) == (-0.066, -0.05)                         # C245
    assert (m["charge"][3]) == -0.3                                                     # N287
    assert (m["epsilon"][4], m["rmin2"][4], m["charge"][4]) == (0.0, 0.0, 0.0)          # ILE CD1: unparameterized
    assert (m["epsilon"][5], m["charge"][5]) == (0.0, 0.0)                              # OXT
    assert (m["epsilon"][6], m["rmin2"][6], m["charge"][6]) == (-0.425, 2.0204317, -0.335)  # S200


# ------------------------------------------------------------ clustering.rs
def test_fcc_known_answers():  # clustering.rs:308-355
    a = {"A 1 B 2", "A 3 B 4"}
    assert orc.calculate_fcc(a, a) == (1.0, 1.0)
    assert orc.calculate_fcc({"A 1 B 2"}, {"A 5 B 6"}) == (0.0, 0.0)
    assert orc.calculate_fcc(a, {"A 1 B 2", "A 5 B 6"}) == (0.5, 0.5)
    assert orc.calculate_fcc(set(), set()) == (0.0, 0.0)


def test_contacts_real_data():  # clustering.rs:365-392
    rec = orc.read_pdb_text(read_text("2oob_A.pdb"))[0]
    lig = orc.read_pdb_text(read_text("2oob_B.pdb"))[0]
    c = orc.contacts(np.concatenate([rec, lig]), 5.0)
    assert c
    for s in c:
        p = s.split()
        assert len(p) == 4 and p[0] != p[2] and p[0] < p[2]


def test_contacts_brute_force_and_monotone():
    """Contacts equal an all-atom-pair transcription; they grow with the cutoff."""
    rec = orc.read_pdb_text(read_text("2oob_A.pdb"))[0]
    lig = orc.read_pdb_text(read_text("2oob_B.pdb"))[0]
    cplx = np.concatenate([rec, lig])
    heavy = cplx[cplx["element"] != b"H"]
    xyz = np.stack([heavy["x"], heavy["y"], heavy["z"]], 1)
    d = np.sqrt(((xyz[:, None] - xyz[None]) ** 2).sum(-1))
    ch = heavy["chainid"]
    ii, jj = np.nonzero((d <= 5.0) & (ch[:, None] != ch[None]))
    brute = set()
    for i, j in zip(ii, jj):
        a, b = (i, j) if ch[i] < ch[j] else (j, i)
        brute.add(f"{heavy['chainid'][a].decode()} {heavy['resseq'][a]} {heavy['chainid'][b].decode()} {heavy['resseq'][b]}")
    assert orc.contacts(cplx, 5.0) == brute
    assert orc.contacts(cplx, 4.0) <= orc.contacts(cplx, 5.0) <= orc.contacts(cplx, 6.0)


@pytest.mark.parametrize("name", sorted(TOPOLOGIES))
def test_contacts_chain_layouts_brute_force(name):
    """Every chain layout (two receptor chains, two ligand chains, one chain ID
    for both, residues shared by receptor and ligand): the oracle's contacts
    equal the all-atom-pair transcription at two cutoffs."""
    rr, lr = TOPOLOGIES[name]
    rec = relabel(orc.read_pdb_text(read_text("2oob_A.pdb"))[0], rr)
    lig = relabel(orc.read_pdb_text(read_text("2oob_B.pdb"))[0], lr)
    for genes in (np.zeros(6), np.array([0.3, 1.1, 2.0, 4.0, -3.0, 1.5])):
        cplx = np.concatenate([rec, orc.apply_genes(lig, genes)])
        for cutoff in (3.5, 5.0):
            got = orc.contacts(cplx, cutoff)
            assert got == contacts_brute_force(cplx, cutoff), (name, cutoff)
            if name == "shared_chain_id":
                assert got == set()
            else:
                assert got


def _greedy_transcription(nb, min_size):
    """clustering.rs:183-254 written over Python sets."""
    n = len(nb)
    nbrs = {i: {j for j in range(n) if nb[i][j]} for i in range(n)}
    used, out = set(), []
    while True:
        cand = [k for k in nbrs if k not in used]
        if not cand:
            break
        center = max(cand, key=lambda a: (len(nbrs[a] - used), -a))
        members = sorted({center} | (nbrs[center] - used))
        used |= set(members)
        out.append((center, members))
        del nbrs[center]
    out.sort(key=lambda c: -len(c[1]))  # stable
    return [c for c in out if len(c[1]) >= min_size]


@pytest.mark.parametrize("seed", range(40))
def test_greedy_clustering_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 13))
    nb = (rng.random((n, n)) < rng.uniform(0.1, 0.7)).astype(np.uint8)
    np.fill_diagonal(nb, rng.random(n) < 0.9)
    for min_size in (1, 2, 4):
        assert orc.cluster_elements(nb, min_size) == _greedy_transcription(nb, min_size)


def test_cluster_structures_empty():  # clustering.rs:357-363
    rec = orc.read_pdb_text(read_text("2oob_A.pdb"))[0]
    lig = orc.read_pdb_text(read_text("2oob_B.pdb"))[0]
    assert orc.cluster_poses(rec, lig, np.zeros((0, 6)))[0] == []


# ------------------------------------------------------------ run.rs HoF, population.rs
def test_hall_of_fame():  # run.rs:914-1033
    h = orc.Hof()
    assert h.try_add([0] * 6, -100.0, 1)
    assert not h.try_add([0] * 6, -100.0, 2)
    assert h.try_add([PI, PI, PI, 0, 0, 0], -90.0, 2)
    assert h.try_add([0, 0, 0, 10, 10, 10], -90.0, 2)
    assert not h.try_add([0.1, 0.1, 0.1, 0.5, 0.5, 0.5], -90.0, 3)
    h3 = orc.Hof(3)
    for i in range(5):
        h3.try_add([i * 1.0, 0, 0, i * 10.0, 0, 0], -(100.0 - i * 10.0), i)
    f = h3.fitness()
    assert len(f) == 3 and f[0] <= f[1] <= f[2]
    assert orc.angular_difference(0.0, 2 * PI - 0.1) < 0.2
    assert orc.angular_difference(0.0, PI) == pytest.approx(PI)


def test_evolve_preserves_elites_and_size():  # population.rs:469-530
    rng = np.random.default_rng(0)
    genes = np.concatenate([rng.random((20, 3)) * 2 * PI, rng.random((20, 3)) * 100 - 50], 1)
    fit = np.arange(20) * 5.0
    g2, f2 = orc.evolve(genes, fit, 42)
    assert len(f2) == 20
    for e in fit[:5]:
        assert np.any(np.abs(f2 - e) < 1e-10)
    # elites keep their genes too
    for k in range(5):
        assert any(np.array_equal(genes[k], row) for row in g2)
    assert np.all((g2[:, :3] >= 0) & (g2[:, :3] <= 2 * PI))
    assert np.all(np.abs(g2[:, 3:]) <= 50.0)


def test_rigid_motion_invariance():
    """Energies of a real complex do not change when both partners move rigidly."""
    rec = orc.read_pdb_text(read_text("2oob_A.pdb"))[0]
    lig = orc.read_pdb_text(read_text("2oob_B.pdb"))[0]
    g = [0.7, 1.9, 2.6, 3.0, -4.0, 5.0]
    e0 = [orc.vdw_energy(rec, lig), orc.elec_energy(rec, lig), orc.desolv_energy(rec, lig),
          orc.air_energy(RESTRAINTS_2OOB, rec, lig)]
    # move both by the same rotation about the same centre: combine, transform, split
    both = np.concatenate([rec, lig])
    moved = orc.apply_genes(both, g)
    r2, l2 = moved[:len(rec)], moved[len(rec):]
    e1 = [orc.vdw_energy(r2, l2), orc.elec_energy(r2, l2), orc.desolv_energy(r2, l2),
          orc.air_energy(RESTRAINTS_2OOB, r2, l2)]
    assert np.allclose(e0, e1, rtol=1e-9, atol=1e-9)
    assert e0[0] < 0 and e0[3] > 0.0  # 2oob_A/B are given apart: restraints violated


# ---- evaluator.rs:319-490 (oracle restatement of the DockQ evaluator) ----------
def _ev_atom(x, resseq, chain):
    return A(x, 0.0, 0.0, resseq=resseq, chainid=chain, name="CA", element="C")


def test_evaluator_contacts_within_and_beyond():
    assert orc.eval_contacts(mol(_ev_atom(0, 1, "A")), mol(_ev_atom(3, 10, "B"))) == 1
    assert orc.eval_contacts(mol(_ev_atom(0, 1, "A")), mol(_ev_atom(10, 10, "B"))) == 0
    assert orc.eval_contacts(mol(_ev_atom(0, 1, "A")), mol(_ev_atom(5.0, 10, "B"))) == 1  # <= 5.0 inclusive


def test_evaluator_interface_and_creation():
    rec = mol(A(0, 0, 0, resseq=1), A(0, 0, 0, resseq=2, serial=2), A(20, 0, 0, resseq=3, serial=3))
    ref = mol(_ev_atom(3, 10, "B"))
    ev = orc.Evaluator(rec, ref)
    assert ev.receptor_interface == {1, 2} and ev.reference_interface == {10}
    ev1 = orc.Evaluator(mol(_ev_atom(0, 1, "A")), ref)
    assert len(ev1.receptor_interface) == 1 and len(ev1.reference_interface) == 1


def test_evaluator_rmsd_fnat_dockq_known_answers():
    rec, ref = mol(_ev_atom(0, 1, "A")), mol(_ev_atom(3, 10, "B"))
    ev = orc.Evaluator(rec, ref)
    same = ev.calc_metrics([ref])[0]
    assert same[0] == 0.0 and same[1] == 0.0 and same[2] == 1.0
    assert abs(same[3] - 1.0) < 1e-10
    # one atom 3 A from the reference: sqrt(sum of distances / n) = sqrt(3), evaluator.rs:393-410
    moved = ev.calc_metrics([mol(_ev_atom(6, 10, "B"))])[0]
    assert abs(moved[0] - math.sqrt(3.0)) < 1e-10
    far = ev.calc_metrics([mol(_ev_atom(20, 10, "B"))])[0]
    assert far[2] == 0.0
    # DockQ from its closed form (evaluator.rs:73-78)
    r, ir, f = far[0], far[1], far[2]
    assert far[3] == (f + 1.0 / (1.0 + (ir / 1.5) ** 2) + 1.0 / (1.0 + (r / 8.5) ** 2)) / 3.0


def test_evaluator_fnat_counts_atom_contacts_against_a_set_model():
    """fnat on a small multi-atom case checked against an independent set model
    of the reference's Vec semantics (each docked atom pair counts when its
    (chain, res, name) pair occurs among the native pairs)."""
    rng = np.random.default_rng(5)
    names = ["N", "CA", "C", "O", "CB"]
    rec = mol(*[A(*rng.uniform(-6, 6, 3), resseq=1 + i // 5, name=names[i % 5], serial=i + 1) for i in range(30)])
    ref = mol(*[A(*(rng.uniform(-6, 6, 3) + [4, 0, 0]), resseq=20 + i // 5, chainid="B", name=names[i % 5],
                  serial=100 + i) for i in range(25)])
    ev = orc.Evaluator(rec, ref)

    def keyed_pairs(r, l, iface_r, iface_l):
        out = []
        for a in r:
            if int(a["resseq"]) not in iface_r:
                continue
            for b in l:
                if int(b["resseq"]) not in iface_l:
                    continue
                d = math.sqrt((a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 + (a["z"] - b["z"]) ** 2)
                if d <= 5.0:
                    out.append((a["chainid"], int(a["resseq"]), a["name"], b["chainid"], int(b["resseq"]), b["name"]))
        return out

    native = keyed_pairs(rec, ref, ev.receptor_interface, ev.reference_interface)
    assert len(native) == ev.n_native > 0
    for k in range(6):
        model = ref.copy()
        model["x"] += rng.normal(0, 1.5, len(model))
        model["y"] += rng.normal(0, 1.5, len(model))
        docked = keyed_pairs(rec, model, ev.receptor_interface, ev.reference_interface)
        exp = sum(1 for c in docked if c in set(native)) / len(native)
        assert ev.calc_metrics([model])[0][2] == exp, k

"""gdock_score_poses_device_noalloc: the device-resident scoring entry point
that allocates nothing (its per-pose scratch is a static device array).

Bar: bit-identical to gdock_score_poses_device on the same inputs (same
kernels, same summation order), on two calls in a row with the second one
smaller, including poses that take the exact rescoring path; free device
memory unchanged across the calls; sizes beyond the scratch rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import _lib, structure, synth
from gdock_amd.device import Complex, default_context
from helpers import assert_energies_close, mol, test_atom as A

pytestmark = pytest.mark.gpu
W = (0.4, 0.05, 3.4, 1.0)


def _hip():
    import torch  # noqa: F401  (the runtime torch loaded)
    h = C.CDLL("libamdhip64.so")
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    h.hipMemGetInfo.restype = C.c_int
    return h


def _free_bytes(h):
    free, total = C.c_size_t(), C.c_size_t()
    assert h.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _poses(genes):
    p = np.zeros((len(genes), 16))
    _lib.check(_lib.lib().gdock_prepare_poses(_lib.ptr(np.ascontiguousarray(genes)), len(genes), _lib.ptr(p)),
               what="prepare_poses")
    return p


def _score(entry, cx, d_poses, n, d_out, d_ns, stream):
    import torch
    L = _lib.lib()
    w = _lib.Weights(*W)
    d_out.fill_(np.nan)
    d_ns.fill_(-1)
    st = getattr(L, entry)(cx.ctx.handle, cx.handle, C.c_void_p(d_poses.data_ptr()), n, C.byref(w),
                           C.c_void_p(d_out.data_ptr()), C.c_void_p(d_ns.data_ptr()), C.c_void_p(stream.cuda_stream))
    _lib.check(st, cx.ctx.handle, entry)
    stream.synchronize()
    torch.cuda.synchronize()
    return d_out[:n].cpu().numpy().copy(), d_ns[:n].cpu().numpy().copy()


def _lattice_complex():
    """Receptor and ligand on one integer lattice: pairs exactly on the 8, 12
    and 15 A cutoffs overflow the settle buffers, so poses go through the
    exact kernel, which reads the per-pose flags from the scratch."""
    rec = structure.Molecule(mol(*[A(float(x), float(y), float(z), charge=0.5 if (x + y) % 2 else -0.5,
                                     element="O" if z % 2 else "C", serial=i + 1, resseq=i // 8 + 1)
                                   for i, (x, y, z) in enumerate((x, y, z) for x in range(16) for y in range(16)
                                                                  for z in range(12))]))
    lig = structure.Molecule(mol(*[A(float(x), float(y), float(z) + 13.0, charge=-0.25 if x % 3 else 0.4,
                                     element="N" if y % 4 else "C", serial=9000 + i, chainid="B",
                                     resseq=i // 8 + 1)
                                   for i, (x, y, z) in enumerate((x, y, z) for x in range(12) for y in range(12)
                                                                  for z in range(6))]))
    return rec, lig


def test_noalloc_equals_device_entry_twice_second_smaller():
    import torch
    dev = torch.device("cuda", 0)
    c = synth.make_complex(7, rec_atoms=1800, lig_atoms=1100, n_restraints=6)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints, default_context(0))
    lrec, llig = _lattice_complex()
    lcx = Complex(lrec, llig, (), default_context(0))
    q = np.pi / 2
    lattice_genes = np.array([[0, 0, 0, 0, 0, 0], [q, 0, 0, 0, 0, 0], [0, q, 0, 0, 0, 0], [0, 0, q, 0, 0, 1.0]])

    # 1500 poses: more than one round of workgroups, so launched in cost order
    # with the order in the scratch; then 90, in pose order
    g1 = np.concatenate([synth.random_genes(11, 1400, 16.0), synth.ga_genes(12, 100)])
    g2 = synth.random_genes(13, 90, 12.0)
    stream = torch.cuda.Stream(device=dev)
    d_out = torch.zeros((1600, 6), dtype=torch.float64, device=dev)
    d_ns = torch.zeros(1600, dtype=torch.int32, device=dev)
    d_p1 = torch.from_numpy(_poses(g1)).to(dev)
    d_p2 = torch.from_numpy(_poses(g2)).to(dev)
    d_pl = torch.from_numpy(_poses(lattice_genes)).to(dev)
    torch.cuda.synchronize()

    # the existing entry point first: results to compare against, and every
    # kernel of the path loaded before free memory is measured
    ref1 = _score("gdock_score_poses_device", cx, d_p1, len(g1), d_out, d_ns, stream)
    ref2 = _score("gdock_score_poses_device", cx, d_p2, len(g2), d_out, d_ns, stream)
    refl = _score("gdock_score_poses_device", lcx, d_pl, len(lattice_genes), d_out, d_ns, stream)

    hip = _hip()
    free0 = _free_bytes(hip)
    got1 = _score("gdock_score_poses_device_noalloc", cx, d_p1, len(g1), d_out, d_ns, stream)
    got2 = _score("gdock_score_poses_device_noalloc", cx, d_p2, len(g2), d_out, d_ns, stream)
    gotl = _score("gdock_score_poses_device_noalloc", lcx, d_pl, len(lattice_genes), d_out, d_ns, stream)
    free1 = _free_bytes(hip)
    assert free1 == free0, (free0, free1)

    for got, ref, what in ((got1, ref1, "first call (1500 poses)"), (got2, ref2, "second call (90 poses)"),
                           (gotl, refl, "lattice poses (exact path)")):
        assert np.array_equal(got[0], ref[0]), what
        assert np.array_equal(got[1], ref[1]), what
    # and both within the scoring tolerance of the FP64 oracle on a sample
    oref, ons = orc.score_poses(rec.atoms, lig.atoms, c.restraints, g2[:12], W, nthreads=8)
    assert_energies_close(got2[0][:12], oref, "noalloc vs oracle")
    assert np.array_equal(got2[1][:12], ons)
    lref, _ = orc.score_poses(lrec.atoms, llig.atoms, [], lattice_genes, W, nthreads=8)
    assert_energies_close(gotl[0], lref, "noalloc lattice vs oracle")


def test_noalloc_rejects_more_poses_than_its_scratch():
    import torch
    c = synth.make_complex(8, rec_atoms=300, lig_atoms=200, n_restraints=2)
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints, default_context(0))
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    w = _lib.Weights(*W)
    L = _lib.lib()
    st = L.gdock_score_poses_device_noalloc(cx.ctx.handle, cx.handle, C.c_void_p(buf.data_ptr()), 65537,
                                            C.byref(w), C.c_void_p(buf.data_ptr()), None, None)
    assert st == _lib.INVALID_ARG
    assert b"65536" in L.gdock_last_error(cx.ctx.handle)
    assert L.gdock_score_poses_device_noalloc(cx.ctx.handle, cx.handle, None, 0, C.byref(w), None, None,
                                              None) == _lib.OK

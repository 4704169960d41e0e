"""Device context and device-resident complexes (gdock_ctx / gdock_complex)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import constants
from ._lib import (ChromosomeC, ClusterResultC, ClusteringConfigC, GdockError, Weights, check, lib, ptr)
from .structure import Molecule

_contexts: dict[int, "Context"] = {}


class Context:
    """One device, one stream, reusable staging memory."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().gdock_ctx_create(device, C.byref(h))
        check(st, what=f"gdock_ctx_create(device={device})")
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            lib().gdock_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context(device: int = 0) -> Context:
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


def weights_c(w: constants.EnergyWeights | None) -> Weights:
    w = w or constants.EnergyWeights()
    return Weights(w.vdw, w.elec, w.desolv, w.air)


class Complex:
    """Receptor + positioned ligand + resolved restraints, resident on the device."""

    def __init__(self, receptor: Molecule, ligand: Molecule, restraint_pairs=(), ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.receptor, self.ligand = receptor, ligand
        pairs = np.ascontiguousarray(np.array(list(restraint_pairs), dtype=np.int32).reshape(-1, 2))
        self._pairs = pairs
        h = C.c_void_p()
        st = lib().gdock_complex_create(self.ctx.handle, receptor.atoms.ctypes.data, len(receptor),
                                        ligand.atoms.ctypes.data, len(ligand), ptr(pairs, C.c_int32),
                                        len(pairs), C.byref(h))
        check(st, self.ctx.handle, "gdock_complex_create")
        self.handle = h

    @property
    def n_restraints(self) -> int:
        return lib().gdock_complex_n_restraints(self.handle)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().gdock_complex_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scoring -----------------------------------------------------------
    def score_poses(self, genes: np.ndarray, weights: constants.EnergyWeights | None = None):
        """genes (n, 6) -> (out (n, 6) = vdw, elec, desolv, air, restraint_penalty, fitness; n_satisfied)."""
        g = np.ascontiguousarray(genes, dtype=np.float64).reshape(-1, 6)
        out = np.zeros((len(g), 6))
        ns = np.zeros(len(g), np.int32)
        w = weights_c(weights)
        check(lib().gdock_score_poses(self.ctx.handle, self.handle, ptr(g), len(g), C.byref(w), ptr(out),
                                      ptr(ns, C.c_int32)), self.ctx.handle, "gdock_score_poses")
        return out, ns

    def count_pairs(self, genes: np.ndarray) -> tuple[int, int]:
        """(lane-pairs the scoring kernel steps through after culling, those within 15 A)."""
        g = np.ascontiguousarray(genes, dtype=np.float64).reshape(-1, 6)
        c = np.zeros(2, np.int64)
        check(lib().gdock_count_pairs(self.ctx.handle, self.handle, ptr(g), len(g), ptr(c, C.c_int64)),
              self.ctx.handle, "gdock_count_pairs")
        return int(c[0]), int(c[1])

    def score_coords(self, lig_xyz: np.ndarray, weights: constants.EnergyWeights | None = None):
        """Ligand coordinates (n, n_lig, 3) scored as given (commands/score.rs)."""
        x = np.ascontiguousarray(lig_xyz, dtype=np.float64).reshape(-1, len(self.ligand), 3)
        out = np.zeros((len(x), 6))
        ns = np.zeros(len(x), np.int32)
        w = weights_c(weights)
        check(lib().gdock_score_coords(self.ctx.handle, self.handle, ptr(x), len(x), C.byref(w), ptr(out),
                                       ptr(ns, C.c_int32)), self.ctx.handle, "gdock_score_coords")
        return out, ns

    def eval_fitness(self, chromosomes, weights: constants.EnergyWeights | None = None) -> None:
        arr = (ChromosomeC * len(chromosomes))()
        for i, c in enumerate(chromosomes):
            arr[i].genes[:] = list(c.genes)
        w = weights_c(weights)
        check(lib().gdock_eval_fitness(self.ctx.handle, self.handle, arr, len(chromosomes), C.byref(w)),
              self.ctx.handle, "gdock_eval_fitness")
        for i, c in enumerate(chromosomes):
            a = arr[i]
            c.fitness, c.vdw, c.elec, c.desolv, c.air = a.fitness, a.vdw, a.elec, a.desolv, a.air
            c.restraint_penalty = a.restraint_penalty

    # -- FCC -----------------------------------------------------------------
    def contact_words(self, contact_distance: float = 5.0) -> int:
        """uint64 words of a pose's contact bitset at this cutoff."""
        w = lib().gdock_contact_words(self.handle, contact_distance)
        if w < 0:
            raise GdockError(f"contacts unsupported at cutoff {contact_distance} for this complex")
        return w

    def contact_pair(self, bit: int, contact_distance: float = 5.0) -> tuple[str, int, str, int]:
        ca, cb = C.create_string_buffer(1), C.create_string_buffer(1)
        ra, rb = C.c_int16(), C.c_int16()
        check(lib().gdock_contact_pair(self.handle, contact_distance, bit, ca, C.byref(ra), cb, C.byref(rb)),
              what="contact_pair")
        return ca.raw.decode(), ra.value, cb.raw.decode(), rb.value

    def contacts(self, genes: np.ndarray, contact_distance: float = 5.0):
        """(bitsets (n, words) uint64, set sizes (n,))."""
        g = np.ascontiguousarray(genes, dtype=np.float64).reshape(-1, 6)
        w = self.contact_words(contact_distance)
        bits = np.zeros((len(g), max(w, 1)), np.uint64)
        sizes = np.zeros(len(g), np.int32)
        check(lib().gdock_contacts(self.ctx.handle, self.handle, ptr(g), len(g), contact_distance,
                                   ptr(bits, C.c_uint64), ptr(sizes, C.c_int32)), self.ctx.handle, "gdock_contacts")
        return bits[:, :w], sizes

    def contact_set(self, bitset: np.ndarray, contact_distance: float = 5.0) -> set[str]:
        """Decode one pose's bitset (at this cutoff) into the reference's contact strings."""
        out = set()
        for wi, word in enumerate(bitset.tolist()):
            while word:
                b = (word & -word).bit_length() - 1
                word &= word - 1
                ca, ra, cb, rb = self.contact_pair(wi * 64 + b, contact_distance)
                out.add(f"{ca} {ra} {cb} {rb}")
        return out

    def cluster(self, genes: np.ndarray, config=None):
        from .clustering import ClusterResult, ClusteringConfig
        cfg = config or ClusteringConfig()
        g = np.ascontiguousarray(genes, dtype=np.float64).reshape(-1, 6)
        n = len(g)
        res = (ClusterResultC * max(n, 1))()
        members = np.zeros(max(n, 1), np.int64)
        nc = C.c_int32()
        cc = ClusteringConfigC(cfg.contact_distance, cfg.fcc_cutoff, cfg.strictness, cfg.min_cluster_size)
        check(lib().gdock_cluster_structures(self.ctx.handle, self.handle, ptr(g), n, C.byref(cc), res,
                                             ptr(members, C.c_int64), C.byref(nc)), self.ctx.handle,
              "gdock_cluster_structures")
        out = []
        for k in range(nc.value):
            r = res[k]
            m = members[r.members_offset:r.members_offset + r.size].tolist()
            out.append(ClusterResult(r.center_idx, m, r.size))
        return out


class Batch:
    """Several complexes scored together: one persistent launch per generation,
    replayed from a captured HIP graph (gdock_batch_*)."""

    def __init__(self, complexes: list[Complex], poses_per_complex, ctx: Context | None = None):
        self.ctx = ctx or (complexes[0].ctx if complexes else default_context())
        self.complexes = list(complexes)  # kept alive: the batch points at their device memory
        counts = np.ascontiguousarray(np.array(list(poses_per_complex), dtype=np.int32))
        assert len(counts) == len(self.complexes)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        arr = (C.c_void_p * max(len(self.complexes), 1))(*[c.handle.value for c in self.complexes])
        h = C.c_void_p()
        check(lib().gdock_batch_create(self.ctx.handle, arr, len(self.complexes), ptr(counts, C.c_int32),
                                       C.byref(h)), self.ctx.handle, "gdock_batch_create")
        self.handle = h
        self.n_poses = int(lib().gdock_batch_n_poses(h))

    def plan(self) -> dict:
        n = len(self.complexes)
        k = np.zeros(max(n, 1), np.int32)
        parts = np.zeros(max(n, 1), np.int32)
        grid, lds = C.c_int32(), C.c_int32()
        check(lib().gdock_batch_plan(self.handle, ptr(k, C.c_int32), ptr(parts, C.c_int32), C.byref(grid),
                                     C.byref(lds)), what="gdock_batch_plan")
        return {"lanes_k": k[:n].tolist(), "parts": parts[:n].tolist(), "grid": grid.value, "lds_bytes": lds.value}

    def last_items(self) -> int:
        """Work items of the last launch: one per pose, plus one per extra
        part of each pose the launch split."""
        n = C.c_int32()
        check(lib().gdock_batch_last_items(self.handle, C.byref(n)), self.ctx.handle, "gdock_batch_last_items")
        return n.value

    def score(self, genes: np.ndarray, weights: constants.EnergyWeights | None = None):
        g = np.ascontiguousarray(genes, dtype=np.float64).reshape(-1, 6)
        assert len(g) == self.n_poses, (len(g), self.n_poses)
        out = np.zeros((len(g), 6))
        ns = np.zeros(len(g), np.int32)
        w = weights_c(weights)
        check(lib().gdock_batch_score(self.ctx.handle, self.handle, ptr(g), C.byref(w), ptr(out),
                                      ptr(ns, C.c_int32)), self.ctx.handle, "gdock_batch_score")
        return out, ns

    def device_buffers(self):
        p, o, n = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().gdock_batch_device_buffers(self.handle, C.byref(p), C.byref(o), C.byref(n)),
              what="gdock_batch_device_buffers")
        return p.value, o.value, n.value

    def launch(self, weights: constants.EnergyWeights | None = None, stream: int | None = None) -> None:
        w = weights_c(weights)
        check(lib().gdock_batch_launch(self.ctx.handle, self.handle, C.byref(w), C.c_void_p(stream or 0)),
              self.ctx.handle, "gdock_batch_launch")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().gdock_batch_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""DockQ evaluator (evaluator.rs) on the device: gdock_evaluator_*."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, ptr
from .device import Complex
from .structure import Molecule


@dataclass
class Metrics:
    """evaluator.rs:29-47"""
    rmsd: float
    irmsd: float
    fnat: float
    dockq: float

    def rank(self) -> str:
        if self.dockq >= 0.8:
            return "High"
        if self.dockq >= 0.5:
            return "Medium"
        return "Low"


class Evaluator:
    """Evaluator::new(receptor, reference) for a device complex; models are the
    complex's positioned ligand moved by genes."""

    def __init__(self, cx: Complex, reference: Molecule):
        self.cx = cx
        h = C.c_void_p()
        check(lib().gdock_evaluator_create(cx.ctx.handle, cx.handle, reference.atoms.ctypes.data, len(reference),
                                           C.byref(h)), cx.ctx.handle, "gdock_evaluator_create")
        self.handle = h
        a, b, n = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().gdock_evaluator_info(h, C.byref(a), C.byref(b), C.byref(n)), what="gdock_evaluator_info")
        self.n_receptor_interface_res, self.n_reference_interface_res, self.n_native = a.value, b.value, n.value

    def calc_metrics(self, genes: np.ndarray) -> np.ndarray:
        """genes (n, 6) -> (n, 4): rmsd, irmsd, fnat, dockq."""
        g = np.ascontiguousarray(genes, dtype=np.float64).reshape(-1, 6)
        out = np.zeros((len(g), 4))
        check(lib().gdock_calc_metrics(self.cx.ctx.handle, self.handle, ptr(g), len(g), ptr(out)),
              self.cx.ctx.handle, "gdock_calc_metrics")
        return out

    def metrics(self, genes: np.ndarray) -> list[Metrics]:
        return [Metrics(*row) for row in self.calc_metrics(genes).tolist()]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().gdock_evaluator_destroy(self.cx.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

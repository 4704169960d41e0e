"""Docking server (gdock_server_*): GA jobs admitted by device memory and
batched generation by generation into one persistent launch."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import constants
from ._lib import JobStats, ServerConfig, ServerStats, check, lib, ptr
from .device import weights_c
from .ga import GaResult, default_params, take_result
from .structure import Molecule


@dataclass
class JobTimes:
    submit_s: float
    admit_s: float
    finish_s: float
    generations: int

    @property
    def latency_s(self) -> float:
        return self.finish_s - self.submit_s


class Server:
    def __init__(self, device: int = 0, max_jobs: int = 16, max_population: int = 150,
                 max_receptor_atoms: int = 8192, memory_budget: int = 0,
                 weights: constants.EnergyWeights | None = None, ga_on_device: bool = False):
        c = ServerConfig()
        check(lib().gdock_server_default_config(C.byref(c)), what="server_default_config")
        c.max_jobs, c.max_population, c.max_receptor_atoms = max_jobs, max_population, max_receptor_atoms
        c.ga_on_device = 1 if ga_on_device else 0
        c.memory_budget = memory_budget
        c.weights = weights_c(weights)
        h = C.c_void_p()
        check(lib().gdock_server_create(device, C.byref(c), C.byref(h)), what="gdock_server_create")
        self.handle = h
        self._params = {}

    def submit(self, receptor: Molecule, ligand: Molecule, restraint_pairs=(), seed: int = 42, **params) -> int:
        pairs = np.ascontiguousarray(np.array(list(restraint_pairs), dtype=np.int32).reshape(-1, 2))
        p = default_params(**params)
        jid = C.c_int64()
        check(lib().gdock_server_submit(self.handle, receptor.atoms.ctypes.data, len(receptor),
                                        ligand.atoms.ctypes.data, len(ligand), ptr(pairs, C.c_int32), len(pairs),
                                        seed, C.byref(p), C.byref(jid)), what="gdock_server_submit")
        self._params[jid.value] = p
        return jid.value

    def wait(self, job_id: int, timeout_s: float = -1.0) -> tuple[GaResult, JobTimes]:
        h = C.c_void_p()
        st = JobStats()
        check(lib().gdock_server_wait(self.handle, job_id, timeout_s, C.byref(h), C.byref(st)),
              what=f"gdock_server_wait(job {job_id})")
        p = self._params.pop(job_id)
        return take_result(h, p), JobTimes(st.submit_s, st.admit_s, st.finish_s, st.generations)

    def stats(self) -> dict:
        s = ServerStats()
        check(lib().gdock_server_stats(self.handle, C.byref(s)), what="gdock_server_stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_ if k != "reserved"}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().gdock_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

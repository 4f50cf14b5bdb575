"""Host-side mirror of the reference's simulation interface
(src/simulation.h:53-101) for Python callers: a Simulation owns one sim_t,
calls the C ABI entry points with the reference's semantics, and exposes the
blocks of sim.mem as numpy views (the sim_find layout).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Simulation:
    """sim_init on construction, sim_free on close()."""

    def __init__(self, N: int, **params):
        self.lib = L.load()
        self.params = L.set_params(**params)
        self.sim = L.SimT()
        L.check(self.lib.sim_init(C.byref(self.sim), C.c_uint(N)), "sim_init")
        self.N = N
        self.dof = int(self.lib.sf_sim_dof())
        self._closed = False

    # --- sim.mem views (reference src/simulation.c:33-53) ---
    def _block(self, e: int, count: int) -> np.ndarray:
        if e + count > self.dof:
            raise IndexError(f"block {e} is not in sim.mem (SIM_DOF = {self.dof})")
        n = self.N * count
        if n == 0:
            return np.zeros(0, np.float32)
        addr = self.sim.mem + 4 * self.N * e
        return np.ctypeslib.as_array((C.c_float * n).from_address(addr))

    @property
    def mem(self) -> np.ndarray:
        return self._block(0, self.dof)

    @property
    def pos(self) -> np.ndarray:
        return self._block(L.SIM_POS, 3).reshape(self.N, 3)

    @property
    def vel(self) -> np.ndarray:
        return self._block(L.SIM_VEL, 3).reshape(self.N, 3)

    @property
    def acc(self) -> np.ndarray:
        return self._block(L.SIM_ACC, 3).reshape(self.N, 3)

    @property
    def mas(self) -> np.ndarray:
        return self._block(L.SIM_MAS, 1)

    @property
    def pot(self) -> np.ndarray:
        return self._block(L.SIM_POT, 1)

    @property
    def ken(self) -> np.ndarray:
        return self._block(L.SIM_KEN, 1)

    @property
    def pen(self) -> np.ndarray:
        return self._block(L.SIM_PEN, 1)

    @property
    def rad(self) -> np.ndarray:
        return self._block(L.SIM_RAD, 1)

    @property
    def rho(self) -> np.ndarray:
        return self._block(L.SIM_RHO, 1)

    @property
    def prs(self) -> np.ndarray:
        return self._block(L.SIM_PRS, 1)

    @property
    def step_number(self) -> int:
        return self.sim.step_number

    @step_number.setter
    def step_number(self, v: int) -> None:
        self.sim.step_number = v

    # --- reference entry points ---
    def solve(self) -> None:
        L.check(self.lib.sim_solv(C.byref(self.sim)), "sim_solv")

    def step(self) -> None:
        L.check(self.lib.sim_step(C.byref(self.sim)), "sim_step")

    def solve_gravity_pp(self) -> None:
        L.check(self.lib.solve_gravity_part_part(C.byref(self.sim)), "solve_gravity_part_part")

    def solve_gravity_tree(self) -> None:
        L.check(self.lib.solve_gravity_part_tree(C.byref(self.sim)), "solve_gravity_part_tree")

    def solve_sph(self) -> None:
        L.check(self.lib.solve_sph(C.byref(self.sim)), "solve_sph")

    def initcond_generate(self) -> None:
        L.check(self.lib.initcond_generate(C.byref(self.sim)), "initcond_generate")

    def save_raw(self, path: str) -> None:
        L.check(self.lib.sim_save_raw(C.byref(self.sim), path.encode()), "sim_save_raw")

    def read_raw(self, path: str) -> None:
        L.check(self.lib.sim_read_raw(C.byref(self.sim), path.encode()), "sim_read_raw")

    # --- additive entry points ---
    def to_device(self) -> None:
        L.check(self.lib.sf_sync(C.byref(self.sim), L.SF_SYNC_TO_DEVICE), "sf_sync")

    def to_host(self) -> None:
        L.check(self.lib.sf_sync(C.byref(self.sim), L.SF_SYNC_TO_HOST), "sf_sync")

    def device_ptr(self, e: int) -> int:
        p = self.lib.sf_device_find(C.byref(self.sim), e)
        if not p:
            raise RuntimeError(L.last_error())
        return p

    def set_u(self, u) -> None:
        """Specific internal energy of the cubic SPH model (device array)."""
        u = np.ascontiguousarray(np.broadcast_to(np.asarray(u, np.float32), (self.N,)))
        L.check(self.lib.sf_ext_copy(C.byref(self.sim), L.SF_EXT_U, u.ctypes.data, L.SF_SYNC_TO_DEVICE),
                "sf_ext_copy")

    def _get_ext(self, which: int) -> np.ndarray:
        out = np.zeros(self.N, np.float32)
        L.check(self.lib.sf_ext_copy(C.byref(self.sim), which, out.ctypes.data, L.SF_SYNC_TO_HOST), "sf_ext_copy")
        return out

    def get_u(self) -> np.ndarray:
        return self._get_ext(L.SF_EXT_U)

    def get_dudt(self) -> np.ndarray:
        return self._get_ext(L.SF_EXT_DUDT)

    def stats(self) -> dict:
        s = L.Stats()
        L.check(self.lib.sf_get_stats(C.byref(self.sim), C.byref(s)), "sf_get_stats")
        return {"T": s.T, "V": s.V, "E": s.E, "W": s.W, "U": s.U, "P": list(s.P), "L": list(s.L), "K": s.K}

    def ic_uniform_ball(self, seed: int, radius: float = 1.0, total_mass: float = 1.0, v_hubble: float = 0.0):
        L.check(self.lib.sf_ic_uniform_ball(C.byref(self.sim), seed, radius, total_mass, v_hubble),
                "sf_ic_uniform_ball")

    def ic_plummer(self, seed: int, first: int, count: int, a: float = 1.0, M: float = 1.0,
                   centre=(0.0, 0.0, 0.0), bulk_v=(0.0, 0.0, 0.0)):
        c3 = (C.c_double * 3)(*centre)
        v3 = (C.c_double * 3)(*bulk_v)
        L.check(self.lib.sf_ic_plummer(C.byref(self.sim), seed, first, count, a, M, c3, v3), "sf_ic_plummer")

    def direct_f64(self, idx: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """FP64 direct-sum acc (n,3) and pot (n,) for the listed particles."""
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        acc = np.zeros((len(idx), 3), np.float64)
        pot = np.zeros(len(idx), np.float64)
        L.check(self.lib.sf_pp_direct_f64(C.byref(self.sim), idx.ctypes.data, len(idx), acc.ctypes.data,
                                          pot.ctypes.data), "sf_pp_direct_f64")
        return acc, pot

    def sph_direct_f64(self, idx: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """FP64 cubic-model rho (n,), SPH acc (n,3) and du/dt (n,) for the
        listed particles."""
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        rho = np.zeros(len(idx), np.float64)
        acc = np.zeros((len(idx), 3), np.float64)
        du = np.zeros(len(idx), np.float64)
        L.check(self.lib.sf_sph_direct_f64(C.byref(self.sim), idx.ctypes.data, len(idx), rho.ctypes.data,
                                           acc.ctypes.data, du.ctypes.data), "sf_sph_direct_f64")
        return rho, acc, du

    def lbvh_export(self, nodes: bool = False) -> dict:
        """Sorted Morton codes, permutation and walk counters of the last
        SF_TREE_LBVH build (and node arrays when nodes=True)."""
        N = self.N
        keys = np.zeros(N, np.uint64)
        order = np.zeros(N, np.uint32)
        counts = np.zeros(4, np.uint64)
        cm = np.zeros((2 * N - 1, 4), np.float32) if nodes else None
        lr = np.zeros((max(N - 1, 1), 4), np.int32) if nodes else None
        cell = np.zeros((2 * N - 1, 4), np.float32) if nodes else None
        L.check(self.lib.sf_lbvh_export(C.byref(self.sim.tree), keys.ctypes.data, order.ctypes.data,
                                        counts.ctypes.data, cm.ctypes.data if nodes else None,
                                        lr.ctypes.data if nodes else None, cell.ctypes.data if nodes else None),
                "sf_lbvh_export")
        out = {"keys": keys, "order": order, "cells": int(counts[0]), "parts": int(counts[1]), "bad": int(counts[2]), "batches": int(counts[3])}
        if nodes:
            out["node_cm"] = cm
            out["node_lr_range"] = lr[:N - 1]
            out["node_cell"] = cell
        return out

    def tree_nodes(self) -> np.ndarray:
        """The host node_t array of the last reference-mode build."""
        n = self.sim.tree.num_nodes
        raw = np.ctypeslib.as_array(C.cast(self.sim.tree.node, C.POINTER(C.c_uint8)), shape=(n * 80,))
        return raw.copy()

    def close(self) -> None:
        if not self._closed:
            L.check(self.lib.sim_free(C.byref(self.sim)), "sim_free")
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Host-side mirror of the reference's visualisation interface
(src/visualization.h:10-39): a Visualization owns one vis_t, draws a
Simulation into it with vis_draw and writes image files with vis_save; its
accumulation and render buffers are numpy views (sizey, sizex, 4).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .sim import Simulation


class Visualization:
    """vis_init on construction, vis_free on close()."""

    def __init__(self, sizex: int, sizey: int):
        self.lib = L.load()
        self.vis = L.VisT()
        L.check(self.lib.vis_init(C.byref(self.vis), sizex, sizey), "vis_init")
        self.sizex, self.sizey = sizex, sizey
        self._closed = False

    def _view(self, addr: int, ctype) -> np.ndarray:
        n = 4 * self.sizex * self.sizey
        return np.ctypeslib.as_array((ctype * n).from_address(addr)).reshape(self.sizey, self.sizex, 4)

    @property
    def atomic(self) -> np.ndarray:
        return self._view(self.vis.atomic_buffer, C.c_int32)

    @property
    def render(self) -> np.ndarray:
        return self._view(self.vis.render_buffer, C.c_float)

    def draw(self, sim: Simulation) -> None:
        L.check(self.lib.vis_draw(C.byref(self.vis), C.byref(sim.sim)), "vis_draw")

    def save(self, path: str, filetype: int) -> None:
        L.check(self.lib.vis_save(C.byref(self.vis), str(path).encode(), filetype), "vis_save")

    def close(self) -> None:
        if not self._closed:
            L.check(self.lib.vis_free(C.byref(self.vis)), "vis_free")
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

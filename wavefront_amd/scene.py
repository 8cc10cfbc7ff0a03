"""Text scenes (mirror of the reference's src/scene.rs), voxelized by the
library's C++ host code with the reference's exact f32 predicates."""
from __future__ import annotations

from ctypes import c_int32, c_uint8

import numpy as np

from ._lib import WavefrontError, check, lib, ptr
from .layout import MaterialGrid, MaterialTable


class Scene:
    def __init__(self, text: str):
        self.text = text

    @classmethod
    def load(cls, path) -> "Scene":
        with open(path, "r") as f:
            return cls(f.read())

    @classmethod
    def parse(cls, text: str) -> "Scene":
        return cls(text)

    def voxelize(self, grid: MaterialGrid, table: MaterialTable, dt: float, dx: float) -> int:
        """Rasterizes every primitive in file order; returns the number of
        distinct non-vacuum materials (src/scene.rs:149-188)."""
        d = grid.dims
        q16 = np.ascontiguousarray(table.q16)
        n = check(lib().wf_scene_voxelize(self.text.encode(), d.nx, d.ny, d.nz, dt, dx,
                                          ptr(grid.ids, c_uint8), ptr(q16, c_int32)))
        table.q16[:] = q16
        return n

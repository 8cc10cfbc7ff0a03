"""Grid, material and CPML layout types (mirror of the reference's src/layout.rs).

All per-run constants are computed by the library's C++ host code
(libwavefront_hip.so), which follows the reference's f32 arithmetic and calls
the same C library functions, so values here are bit-identical to the
reference's.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from ctypes import c_float, c_int32, c_uint8

from ._lib import lib, ptr

BLOCK_DIM = 8
VOXELS_PER_BLOCK = BLOCK_DIM ** 3


@dataclass(frozen=True)
class GridDims:
    """src/layout.rs:52-89.  Each axis must be a positive multiple of 8."""

    nx: int
    ny: int
    nz: int

    def __post_init__(self):
        for n in (self.nx, self.ny, self.nz):
            if n <= 0 or n % BLOCK_DIM:
                raise ValueError(
                    f"grid dimensions must be multiples of BLOCK_DIM ({BLOCK_DIM}) so the field "
                    f"grid tiles exactly with no partial edge blocks (got {self.nx}x{self.ny}x{self.nz})")

    def voxel_count(self) -> int:
        return self.nx * self.ny * self.nz

    def block_dims(self) -> tuple[int, int, int]:
        return (self.nx // BLOCK_DIM, self.ny // BLOCK_DIM, self.nz // BLOCK_DIM)

    def block_count(self) -> int:
        bx, by, bz = self.block_dims()
        return bx * by * bz


class FixedQ16_16:
    """src/layout.rs:105-126."""

    @staticmethod
    def from_f32(v: float) -> int:
        return int(lib().wf_q16_from_f32(float(v)))

    @staticmethod
    def to_f32(q: int) -> float:
        return float(lib().wf_q16_to_f32(int(q)))


class MaterialTable:
    """256 materials x {ca, cb, da, db} in Q16.16 (src/layout.rs:167-216)."""

    VACUUM = 0

    def __init__(self, q16: np.ndarray):
        self.q16 = np.ascontiguousarray(q16, dtype=np.int32).reshape(256, 4)

    @classmethod
    def vacuum_filled(cls, dt: float, d: float) -> "MaterialTable":
        t = np.zeros((256, 4), dtype=np.int32)
        lib().wf_material_table_vacuum(dt, d, ptr(t, c_int32))
        return cls(t)

    @staticmethod
    def coeffs_from_physical(eps_r, mu_r, sigma, dt, d) -> np.ndarray:
        out = np.zeros(4, dtype=np.int32)
        lib().wf_material_coeffs(eps_r, mu_r, sigma, dt, d, ptr(out, c_int32))
        return out

    def set_material(self, mat_id: int, eps_r, mu_r, sigma, dt, d) -> None:
        self.q16[mat_id] = self.coeffs_from_physical(eps_r, mu_r, sigma, dt, d)

    def get(self, mat_id: int) -> np.ndarray:
        return self.q16[mat_id].copy()

    def dequantized(self) -> np.ndarray:
        """(256, 4) float32 {ca, cb, da, db}: FixedQ16_16::to_f32 of every entry."""
        return (self.q16.astype(np.float32) * np.float32(1.0 / 65536.0)).astype(np.float32)


class MaterialGrid:
    """u8 material ids, x-fastest (z*ny + y)*nx + x (src/layout.rs:283-364).

    Held in host memory (the reference memory-maps a file; the device engine
    uploads the ids once at setup)."""

    def __init__(self, dims: GridDims, ids: np.ndarray | None = None):
        self.dims = dims
        if ids is None:
            ids = np.zeros(dims.voxel_count(), dtype=np.uint8)
        ids = np.ascontiguousarray(ids, dtype=np.uint8).reshape(-1)
        if ids.size != dims.voxel_count():
            raise ValueError("material grid size does not match the requested dimensions")
        self.ids = ids

    @classmethod
    def create(cls, dims: GridDims) -> "MaterialGrid":
        return cls(dims)

    def _index(self, x, y, z):
        return (z * self.dims.ny + y) * self.dims.nx + x

    def material_at(self, x, y, z) -> int:
        return int(self.ids[self._index(x, y, z)])

    def set_material_at(self, x, y, z, mat_id: int) -> None:
        self.ids[self._index(x, y, z)] = mat_id

    def voxelize_demo_sphere(self) -> None:
        """src/main.rs:403-425 (material id 1)."""
        d = self.dims
        lib().wf_voxelize_demo_sphere(d.nx, d.ny, d.nz, ptr(self.ids, c_uint8))


class CoeffGrid:
    """Per-voxel coefficients (src/layout.rs:224-273).

    Stored as the material ids plus the table: the reference copies
    table[id] into every voxel, so coefficient(voxel) == table[id(voxel)]
    bit for bit; the device keeps the 256-entry table on chip."""

    def __init__(self, material_grid: MaterialGrid, table: MaterialTable):
        self.dims = material_grid.dims
        self.ids = material_grid.ids
        self.table = table

    @classmethod
    def build(cls, material_grid: MaterialGrid, table: MaterialTable) -> "CoeffGrid":
        return cls(material_grid, MaterialTable(table.q16.copy()))


@dataclass
class PmlConfig:
    """src/layout.rs:403-435."""

    thickness: int = BLOCK_DIM
    grading_order: float = 3.0
    kappa_max: float = 7.0
    alpha_max: float = 0.05
    target_reflection: float = 1.0e-6


def effective_pml_thickness_blocks(dims: GridDims, requested_thickness_voxels: int) -> int:
    """src/layout.rs:525-533."""
    return int(lib().wf_effective_pml_thickness_blocks(dims.nx, dims.ny, dims.nz,
                                                      int(requested_thickness_voxels)))


class PmlProfile1D:
    """src/layout.rs:462-519: b, a, inv_kappa per position along one axis."""

    def __init__(self, b: np.ndarray, a: np.ndarray, inv_kappa: np.ndarray):
        self.b, self.a, self.inv_kappa = b, a, inv_kappa

    @classmethod
    def build(cls, n: int, config: PmlConfig, dt: float, dx: float,
              thickness_voxels: int) -> "PmlProfile1D":
        b = np.empty(n, np.float32)
        a = np.empty(n, np.float32)
        k = np.empty(n, np.float32)
        lib().wf_pml_profile(n, config.grading_order, config.kappa_max, config.alpha_max,
                             config.target_reflection, dt, dx, thickness_voxels,
                             ptr(b, c_float), ptr(a, c_float), ptr(k, c_float))
        return cls(b, a, k)

    @classmethod
    def identity(cls, n: int) -> "PmlProfile1D":
        return cls(np.ones(n, np.float32), np.zeros(n, np.float32), np.ones(n, np.float32))

    def get(self, i: int):
        return (float(self.b[i]), float(self.a[i]), float(self.inv_kappa[i]))

    def packed(self) -> np.ndarray:
        return np.concatenate([self.b, self.a, self.inv_kappa]).astype(np.float32)


class PmlContext:
    """src/layout.rs:632-671: the three axis profiles plus the shell depth.

    ``shell_blocks`` is the per-axis PML shell depth in blocks (the engine
    uses the same value on all three axes; PmlAuxGrid::build)."""

    def __init__(self, profiles, shell_blocks):
        self.profile_x, self.profile_y, self.profile_z = profiles
        self.shell_blocks = tuple(int(s) for s in shell_blocks)

    @classmethod
    def build(cls, dims: GridDims, config: PmlConfig, dt: float, dx: float) -> "PmlContext":
        tb = effective_pml_thickness_blocks(dims, config.thickness)
        tv = tb * BLOCK_DIM
        profs = [PmlProfile1D.build(n, config, dt, dx, tv) for n in (dims.nx, dims.ny, dims.nz)]
        return cls(profs, (tb, tb, tb))

    def packed_profiles(self) -> np.ndarray:
        return np.concatenate([p.packed() for p in (self.profile_x, self.profile_y, self.profile_z)])


def field_grid_zeroed(dims: GridDims) -> np.ndarray:
    """FieldGrid::zeroed (src/layout.rs:740-744) in the snapshot byte layout:
    (blocks, 6, 512) float32, component order ex ey ez hx hy hz."""
    return np.zeros((dims.block_count(), 6, VOXELS_PER_BLOCK), dtype=np.float32)


def aosoa_to_xyz(fields: np.ndarray, dims: GridDims) -> np.ndarray:
    """Reference AoSoA snapshot -> (6, nz, ny, nx) float32 view copy."""
    bx, by, bz = dims.block_dims()
    f = np.asarray(fields, dtype=np.float32).reshape(bz, by, bx, 6, 8, 8, 8)
    # axes: bz by bx c lz ly lx -> c (bz lz) (by ly) (bx lx)
    return f.transpose(3, 0, 4, 1, 5, 2, 6).reshape(6, dims.nz, dims.ny, dims.nx).copy()


def xyz_to_aosoa(xyz: np.ndarray, dims: GridDims) -> np.ndarray:
    bx, by, bz = dims.block_dims()
    f = np.asarray(xyz, dtype=np.float32).reshape(6, bz, 8, by, 8, bx, 8)
    return f.transpose(1, 3, 5, 0, 2, 4, 6).reshape(dims.block_count(), 6, VOXELS_PER_BLOCK).copy()

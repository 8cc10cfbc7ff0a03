"""Synthetic workloads: the scenes the engine is measured and checked on.

Each builder returns a ``Workload`` (geometry, material ids, coefficient
table, CPML, sources, probes, step count), built only from the library's
host code.  The same description feeds the CPU oracle in tests.

  plain_64      64^3 vacuum, 10-cell (effective 16) CPML, Gaussian Ez point
                source at the centre, 6 probes at +-20 cells, 200 steps.
  spheres_512   512^3 two dielectric spheres, 16-cell CPML, Ricker 10 GHz,
                32 seeded probes, 2000 steps, snapshots every 100.
  random_1024   1024^3 seeded random fill of 200 spheres + 200 boxes from a
                255-material palette, 20-cell (effective 24) CPML, 5 GHz
                sinusoid, 1000 steps.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .layout import GridDims, MaterialGrid, MaterialTable, PmlConfig, PmlContext
from .probe import FieldComponent as PC
from .probe import Probe
from .scene import Scene
from .source import FieldComponent as SC
from .source import Source, Waveform, default_freq, gaussian_from_freq, ricker_from_freq

C0 = 299792458.0
DX = 1.0e-3
# Courant number 0.99/sqrt(3) at dx = 1 mm, as the f32 literal 1.9065752e-12
# (0x1.0c539cp-39; c*dt*sqrt(3)/dx = 0.990000004 < 1).
DT_COURANT_099 = float(np.float32(1.9065752e-12))


@dataclass
class Workload:
    name: str
    dims: GridDims
    dx: float
    dt: float
    ids: np.ndarray | None
    table: MaterialTable
    pml_config: PmlConfig | None
    sources: list
    probes: list
    num_steps: int
    snapshot_every: int = 0
    scene_text: str = ""
    notes: dict = field(default_factory=dict)

    def pml(self) -> PmlContext | None:
        if self.pml_config is None or self.pml_config.thickness == 0:
            return None
        return PmlContext.build(self.dims, self.pml_config, self.dt, self.dx)

    def cells(self) -> int:
        return self.dims.voxel_count()


def _scene(dims: GridDims, text: str, dt: float, dx: float):
    grid = MaterialGrid.create(dims)
    table = MaterialTable.vacuum_filled(dt, dx)
    n = Scene.parse(text).voxelize(grid, table, dt, dx)
    return grid.ids, table, n


def _f32s(v) -> str:
    """Decimal text that parses back to exactly the f32 value v."""
    return repr(float(np.float32(v)))


def plain_64(steps: int = 200) -> Workload:
    n = 64
    dims = GridDims(n, n, n)
    dt, dx = DT_COURANT_099, DX
    f0 = default_freq(dt)
    src = Source(32, 32, 32, SC.Ez, 1.0, gaussian_from_freq(f0))
    f32 = np.float32
    freqs = [float(f32(f0) / f32(2.0)), float(f32(f0)), float(f32(2.0) * f32(f0))]
    pos = [(12, 32, 32), (52, 32, 32), (32, 12, 32), (32, 52, 32), (32, 32, 12), (32, 32, 52)]
    probes = [Probe(x, y, z, PC.Ez, list(freqs), 0.0) for (x, y, z) in pos]
    return Workload("plain_64", dims, dx, dt, None, MaterialTable.vacuum_filled(dt, dx),
                    PmlConfig(thickness=10), [src], probes, steps)


def spheres_512(steps: int = 2000, n: int = 512) -> Workload:
    dims = GridDims(n, n, n)
    dt, dx = DT_COURANT_099, DX
    s = n / 512.0
    text = (f"sphere 4.0 1.0 0.0 {160 * s} {256 * s} {256 * s} {80 * s}\n"
            f"sphere 2.25 1.0 0.01 {360 * s} {256 * s} {256 * s} {60 * s}\n")
    ids, table, _ = _scene(dims, text, dt, dx)
    c = n // 2
    src = Source(c, c, c, SC.Ez, 1.0, ricker_from_freq(1.0e10))
    rng = np.random.default_rng(2)
    pos = rng.integers(0, n, size=(32, 3))
    probes = [Probe(int(x), int(y), int(z), PC.Ez, [1.0e10], 0.0) for x, y, z in pos]
    return Workload("spheres_512", dims, dx, dt, ids, table, PmlConfig(thickness=16), [src],
                    probes, steps, snapshot_every=100, scene_text=text)


def random_fill_text(n: int, seed: int = 3, nspheres: int = 200, nboxes: int = 200) -> str:
    """200 spheres then 200 axis-aligned boxes, radii / sides U[10, 60],
    centres U[0, n)^3, each drawing (eps_r, sigma) from a seeded 255-entry
    palette (eps_r U[1, 10], sigma U[0, 0.05]) so the scene stays within the
    255-material limit (src/scene.rs:173-175)."""
    rng = np.random.default_rng(seed)
    k = n / 1024.0
    pal_eps = rng.uniform(1.0, 10.0, 255).astype(np.float32)
    pal_sig = rng.uniform(0.0, 0.05, 255).astype(np.float32)
    lines = []
    for _ in range(nspheres):
        m = rng.integers(0, 255)
        cx, cy, cz = rng.uniform(0, n, 3)
        r = rng.uniform(10, 60) * k
        lines.append(f"sphere {_f32s(pal_eps[m])} 1.0 {_f32s(pal_sig[m])} "
                     f"{_f32s(cx)} {_f32s(cy)} {_f32s(cz)} {_f32s(r)}")
    for _ in range(nboxes):
        m = rng.integers(0, 255)
        c = rng.uniform(0, n, 3)
        h = rng.uniform(10, 60, 3) * k / 2.0
        lo, hi = c - h, c + h
        lines.append(f"box {_f32s(pal_eps[m])} 1.0 {_f32s(pal_sig[m])} "
                     + " ".join(_f32s(v) for v in lo) + " " + " ".join(_f32s(v) for v in hi))
    return "\n".join(lines) + "\n"


def random_1024(steps: int = 1000, n: int = 1024, seed: int = 3, pml: int = 20) -> Workload:
    dims = GridDims(n, n, n)
    dt, dx = DT_COURANT_099, DX
    text = random_fill_text(n, seed)
    ids, table, nmat = _scene(dims, text, dt, dx)
    c = n // 2
    src = Source(c, c, c, SC.Ez, 1.0, Waveform.sinusoid(5.0e9))
    probes = [Probe(c + 20, c, c, PC.Ez, [5.0e9], 0.0)]
    return Workload(f"random_{n}", dims, dx, dt, ids, table, PmlConfig(thickness=pml), [src],
                    probes, steps, scene_text=text, notes={"materials": nmat})


def batch_scenes(nscenes: int = 96, steps: int = 500, seed: int = 4, edges=None,
                 max_edge: int = 192) -> list[Workload]:
    """Independent small scenes for the batched engine: each edge drawn from
    {32, 40, ..., max_edge} (multiples of 8, non-cubic), 1-4 spheres with
    r ~ U[4, min_edge/4], eps_r ~ U[1, 10], sigma ~ U[0, 0.05], mu_r = 1,
    requested CPML ~ U{8..16} (effective 8 or 16 voxels, at most half the
    smallest axis), one Gaussian Ez source at the centre and one Ez probe.
    `edges` overrides the drawn edge list (for small test batches)."""
    rng = np.random.default_rng(seed)
    dt, dx = DT_COURANT_099, DX
    f0 = default_freq(dt)
    choices = np.arange(32, max_edge + 1, 8)
    out = []
    for k in range(nscenes):
        nx, ny, nz = (int(v) for v in (edges[k] if edges is not None
                                        else rng.choice(choices, 3)))
        dims = GridDims(nx, ny, nz)
        m = min(nx, ny, nz)
        lines = []
        for _ in range(int(rng.integers(1, 5))):
            r = float(rng.uniform(4.0, m / 4.0))
            c = [float(rng.uniform(r, n - r)) for n in (nx, ny, nz)]
            eps, sig = float(rng.uniform(1.0, 10.0)), float(rng.uniform(0.0, 0.05))
            lines.append(f"sphere {_f32s(eps)} 1.0 {_f32s(sig)} "
                         + " ".join(_f32s(v) for v in c) + f" {_f32s(r)}")
        text = "\n".join(lines) + "\n"
        ids, table, _ = _scene(dims, text, dt, dx)
        pml = int(rng.integers(8, 17))
        src = Source(nx // 2, ny // 2, nz // 2, SC.Ez, 1.0, gaussian_from_freq(f0))
        probes = [Probe(nx // 4, ny // 2, nz // 2, PC.Ez, [float(np.float32(f0))], 0.0)]
        out.append(Workload(f"scene{k}_{nx}x{ny}x{nz}", dims, dx, dt, ids, table,
                            PmlConfig(thickness=pml), [src], probes, steps, scene_text=text))
    return out

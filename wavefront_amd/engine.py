"""Timestep engine on the GPU (mirror of the reference's src/engine.rs).

``DeviceEngine`` owns one wf_ctx: the field, coefficient and psi state live
in HBM for the whole run, and each ``step`` advances H, E, sources and probes
exactly as one iteration of engine::run's loop (src/engine.rs:747-782).
``run`` mirrors engine::run itself (src/engine.rs:653-819), including the
snapshot cadence and file format.
"""
from __future__ import annotations

import os
import sys
from ctypes import (byref, c_double, c_float, c_int, c_int32, c_uint8, c_uint32, c_uint64,
                    c_void_p)
from dataclasses import dataclass

import numpy as np

from ._lib import ProbeSite, SourceSite, WavefrontError, check, lib, ptr
from .layout import (CoeffGrid, GridDims, MaterialTable, PmlContext, VOXELS_PER_BLOCK,
                     field_grid_zeroed)
from .probe import Probe, probe_tables
from .source import Source, source_columns, source_values

PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 1, 2, 4


def device_count() -> int:
    return int(lib().wf_device_count())


def hbm_copy_gbs(nbytes: int, device: int = 0, iters: int = 10) -> float:
    """Read+write GB/s of a plain device copy of nbytes (the HBM ceiling)."""
    ms = c_double()
    check(lib().wf_hbm_copy(nbytes, device, iters, byref(ms)))
    return 2.0 * nbytes / (ms.value / 1e3) / 1e9


class DeviceEngine:
    """A device-resident simulation: fields, material ids, coefficient table,
    CPML profiles and psi memory on one GPU."""

    def __init__(self, dims: GridDims, material_ids=None, table: MaterialTable | None = None,
                 pml: PmlContext | None = None, periodic: int = 0, device: int = 0,
                 dt: float | None = None, dx: float | None = None):
        self.dims = dims
        if table is None:
            if dt is None or dx is None:
                raise ValueError("DeviceEngine needs a MaterialTable or dt and dx for vacuum")
            table = MaterialTable.vacuum_filled(dt, dx)
        self.table = table
        ids_p = None
        if material_ids is not None:
            self._ids = np.ascontiguousarray(material_ids, dtype=np.uint8).reshape(-1)
            if self._ids.size != dims.voxel_count():
                raise ValueError("material_ids size does not match dims")
            ids_p = ptr(self._ids, c_uint8)
        q16 = np.ascontiguousarray(table.q16, dtype=np.int32)
        shell_p = prof_p = None
        self.pml = pml
        if pml is not None:
            self._shell = np.array(pml.shell_blocks, dtype=np.uint32)
            self._prof = np.ascontiguousarray(pml.packed_profiles(), dtype=np.float32)
            shell_p = ptr(self._shell, c_uint32)
            prof_p = ptr(self._prof, c_float)
        handle = c_void_p()
        check(lib().wf_create(dims.nx, dims.ny, dims.nz, ids_p, ptr(q16, c_int32), shell_p,
                              prof_p, periodic, device, byref(handle)))
        self._h = handle
        self.nsources = 0
        self.probes: list[Probe] = []
        self._nf = 0
        self._probe_first = 0

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wf_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> c_void_p:
        return self._h

    # -- tables --------------------------------------------------------------
    def set_sources(self, sources: list[Source], dt: float, nsteps: int) -> None:
        """Precomputes amplitude*waveform(t) for the next nsteps steps, one
        value column per distinct (amplitude, waveform)."""
        first = self.step_count()
        n = len(sources)
        if n == 0:
            check(lib().wf_set_sources(self._h, 0, None, nsteps, None))
            self.nsources = 0
            return
        cols, vals = source_columns(sources, dt, first, nsteps)
        vals = np.ascontiguousarray(vals, np.float32)
        sites = (SourceSite * n)(*[s.site() for s in sources])
        check(lib().wf_set_source_columns(self._h, n, sites, ptr(cols, c_uint32), vals.shape[1],
                                          nsteps, ptr(vals, c_float)))
        self.nsources = n

    def set_source_values(self, sites: list[SourceSite], values: np.ndarray) -> None:
        values = np.ascontiguousarray(values, dtype=np.float32)
        n = len(sites)
        arr = (SourceSite * max(n, 1))(*sites)
        check(lib().wf_set_sources(self._h, n, arr, values.shape[0], ptr(values, c_float)))
        self.nsources = n

    def set_probes(self, probes: list[Probe], dt: float, nsteps: int, trace: bool = False) -> None:
        first = self.step_count()
        sites, cos_tab, sin_tab, active, nf = probe_tables(probes, dt, first, nsteps)
        self._keep = (cos_tab, sin_tab, active)
        check(lib().wf_set_probes(self._h, len(probes), sites, nsteps, ptr(cos_tab, c_double),
                                  ptr(sin_tab, c_double), ptr(active, c_uint8), 1 if trace else 0))
        self.probes = list(probes)
        self._nf = nf
        self._probe_first = first

    # -- stepping ------------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        check(lib().wf_step(self._h, nsteps))

    def sync(self) -> None:
        check(lib().wf_sync(self._h))

    def set_fp_contract(self, on: bool) -> None:
        """Contracted (fused multiply-add) step kernels instead of the exact ones."""
        check(lib().wf_set_fp_contract(self._h, 1 if on else 0))

    def run(self, num_steps: int, snapshot_every: int, output_path: str) -> float:
        """engine::run's loop with snapshot egress (wf_run): num_steps steps,
        a snapshot appended to output_path after every step whose loop index
        k satisfies k % snapshot_every == 0.  Returns wall seconds."""
        secs = c_double()
        check(lib().wf_run(self._h, num_steps, snapshot_every, os.fsencode(output_path),
                           byref(secs)))
        return secs.value

    def step_count(self) -> int:
        return int(lib().wf_step_count(self._h))

    def time_steps(self, nsteps: int) -> float:
        ms = c_double()
        check(lib().wf_time_steps(self._h, nsteps, byref(ms)))
        return ms.value

    def _kernel_names(self) -> tuple:
        v = self.kernel_variant()
        if v == "twostep":
            return ("k_twostep4", "k_twostep4_xshell", "k_twostep4_yzshell")
        if v == "twostep-rows":
            return ("k_twostep4_batch",)
        return ("k_fused4", "k_h4", "k_e4")

    def time_kernels(self, nsteps: int) -> dict:
        """Mean ms per launch of each step kernel over nsteps full steps (HIP
        events on the engine's stream): fused interior kernel, H and E
        kernels; with two steps per pass (variant "twostep"), the two-step
        kernels per pass over the interior region, the x shell regions and
        the y and z shell regions; in rows (variant "twostep-rows"), the
        batched two-step launches of a pass together."""
        f, h, e = c_double(), c_double(), c_double()
        check(lib().wf_time_kernels(self._h, nsteps, byref(f), byref(h), byref(e)))
        return dict(zip(self._kernel_names(), (f.value, h.value, e.value)))

    def kernel_bytes(self) -> dict:
        """Algorithmic HBM bytes per launch of each step kernel (the same keys
        as time_kernels)."""
        f, h, e = c_uint64(), c_uint64(), c_uint64()
        check(lib().wf_kernel_bytes(self._h, byref(f), byref(h), byref(e)))
        return dict(zip(self._kernel_names(), (f.value, h.value, e.value)))

    def kernel_variant(self) -> str:
        return lib().wf_kernel_variant(self._h).decode()

    def schedule(self):
        """(schedule chosen at creation: 0 regions, 1 rows, 2 single steps;
        the creation probe's ms per step of each, -1 where not timed)."""
        k = c_int()
        ms = (c_double * 3)()
        check(lib().wf_schedule(self._h, byref(k), ms))
        return k.value, [ms[0], ms[1], ms[2]]

    def placement(self) -> dict:
        """Field placement chosen at creation: ms per step of each arrangement
        timed (initial buffers, then each extra state buffer as state A and as
        state B; empty when no probe ran) and the index of the one kept."""
        ms = (c_double * 17)()
        n, kept = c_int(), c_int()
        check(lib().wf_placement(self._h, ms, 17, byref(n), byref(kept)))
        return {"ms_per_step": [round(ms[i], 4) for i in range(n.value)], "kept": kept.value}

    # -- state transfer --------------------------------------------------------
    def read_fields(self) -> np.ndarray:
        """Field state as the reference snapshot layout (blocks, 6, 512) f32."""
        out = field_grid_zeroed(self.dims)
        check(lib().wf_read_fields(self._h, out.ctypes.data_as(c_void_p)))
        return out

    def snapshot_async(self, buf: np.ndarray) -> None:
        """Starts copying the current state (snapshot byte format) into
        `buf` (wf_field_bytes bytes, C-contiguous); wait() before reading it."""
        if not buf.flags.c_contiguous or buf.nbytes != int(lib().wf_field_bytes(self._h)):
            raise ValueError("snapshot buffer must be C-contiguous and wf_field_bytes long")
        check(lib().wf_snapshot_async(self._h, buf.ctypes.data_as(c_void_p)))
        self._snap_buf = buf  # keep it alive until wait()

    def wait(self) -> None:
        check(lib().wf_wait(self._h))
        self._snap_buf = None

    def write_fields(self, fields: np.ndarray) -> None:
        f = np.ascontiguousarray(fields, dtype=np.float32)
        if f.size != self.dims.voxel_count() * 6:
            raise ValueError("field array size does not match dims")
        check(lib().wf_write_fields(self._h, f.ctypes.data_as(c_void_p)))

    def read_psi(self) -> np.ndarray:
        """psi in the reference PmlAux layout (shell blocks, 12, 512) f32."""
        n = int(lib().wf_psi_shell_blocks(self._h))
        out = np.zeros((n, 12, VOXELS_PER_BLOCK), np.float32)
        if n:
            check(lib().wf_read_psi(self._h, ptr(out, c_float)))
        return out

    def probe_sums(self):
        """(sum_cos[nf], sum_sin[nf], n[probes]) and updates the Probe objects."""
        nf, npb = self._nf, len(self.probes)
        sc = np.zeros(max(nf, 1), np.float64)
        ss = np.zeros(max(nf, 1), np.float64)
        cnt = np.zeros(max(npb, 1), np.uint64)
        if npb:
            check(lib().wf_read_probe_sums(self._h, ptr(sc, c_double), ptr(ss, c_double),
                                           ptr(cnt, c_uint64)))
        off = 0
        for i, p in enumerate(self.probes):
            k = len(p.frequencies_hz)
            p.sum_cos = sc[off:off + k].copy()
            p.sum_sin = ss[off:off + k].copy()
            p.samples_recorded = int(cnt[i])
            off += k
        return sc[:nf], ss[:nf], cnt[:npb]

    def read_trace(self, max_steps: int) -> np.ndarray:
        out = np.zeros((max_steps, max(len(self.probes), 1)), np.float32)
        check(lib().wf_read_trace(self._h, ptr(out, c_float), max_steps))
        return out[:, :len(self.probes)]


class DeviceBatch:
    """Many independent scenes (DeviceEngines on one GPU) advanced together:
    two steps per pass, one two-step launch per CPML variant over every
    scene's regions, then the point work of all scenes (or, with
    WF_BATCH_TWOSTEP=0, a periodic axis or a split-variant engine, one H and
    one E launch per step).  Each scene's results equal stepping it alone."""

    def __init__(self, engines: list[DeviceEngine]):
        self.engines = list(engines)
        arr = (c_void_p * len(self.engines))(*[e.handle.value for e in self.engines])
        handle = c_void_p()
        check(lib().wf_batch_create(len(self.engines), arr, byref(handle)))
        self._h = handle

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wf_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, nsteps: int = 1) -> None:
        check(lib().wf_batch_step(self._h, nsteps))

    def sync(self) -> None:
        check(lib().wf_batch_sync(self._h))

    def time_steps(self, nsteps: int) -> float:
        ms = c_double()
        check(lib().wf_batch_time_steps(self._h, nsteps, byref(ms)))
        return ms.value

    def tune(self, steps_per_candidate: int = 4):
        """Autotunes the split steps' tiling and schedule on real steps;
        returns (zchunk, lanes_per_row, grid, steps_taken), grid 0 = one
        workgroup per tile.  A batch on two-step passes has nothing to tune
        (its regions follow each scene's geometry): the current tiling and 0
        steps."""
        z, l, g, n = c_int(), c_int(), c_int(), c_uint64()
        check(lib().wf_batch_tune(self._h, steps_per_candidate, byref(z), byref(l), byref(g),
                                  byref(n)))
        return z.value, l.value, g.value, n.value

    def set_tiling(self, zchunk: int, lanes_per_row: int, grid: int = 0) -> None:
        check(lib().wf_batch_set_tiling(self._h, zchunk, lanes_per_row, grid))

    def two_step(self):
        """(True, algorithmic bytes of one pass) when the batch runs two
        steps per pass, else (False, 0)."""
        t, nb = c_int(), c_uint64()
        check(lib().wf_batch_two_step(self._h, byref(t), byref(nb)))
        return bool(t.value), int(nb.value)

    def cells(self) -> int:
        return int(lib().wf_batch_cells(self._h))


@dataclass
class EngineConfig:
    """src/engine.rs:608-615."""

    num_steps: int
    snapshot_every: int
    dt: float
    output_path: str


def run(field_grid: np.ndarray | None, coeff_grid: CoeffGrid, pml: PmlContext | None,
        sources: list[Source], probes: list[Probe], config: EngineConfig,
        device: int = 0) -> np.ndarray:
    """engine::run (src/engine.rs:653-819) on the GPU.

    Takes the initial field state (None = zeroed), steps num_steps times,
    writes a snapshot after every step with step % snapshot_every == 0 to
    config.output_path (reference byte format, no header), updates the probes'
    accumulators and returns the final field state."""
    dims = coeff_grid.dims
    if config.snapshot_every < 1:
        raise ValueError("snapshot_every must be at least 1")
    with DeviceEngine(dims, coeff_grid.ids, coeff_grid.table, pml, device=device) as eng:
        if field_grid is not None:
            eng.write_fields(field_grid)
        eng.set_sources(sources, config.dt, config.num_steps)
        eng.set_probes(probes, config.dt, config.num_steps)
        elapsed = eng.run(config.num_steps, config.snapshot_every, str(config.output_path))
        eng.probe_sums()
        final = eng.read_fields()
    print(f"wavefront: {config.num_steps} steps over 1 device in {elapsed:.3f}s "
          f"({config.num_steps / max(elapsed, 1e-9):.1f} steps/s)", file=sys.stderr)
    return final

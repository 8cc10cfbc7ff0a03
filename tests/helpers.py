"""Shared test helpers: build the CPU oracle for a workload independently of
the product's host code (its own tables, profiles and voxelizer), and run
the same workload on the GPU engine."""
from __future__ import annotations

import numpy as np

from oracle import wfo


def oracle_table(w):
    if w.scene_text:
        ids, table, _ = wfo.scene_voxelize(w.scene_text, w.dims.nx, w.dims.ny, w.dims.nz,
                                           w.dt, w.dx)
        return ids, table
    return None, wfo.vacuum_table(w.dt, w.dx)


def oracle_pml(dims, cfg, dt, dx):
    """(shell_blocks, packed profiles) built by the oracle's own functions."""
    if cfg is None or cfg.thickness == 0:
        return None, None
    tb = wfo.lib().wfo_effective_pml_thickness_blocks(dims.nx, dims.ny, dims.nz, cfg.thickness)
    prof = np.concatenate([
        wfo.pml_profile(n, tb * 8, dt, dx, cfg.grading_order, cfg.kappa_max, cfg.alpha_max,
                        cfg.target_reflection).reshape(-1)
        for n in (dims.nx, dims.ny, dims.nz)])
    return (tb, tb, tb), prof


def oracle_sim_for(w, ids=None, table=None):
    if ids is None and table is None:
        ids, table = oracle_table(w)
    shell, prof = oracle_pml(w.dims, w.pml_config, w.dt, w.dx)
    sim = wfo.OracleSim(w.dims.nx, w.dims.ny, w.dims.nz, ids=ids, table=table,
                        shell_blocks=shell, profiles=prof)
    for s in w.sources:
        sim.add_source(s.x, s.y, s.z, int(s.component), s.amplitude, s.waveform.kind,
                       s.waveform.p0, s.waveform.p1, hard=s.hard)
    for p in w.probes:
        sim.add_probe(p.x, p.y, p.z, int(p.component), p.frequencies_hz, p.start_recording_at_s)
    return sim


def device_engine_for(w, device=0):
    from wavefront_amd.engine import DeviceEngine
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml(), device=device)
    eng.set_sources(w.sources, w.dt, w.num_steps)
    eng.set_probes(w.probes, w.dt, w.num_steps)
    return eng


def assert_bitwise(a, b, what=""):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    va = a.view(np.uint32 if a.dtype == np.float32 else np.uint64)
    vb = b.view(np.uint32 if b.dtype == np.float32 else np.uint64)
    bad = np.flatnonzero(va.reshape(-1) != vb.reshape(-1))
    if bad.size:
        i = bad[0]
        raise AssertionError(
            f"{what}: {bad.size} of {va.size} values differ bitwise; first at flat index {i}: "
            f"{a.reshape(-1)[i]!r} vs {b.reshape(-1)[i]!r}")

"""The reference's two validation studies, run on the GPU engine.

  convergence_study     examples/convergence_study.rs: plane-wave phase
                        velocity vs the closed-form Yee dispersion relation.
  pml_reflection_study  examples/pml_reflection_study.rs: CPML reflection
                        coefficient by two-run subtraction.

Both drive a hard sheet source across the full Y/Z extent of a domain one
block wide in Y and Z (wrapped periodically) and record the probe trace on
the device; the f32/f64 post-processing below keeps the examples' exact
arithmetic (numpy float32 scalars are IEEE single operations).
"""
from __future__ import annotations

import math

import numpy as np

from .engine import PERIODIC_Y, PERIODIC_Z, DeviceEngine
from .layout import BLOCK_DIM, GridDims, MaterialTable, PmlConfig, PmlContext, PmlProfile1D
from .probe import FieldComponent as PC
from .probe import Probe
from .source import FieldComponent as SC
from .source import Source, Waveform

f32 = np.float32
C0 = f32(299792458.0)


def _nx8(v: int) -> int:
    return -(-v // BLOCK_DIM) * BLOCK_DIM


def _round_usize(v) -> int:
    # Rust f32::round (half away from zero) then `as usize`
    return int(math.floor(float(v) + 0.5)) if v >= 0 else 0


def run_plane_wave(nx: int, dx, dt, freq_hz, x_source: int, x_probe: int, steps: int,
                   pml_thickness_voxels: int, device: int = 0) -> np.ndarray:
    """convergence_study.rs:81-179 / pml_reflection_study.rs:85-198 on the
    device; returns the probe's Ez trace (f32, one sample per step)."""
    dims = GridDims(nx, BLOCK_DIM, BLOCK_DIM)
    table = MaterialTable.vacuum_filled(float(dt), float(dx))
    pml = None
    if pml_thickness_voxels > 0:
        bx_n = nx // BLOCK_DIM
        tb = min(-(-pml_thickness_voxels // BLOCK_DIM), bx_n // 2)
        cfg = PmlConfig(thickness=pml_thickness_voxels)
        px = PmlProfile1D.build(nx, cfg, float(dt), float(dx), tb * BLOCK_DIM)
        ident = PmlProfile1D.identity(BLOCK_DIM)
        pml = PmlContext([px, ident, PmlProfile1D.identity(BLOCK_DIM)], (tb, 0, 0))
    wave = Waveform.sinusoid(float(freq_hz))
    sources = [Source(x_source, ly, lz, SC.Ez, 1.0, wave, hard=True)
               for ly in range(BLOCK_DIM) for lz in range(BLOCK_DIM)]
    probe = Probe(x_probe, 0, 0, PC.Ez, [], 0.0)
    with DeviceEngine(dims, None, table, pml, periodic=PERIODIC_Y | PERIODIC_Z,
                      device=device) as eng:
        eng.set_sources(sources, float(dt), steps)
        eng.set_probes([probe], float(dt), steps, trace=True)
        eng.step(steps)
        return eng.read_trace(steps)[:, 0].copy()


def _quadrature(trace: np.ndarray, dt, omega: float, skip_before_s):
    sum_cos = 0.0
    sum_sin = 0.0
    count = 0
    for i, v in enumerate(trace):
        t = (f32(i) + f32(1.0)) * dt
        if t > skip_before_s:
            phase = omega * float(t)
            sum_cos += float(v) * math.cos(phase)
            sum_sin += float(v) * math.sin(phase)
            count += 1
    return 2.0 * sum_cos / count, 2.0 * sum_sin / count


def convergence_rows(device: int = 0):
    """examples/convergence_study.rs:204-295, measured columns only:
    (dx, ppw, measured_v, measured_relative_error) per resolution."""
    freq_hz = f32(3.0e10)
    wavelength = C0 / freq_hz
    courant = f32(0.4)
    periods = f32(8.0)
    rows = []
    for ppw in (f32(10.0), f32(15.0), f32(20.0), f32(30.0)):
        dx = wavelength / ppw
        dt = courant * dx / C0
        period_s = f32(1.0) / freq_hz
        period_steps = ppw / courant
        n_vox = _round_usize(periods * ppw)
        x_source = 3 * n_vox
        x_probe = x_source + n_vox
        nx = _nx8(7 * n_vox)
        steps = int(math.ceil(float((periods + f32(12.0)) * period_steps)))
        expected = (f32(n_vox) * dx) / C0
        omega = 2.0 * math.pi * float(freq_hz)
        trace = run_plane_wave(nx, dx, dt, freq_hz, x_source, x_probe, steps, 0, device)
        skip = expected + f32(5.0) * period_s
        a, b = _quadrature(trace, dt, omega, skip)
        raw_delay = math.atan2(-a, b) / omega
        k = (float(expected) - raw_delay) / float(period_s)
        k = math.floor(k + 0.5) if k >= 0 else -math.floor(-k + 0.5)
        measured_delay = f32(raw_delay + k * float(period_s))
        measured_v = (f32(n_vox) * dx) / measured_delay
        err = abs(measured_v - C0) / C0
        rows.append((dx, ppw, measured_v, err))
    return rows


def pml_reflection_rows(device: int = 0):
    """examples/pml_reflection_study.rs:237-320: reflection coefficient per
    thickness (1, 2, 4, 8 blocks)."""
    freq_hz = f32(3.0e10)
    wavelength = C0 / freq_hz
    courant = f32(0.4)
    ppw = f32(20.0)
    dx = wavelength / ppw
    dt = courant * dx / C0
    period_s = f32(1.0) / freq_hz
    probe_off = _round_usize(f32(3.0) * ppw)
    right_gap = _round_usize(f32(2.0) * ppw)
    out = []
    for tb in (1, 2, 4, 8):
        tv = tb * BLOCK_DIM
        xs = tv + _round_usize(f32(2.0) * ppw)
        xp = xs + probe_off
        face = xp + right_gap
        nx_test = _nx8(face + tv)
        incident = f32(probe_off) * dx / C0
        round_trip = f32(2 * (face - xp)) * dx / C0
        skip = incident + round_trip + f32(6.0) * period_s
        steps = int(math.ceil(float((skip + f32(10.0) * period_s) / dt))) + 4
        test = run_plane_wave(nx_test, dx, dt, freq_hz, xs, xp, steps, tv, device)
        required = int(math.ceil(float((f32(steps) * dt + f32(2.0) * period_s) * C0 / dx))) \
            + BLOCK_DIM
        xpr = required + probe_off
        ref = run_plane_wave(_nx8(xpr + required), dx, dt, freq_hz, xpr - probe_off, xpr,
                             steps, 0, device)
        omega = 2.0 * math.pi * float(freq_hz)
        a, b = _quadrature(ref, dt, omega, skip)
        inc = math.sqrt(a * a + b * b)
        a, b = _quadrature((test - ref).astype(np.float32), dt, omega, skip)
        refl = math.sqrt(a * a + b * b)
        out.append(refl / inc)
    return out

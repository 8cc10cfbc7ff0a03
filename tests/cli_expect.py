"""Expected `wavefront` CLI output, written independently of the C++ driver:
Rust's float formatting rebuilt from numpy's shortest-repr printer, the
demo-run setup from the oracle's own builders, and Probe::spectrum
(src/probe.rs:143-167) restated in Python floats (IEEE f64, glibc libm)."""
from __future__ import annotations

import math
import os

import numpy as np

from oracle import wfo

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "wavefront_amd", "bin", "wavefront")


def rust_debug_f32(v) -> str:
    """Rust `{:?}` of an f32: shortest round-trip digits; positional with at
    least one fractional digit when 1e-4 <= |v| < 1e16 (or 0), else d.ddde<exp>."""
    v = np.float32(v)
    if np.isnan(v):
        return "NaN"
    if np.isinf(v):
        return "-inf" if v < 0 else "inf"
    a = np.float32(abs(v))
    if a == 0:
        return "-0.0" if np.signbit(v) else "0.0"
    if np.float32(1e-4) <= a < np.float32(1e16):
        s = np.format_float_positional(v, unique=True, trim="0")
        return s
    s = np.format_float_scientific(v, unique=True, trim="-")
    mant, exp = s.split("e")
    return f"{mant}e{int(exp)}"


def rust_exp(v: float, prec: int) -> str:
    """Rust `{:.N e}` (no exponent sign for positive exponents, no padding)."""
    mant, exp = f"{v:.{prec}e}".split("e")
    return f"{mant}e{int(exp)}"


def default_source_line(nx, ny, nz, dt, waveform="ricker", freq=None, comp="Ez", x=None,
                        y=None, z=None):
    f32 = np.float32
    f = f32(1.0) / (f32(20.0) * f32(dt)) if freq is None else f32(freq)
    if waveform == "ricker":
        wf = f"RickerWavelet {{ peak_freq_hz: {rust_debug_f32(f)}, t0: {rust_debug_f32(f32(1.0) / f)} }}"
    elif waveform == "sinusoid":
        wf = f"Sinusoid {{ freq_hz: {rust_debug_f32(f)} }}"
    else:
        spread = f32(0.5) / f
        wf = f"GaussianPulse {{ t0: {rust_debug_f32(f32(4.0) * spread)}, spread: {rust_debug_f32(spread)} }}"
    x = nx // 2 if x is None else x
    y = ny // 2 if y is None else y
    z = nz // 2 if z is None else z
    return f"wavefront: source at ({x}, {y}, {z}) driving {comp}, {wf}"


def header_line(nx, ny, nz, dx, dt, steps, pml):
    mib = nx * ny * nz * 24 // (1024 * 1024)
    return (f"wavefront: {nx}x{ny}x{nz} voxels ({mib} MiB field state, dx={rust_exp(float(np.float32(dx)), 3)} m, "
            f"dt={rust_exp(float(np.float32(dt)), 3)} s, {steps} steps, pml={pml} voxels)")


def spectrum_lines(freqs, sum_cos, sum_sin, n):
    """Probe::spectrum then the CLI's print format."""
    nn = float(max(int(n), 1))
    out = []
    for f, sc, ss in zip(freqs, sum_cos, sum_sin):
        a = 2.0 * float(sc) / nn
        b = 2.0 * float(ss) / nn
        amp = math.sqrt(a * a + b * b)
        ph = math.atan2(a, b)
        out.append(f"  {rust_exp(float(np.float32(f)), 4)} Hz: amplitude {rust_exp(amp, 6)}, "
                   f"phase {ph:.4f} rad")
    return out


def demo_oracle(nx, ny, nz, dt, dx, pml_req):
    """OracleSim for the CLI's demo-sphere run (src/main.rs:534-537)."""
    ids = np.zeros(nx * ny * nz, np.uint8)
    import ctypes
    wfo.lib().wfo_voxelize_demo_sphere(nx, ny, nz, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    table = wfo.vacuum_table(dt, dx)
    table[1] = wfo.material_coeffs(4.0, 1.0, 0.0, dt, dx)
    shell = prof = None
    if pml_req:
        tb = wfo.lib().wfo_effective_pml_thickness_blocks(nx, ny, nz, pml_req)
        prof = np.concatenate([wfo.pml_profile(n, tb * 8, dt, dx).reshape(-1) for n in (nx, ny, nz)])
        shell = (tb, tb, tb)
    return ids, wfo.OracleSim(nx, ny, nz, ids=ids, table=table, shell_blocks=shell, profiles=prof)

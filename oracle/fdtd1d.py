"""FP64 one-dimensional restatement of the reflection study's plane-wave run
-- test infrastructure only.

examples/pml_reflection_study.rs:85-198 drives a hard Ez sheet source across
a domain one block wide in Y and Z with periodic wrap, so every Y/Z
derivative is zero and the 3D Yee update reduces exactly to the 1D update
along X of the (Ez, Hy) pair (src/fdtd.rs:194-253, 369-428):

    Hy = da*Hy - db*(0 - (Ez[x+1] - Ez[x]))      (zero beyond the last cell)
    Ez = ca*Ez + cb*(Hy[x] - Hy[x-1])            (zero before the first)

with the vacuum coefficients dequantized from Q16.16 (src/layout.rs:194-208).
Here it runs in float64 with a float64 source sin(2 pi f t), t = (n+1) dt,
as the FP64 oracle of the incident wave that the learned boundary's
reflection is measured against.
"""
from __future__ import annotations

import math

import numpy as np


def plane_wave_trace_f64(nx: int, coeffs_q16, dt: float, freq: float, x_source: int,
                         x_probe: int, steps: int) -> np.ndarray:
    """Probe Ez per step of the PML-free study run, in float64."""
    ca, cb, da, db = (float(q) * (1.0 / 65536.0) for q in coeffs_q16)
    ez = np.zeros(nx)
    hy = np.zeros(nx)
    out = np.empty(steps)
    w = 2.0 * math.pi * float(freq)
    for n in range(steps):
        d = np.empty(nx)
        d[:-1] = ez[1:] - ez[:-1]
        d[-1] = -ez[-1]
        hy = da * hy + db * d
        c = np.empty(nx)
        c[1:] = hy[1:] - hy[:-1]
        c[0] = hy[0]
        ez = ca * ez + cb * c
        ez[x_source] = math.sin(w * (n + 1) * float(dt))
        out[n] = ez[x_probe]
    return out

"""Point sources (mirror of the reference's src/source.rs).

The device applies precomputed per-step values; Waveform.evaluate and the
value table come from the library's C++ host code (same f32 arithmetic and C
library calls as the reference)."""
from __future__ import annotations

import enum
from dataclasses import dataclass

import numpy as np
from ctypes import c_float

from ._lib import SourceDesc, SourceSite, check, lib, ptr


class FieldComponent(enum.IntEnum):
    """src/source.rs:20-25"""

    Ex = 0
    Ey = 1
    Ez = 2


WAVE_GAUSSIAN, WAVE_SINUSOID, WAVE_RICKER = 0, 1, 2


@dataclass(frozen=True)
class Waveform:
    """src/source.rs:30-43.  kind: 0 Gaussian(t0, spread), 1 Sinusoid(freq),
    2 Ricker(peak_freq, t0)."""

    kind: int
    p0: float
    p1: float = 0.0

    @classmethod
    def gaussian_pulse(cls, t0: float, spread: float) -> "Waveform":
        return cls(WAVE_GAUSSIAN, t0, spread)

    @classmethod
    def sinusoid(cls, freq_hz: float) -> "Waveform":
        return cls(WAVE_SINUSOID, freq_hz, 0.0)

    @classmethod
    def ricker_wavelet(cls, peak_freq_hz: float, t0: float) -> "Waveform":
        return cls(WAVE_RICKER, peak_freq_hz, t0)

    def evaluate(self, t: float) -> float:
        return float(lib().wf_waveform(self.kind, self.p0, self.p1, t))


@dataclass(frozen=True)
class Source:
    """src/source.rs:65-72.  ``hard`` assigns instead of adding (the sheet
    source of the validation studies)."""

    x: int
    y: int
    z: int
    component: FieldComponent
    amplitude: float
    waveform: Waveform
    hard: bool = False

    def site(self) -> SourceSite:
        return SourceSite(self.x, self.y, self.z, int(self.component), 1 if self.hard else 0)

    def desc(self) -> SourceDesc:
        return SourceDesc(self.site(), self.amplitude, self.waveform.kind,
                          self.waveform.p0, self.waveform.p1)


def source_values(sources, dt: float, first_step: int, nsteps: int) -> np.ndarray:
    """values[k, s] = amplitude * waveform(t), t = (first_step+k+1)*dt."""
    n = len(sources)
    out = np.zeros((nsteps, max(n, 1)), dtype=np.float32)
    if n == 0:
        return out[:, :0]
    arr = (SourceDesc * n)(*[s.desc() for s in sources])
    check(lib().wf_source_values(n, arr, dt, first_step, nsteps, ptr(out, c_float)))
    return out


def source_columns(sources, dt: float, first_step: int, nsteps: int):
    """Sources grouped by (amplitude, waveform): (cols[s], values[k, col]),
    one column per distinct pair, evaluated once (amplitude * waveform(t),
    t = (first_step+k+1)*dt); sources with equal parameters give equal
    values, so a sheet source needs one column whatever its size."""
    key_col: dict = {}
    cols = np.zeros(len(sources), np.uint32)
    reps = []
    for i, s in enumerate(sources):
        w = s.waveform
        key = (np.float32(s.amplitude).tobytes(), int(w.kind), np.float32(w.p0).tobytes(),
               np.float32(w.p1).tobytes())
        if key not in key_col:
            key_col[key] = len(reps)
            reps.append(s)
        cols[i] = key_col[key]
    return cols, source_values(reps, dt, first_step, nsteps)


def gaussian_from_freq(freq_hz: float) -> Waveform:
    """src/main.rs:443-449: spread = 0.5/f, t0 = 4*spread (f32)."""
    spread = np.float32(0.5) / np.float32(freq_hz)
    return Waveform.gaussian_pulse(float(np.float32(4.0) * spread), float(spread))


def ricker_from_freq(freq_hz: float) -> Waveform:
    """src/main.rs:439-442: t0 = 1/f (f32)."""
    return Waveform.ricker_wavelet(freq_hz, float(np.float32(1.0) / np.float32(freq_hz)))


def default_freq(dt: float) -> float:
    """src/main.rs:435: 1/(20*dt) in f32."""
    return float(np.float32(1.0) / (np.float32(20.0) * np.float32(dt)))

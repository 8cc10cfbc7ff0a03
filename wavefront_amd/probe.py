"""Running-DFT probes (mirror of the reference's src/probe.rs).

The f64 accumulation runs on the device in step order; the twiddles
cos/sin(2*pi*f*t) are precomputed by the library's C++ host code with the
reference's exact expressions."""
from __future__ import annotations

import enum
from dataclasses import dataclass, field

import numpy as np
from ctypes import c_double, c_float, c_uint8

from ._lib import ProbeSite, check, lib, ptr


class FieldComponent(enum.IntEnum):
    """src/probe.rs:30-38"""

    Ex = 0
    Ey = 1
    Ez = 2
    Hx = 3
    Hy = 4
    Hz = 5


@dataclass
class FrequencyResponse:
    freq_hz: float
    amplitude: float
    phase_rad: float


@dataclass
class Probe:
    """src/probe.rs:63-75."""

    x: int
    y: int
    z: int
    component: FieldComponent
    frequencies_hz: list
    start_recording_at_s: float = 0.0
    sum_cos: np.ndarray = field(default=None)
    sum_sin: np.ndarray = field(default=None)
    samples_recorded: int = 0

    def __post_init__(self):
        n = len(self.frequencies_hz)
        if self.sum_cos is None:
            self.sum_cos = np.zeros(n, np.float64)
        if self.sum_sin is None:
            self.sum_sin = np.zeros(n, np.float64)

    def site(self) -> ProbeSite:
        return ProbeSite(self.x, self.y, self.z, int(self.component), len(self.frequencies_hz))

    def position(self):
        return (self.x, self.y, self.z)

    def spectrum(self) -> list:
        """src/probe.rs:143-167."""
        n = len(self.frequencies_hz)
        amp = np.zeros(max(n, 1), np.float64)
        ph = np.zeros(max(n, 1), np.float64)
        sc = np.ascontiguousarray(self.sum_cos, np.float64)
        ss = np.ascontiguousarray(self.sum_sin, np.float64)
        if n:
            lib().wf_probe_spectrum(n, ptr(sc, c_double), ptr(ss, c_double),
                                    int(self.samples_recorded), ptr(amp, c_double),
                                    ptr(ph, c_double))
        return [FrequencyResponse(float(np.float32(f)), float(amp[i]), float(ph[i]))
                for i, f in enumerate(self.frequencies_hz)]


def probe_tables(probes, dt: float, first_step: int, nsteps: int):
    """(sites, cos_tab, sin_tab, active) for wf_set_probes."""
    n = len(probes)
    sites = (ProbeSite * max(n, 1))(*[p.site() for p in probes])
    freqs = np.array([f for p in probes for f in p.frequencies_hz], dtype=np.float32)
    nf = freqs.size
    starts = np.array([p.start_recording_at_s for p in probes], dtype=np.float32)
    cos_tab = np.zeros((nsteps, max(nf, 1)), np.float64)
    sin_tab = np.zeros((nsteps, max(nf, 1)), np.float64)
    active = np.zeros((nsteps, max(n, 1)), np.uint8)
    if n:
        check(lib().wf_probe_twiddles(n, sites, ptr(freqs, c_float) if nf else None,
                                      ptr(starts, c_float), dt, first_step, nsteps,
                                      ptr(cos_tab, c_double), ptr(sin_tab, c_double),
                                      ptr(active, c_uint8)))
    return sites, cos_tab, sin_tab, active, nf

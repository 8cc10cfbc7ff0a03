"""Pins the CPU oracle to the reference-held golden vectors.

validation/convergence_data.csv and validation/pml_reflection_data.csv are
the reference's own outputs (examples/convergence_study.rs,
examples/pml_reflection_study.rs).  The oracle must reproduce every printed
value exactly: the f32 columns bit for bit, the f64 reflection coefficients
to the last digit."""
import csv
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rows(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return list(csv.reader(f))


def test_convergence_csv_bit_exact(wfo):
    rows = _rows("convergence_data.csv")
    assert rows[0] == ["dx_m", "points_per_wavelength", "measured_phase_velocity",
                       "theoretical_phase_velocity", "measured_relative_error",
                       "theoretical_relative_error"]
    got = wfo.convergence_study(threads=2)
    for r, g in zip(rows[1:], got):
        exp = np.array([float(v) for v in r], dtype=np.float32)
        assert exp.view(np.uint32).tolist() == g.view(np.uint32).tolist(), (r, g)


def test_convergence_gates(wfo):
    """examples/convergence_study.rs:340-349: slope in [1.5, 2.6), max
    discrepancy < 0.01 (computed from the oracle's rows)."""
    got = wfo.convergence_study(threads=2)
    dx = np.log10(got[:, 0].astype(np.float64))
    for col in (4, 5):
        y = np.log10(got[:, col].astype(np.float64))
        slope = np.polyfit(dx, y, 1)[0]
        assert 1.5 <= slope < 2.6
    assert np.abs(got[:, 4] - got[:, 5]).max() < 0.01


def test_pml_reflection_csv_exact(wfo):
    rows = _rows("pml_reflection_data.csv")
    assert rows[0] == ["thickness_blocks", "thickness_voxels", "reflection_coeff"]
    got = wfo.pml_reflection_study(threads=2)
    assert [float(r[2]) for r in rows[1:]] == got.tolist()
    # pml_reflection_study.rs:351-374 gates
    assert all(b < a for a, b in zip(got, got[1:]))
    assert got[-1] < 1e-2


def test_yee_closed_form_against_math(wfo):
    """examples/convergence_study.rs:67-72 against an f64 evaluation of the
    same dispersion relation (independent of the f32 rounding)."""
    c = 299792458.0
    for ppw in (10.0, 15.0, 20.0, 30.0):
        dx = (c / 3e10) / ppw
        dt = 0.4 * dx / c
        w = 2 * math.pi * 3e10
        k = (2 / dx) * math.asin((dx / (c * dt)) * math.sin(w * dt / 2))
        v64 = w / k
        v32 = wfo.lib().wfo_yee_phase_velocity(dx, dt, 3e10)
        assert abs(v32 - v64) / v64 < 1e-5

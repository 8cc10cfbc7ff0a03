"""ctypes wrapper of the CPU oracle (oracle/libwfo.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline, never by the
product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_size_t, c_uint,
                    c_uint8, c_uint64, c_void_p)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwfo.so")

_SIG = {
    "wfo_q16_from_f32": (c_int32, [c_float]),
    "wfo_q16_to_f32": (c_float, [c_int32]),
    "wfo_material_coeffs": (None, [c_float, c_float, c_float, c_float, c_float, POINTER(c_int32)]),
    "wfo_effective_pml_thickness_blocks": (c_size_t, [c_size_t, c_size_t, c_size_t, c_size_t]),
    "wfo_pml_profile": (None, [c_size_t, c_float, c_float, c_float, c_float, c_float, c_float,
                               c_size_t, POINTER(c_float), POINTER(c_float), POINTER(c_float)]),
    "wfo_waveform": (c_float, [c_int, c_float, c_float, c_float]),
    "wfo_sim_create": (c_void_p, [c_size_t, c_size_t, c_size_t, POINTER(c_uint8), POINTER(c_int32),
                                  POINTER(c_size_t), POINTER(c_float), c_uint]),
    "wfo_sim_destroy": (None, [c_void_p]),
    "wfo_sim_add_source": (c_int, [c_void_p, c_size_t, c_size_t, c_size_t, c_int, c_int, c_float,
                                   c_int, c_float, c_float]),
    "wfo_sim_add_probe": (c_int, [c_void_p, c_size_t, c_size_t, c_size_t, c_int, c_size_t,
                                  POINTER(c_float), c_float]),
    "wfo_sim_run": (None, [c_void_p, c_size_t, c_size_t, c_float, c_int, POINTER(c_float)]),
    "wfo_sim_step_fields": (None, [c_void_p, c_int]),
    "wfo_sim_bytes": (c_size_t, [c_void_p]),
    "wfo_sim_read_aosoa": (None, [c_void_p, c_void_p]),
    "wfo_sim_write_aosoa": (None, [c_void_p, c_void_p]),
    "wfo_sim_shell_block_count": (c_size_t, [c_void_p]),
    "wfo_sim_read_psi": (None, [c_void_p, POINTER(c_float)]),
    "wfo_sim_accumulate_probes": (None, [c_void_p, c_float]),
    "wfo_sim_probe_count": (c_size_t, [c_void_p]),
    "wfo_sim_probe_nfreq": (c_size_t, [c_void_p, c_size_t]),
    "wfo_sim_probe_sums": (None, [c_void_p, c_size_t, POINTER(c_double), POINTER(c_double),
                                  POINTER(c_uint64)]),
    "wfo_probe_spectrum": (None, [c_size_t, POINTER(c_double), POINTER(c_double), c_uint64,
                                  POINTER(c_double), POINTER(c_double)]),
    "wfo_scene_voxelize": (c_int, [c_char_p, c_size_t, c_size_t, c_size_t, c_float, c_float,
                                   POINTER(c_uint8), POINTER(c_int32), c_char_p, c_size_t]),
    "wfo_voxelize_demo_sphere": (None, [c_size_t, c_size_t, c_size_t, POINTER(c_uint8)]),
    "wfo_convergence_study": (None, [POINTER(c_float), c_int]),
    "wfo_pml_reflection_study": (None, [POINTER(c_double), c_int]),
    "wfo_yee_phase_velocity": (c_float, [c_float, c_float, c_float]),
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def material_coeffs(eps_r, mu_r, sigma, dt, d) -> np.ndarray:
    out = np.zeros(4, np.int32)
    lib().wfo_material_coeffs(eps_r, mu_r, sigma, dt, d, _p(out, c_int32))
    return out


def vacuum_table(dt, d) -> np.ndarray:
    t = np.zeros((256, 4), np.int32)
    t[:] = material_coeffs(1.0, 1.0, 0.0, dt, d)
    return t


def pml_profile(n, T, dt, dx, m=3.0, kmax=7.0, amax=0.05, r0=1e-6) -> np.ndarray:
    """(3, n) float32: b, a, inv_kappa."""
    out = np.zeros((3, n), np.float32)
    lib().wfo_pml_profile(n, m, kmax, amax, r0, dt, dx, T, _p(out[0], c_float),
                          _p(out[1], c_float), _p(out[2], c_float))
    return out


class OracleSim:
    """wfo_sim: fields, coefficients and psi in the reference layout."""

    def __init__(self, nx, ny, nz, ids=None, table=None, shell_blocks=None, profiles=None,
                 periodic=0):
        self.dims = (nx, ny, nz)
        self._ids = None if ids is None else np.ascontiguousarray(ids, np.uint8).reshape(-1)
        self._table = np.ascontiguousarray(table, np.int32).reshape(256, 4)
        sb = None
        if shell_blocks is not None:
            sb = (c_size_t * 3)(*[int(s) for s in shell_blocks])
        prof = None
        if profiles is not None:
            self._prof = np.ascontiguousarray(profiles, np.float32).reshape(-1)
            prof = _p(self._prof, c_float)
        self.h = lib().wfo_sim_create(nx, ny, nz,
                                      None if self._ids is None else _p(self._ids, c_uint8),
                                      _p(self._table, c_int32), sb, prof, periodic)
        if not self.h:
            raise MemoryError("wfo_sim_create failed")
        self.probe_nfreq = []

    def close(self):
        if self.h:
            lib().wfo_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_source(self, x, y, z, comp, amp, kind, p0, p1=0.0, hard=False):
        assert lib().wfo_sim_add_source(self.h, x, y, z, comp, 1 if hard else 0, amp, kind, p0,
                                        p1) == 0

    def add_probe(self, x, y, z, comp, freqs, start=0.0):
        f = np.ascontiguousarray(freqs, np.float32)
        assert lib().wfo_sim_add_probe(self.h, x, y, z, comp, f.size,
                                       _p(f, c_float) if f.size else None, start) == 0
        self.probe_nfreq.append(f.size)

    def run(self, nsteps, dt, first_step=0, threads=1, trace=False):
        tr = None
        if trace:
            tr = np.zeros((nsteps, max(len(self.probe_nfreq), 1)), np.float32)
        lib().wfo_sim_run(self.h, first_step, nsteps, dt, threads,
                          None if tr is None else _p(tr, c_float))
        return None if tr is None else tr[:, :len(self.probe_nfreq)]

    def accumulate_probes(self, t):
        """Probe::accumulate(&grid, t) of every probe on the current fields."""
        lib().wfo_sim_accumulate_probes(self.h, t)

    def step_fields(self, threads=1):
        lib().wfo_sim_step_fields(self.h, threads)

    def read_fields(self) -> np.ndarray:
        nb = lib().wfo_sim_bytes(self.h)
        out = np.zeros(nb // 4, np.float32)
        lib().wfo_sim_read_aosoa(self.h, out.ctypes.data_as(c_void_p))
        return out.reshape(-1, 6, 512)

    def write_fields(self, fields):
        f = np.ascontiguousarray(fields, np.float32)
        lib().wfo_sim_write_aosoa(self.h, f.ctypes.data_as(c_void_p))

    def read_psi(self) -> np.ndarray:
        n = lib().wfo_sim_shell_block_count(self.h)
        out = np.zeros((n, 12, 512), np.float32)
        if n:
            lib().wfo_sim_read_psi(self.h, _p(out, c_float))
        return out

    def probe_sums(self):
        cs, ss, ns = [], [], []
        for p, nf in enumerate(self.probe_nfreq):
            c = np.zeros(max(nf, 1), np.float64)
            s = np.zeros(max(nf, 1), np.float64)
            n = c_uint64()
            lib().wfo_sim_probe_sums(self.h, p, _p(c, c_double), _p(s, c_double), ctypes.byref(n))
            cs.append(c[:nf])
            ss.append(s[:nf])
            ns.append(n.value)
        cat = (lambda xs: np.concatenate(xs) if xs else np.zeros(0))
        return cat(cs), cat(ss), np.array(ns, np.uint64)


def scene_voxelize(text: str, nx, ny, nz, dt, dx):
    ids = np.zeros(nx * ny * nz, np.uint8)
    table = vacuum_table(dt, dx)
    msg = ctypes.create_string_buffer(512)
    n = lib().wfo_scene_voxelize(text.encode(), nx, ny, nz, dt, dx, _p(ids, c_uint8),
                                 _p(table, c_int32), msg, 512)
    if n < 0:
        raise ValueError(msg.value.decode())
    return ids, table, n


def convergence_study(threads=1) -> np.ndarray:
    out = np.zeros((4, 6), np.float32)
    lib().wfo_convergence_study(_p(out, c_float), threads)
    return out


def pml_reflection_study(threads=1) -> np.ndarray:
    out = np.zeros(4, np.float64)
    lib().wfo_pml_reflection_study(_p(out, c_double), threads)
    return out

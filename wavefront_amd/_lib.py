"""ctypes binding of libwavefront_hip.so (the C ABI in include/wavefront_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C wavefront_amd/csrc``).  Importing this module fails loudly when the
library is missing: there is no CPU fallback for the device engine.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_uint8,
                    c_uint32, c_uint64, c_void_p)

LIB_PATH = os.environ.get("WF_LIB_PATH") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libwavefront_hip.so")

WF_OK = 0
WF_ERR_INVALID = -1
WF_ERR_HIP = -2
WF_ERR_OOM = -3
WF_ERR_STATE = -4
WF_ERR_IO = -5


class WavefrontError(RuntimeError):
    """A negative status returned by the C ABI, with wf_last_error()'s text."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class SourceSite(ctypes.Structure):
    _fields_ = [("x", c_uint32), ("y", c_uint32), ("z", c_uint32),
                ("component", c_int32), ("hard", c_int32)]


class SourceDesc(ctypes.Structure):
    _fields_ = [("site", SourceSite), ("amplitude", c_float), ("waveform", c_int32),
                ("p0", c_float), ("p1", c_float)]


class FaceRegion(ctypes.Structure):
    _fields_ = [("b0", ctypes.c_int32), ("fast0", ctypes.c_int32), ("nfast", ctypes.c_int32),
                ("slow0", ctypes.c_int32), ("nslow", ctypes.c_int32)]


class ProbeSite(ctypes.Structure):
    _fields_ = [("x", c_uint32), ("y", c_uint32), ("z", c_uint32),
                ("component", c_int32), ("nfreq", c_uint32)]


# name -> (restype, argtypes)
_SIGNATURES = {
    "wf_last_error": (c_char_p, []),
    "wf_device_count": (c_int, []),
    "wf_q16_from_f32": (c_int32, [c_float]),
    "wf_q16_to_f32": (c_float, [c_int32]),
    "wf_material_coeffs": (None, [c_float, c_float, c_float, c_float, c_float, POINTER(c_int32)]),
    "wf_material_table_vacuum": (None, [c_float, c_float, POINTER(c_int32)]),
    "wf_effective_pml_thickness_blocks": (c_uint32, [c_uint32, c_uint32, c_uint32, c_uint32]),
    "wf_pml_profile": (None, [c_uint32, c_float, c_float, c_float, c_float, c_float, c_float,
                              c_uint32, POINTER(c_float), POINTER(c_float), POINTER(c_float)]),
    "wf_waveform": (c_float, [c_int32, c_float, c_float, c_float]),
    "wf_source_values": (c_int, [c_uint32, POINTER(SourceDesc), c_float, c_uint64, c_uint64,
                                 POINTER(c_float)]),
    "wf_probe_twiddles": (c_int, [c_uint32, POINTER(ProbeSite), POINTER(c_float), POINTER(c_float),
                                  c_float, c_uint64, c_uint64, POINTER(c_double),
                                  POINTER(c_double), POINTER(c_uint8)]),
    "wf_probe_spectrum": (None, [c_uint32, POINTER(c_double), POINTER(c_double), c_uint64,
                                 POINTER(c_double), POINTER(c_double)]),
    "wf_scene_voxelize": (c_int, [c_char_p, c_uint32, c_uint32, c_uint32, c_float, c_float,
                                  POINTER(c_uint8), POINTER(c_int32)]),
    "wf_voxelize_demo_sphere": (None, [c_uint32, c_uint32, c_uint32, POINTER(c_uint8)]),
    "wf_create": (c_int, [c_uint32, c_uint32, c_uint32, POINTER(c_uint8), POINTER(c_int32),
                          POINTER(c_uint32), POINTER(c_float), c_uint32, c_int, POINTER(c_void_p)]),
    "wf_destroy": (None, [c_void_p]),
    "wf_set_sources": (c_int, [c_void_p, c_uint32, POINTER(SourceSite), c_uint64, POINTER(c_float)]),
    "wf_set_source_columns": (c_int, [c_void_p, c_uint32, POINTER(SourceSite), POINTER(c_uint32),
                                      c_uint32, c_uint64, POINTER(c_float)]),
    "wf_set_probes": (c_int, [c_void_p, c_uint32, POINTER(ProbeSite), c_uint64, POINTER(c_double),
                              POINTER(c_double), POINTER(c_uint8), c_int]),
    "wf_step": (c_int, [c_void_p, c_uint64]),
    "wf_sync": (c_int, [c_void_p]),
    "wf_set_fp_contract": (c_int, [c_void_p, c_int]),
    "wf_snapshot_async": (c_int, [c_void_p, c_void_p]),
    "wf_wait": (c_int, [c_void_p]),
    "wf_step_count": (c_uint64, [c_void_p]),
    "wf_field_bytes": (c_uint64, [c_void_p]),
    "wf_read_fields": (c_int, [c_void_p, c_void_p]),
    "wf_write_fields": (c_int, [c_void_p, c_void_p]),
    "wf_run": (c_int, [c_void_p, c_uint64, c_uint64, c_char_p, POINTER(c_double)]),
    "wf_batch_create": (c_int, [c_uint32, POINTER(c_void_p), POINTER(c_void_p)]),
    "wf_batch_destroy": (None, [c_void_p]),
    "wf_batch_step": (c_int, [c_void_p, c_uint64]),
    "wf_batch_sync": (c_int, [c_void_p]),
    "wf_batch_time_steps": (c_int, [c_void_p, c_uint64, POINTER(c_double)]),
    "wf_batch_tune": (c_int, [c_void_p, c_uint32, POINTER(c_int), POINTER(c_int),
                              POINTER(c_int), POINTER(c_uint64)]),
    "wf_batch_set_tiling": (c_int, [c_void_p, c_int, c_int, c_int]),
    "wf_batch_two_step": (c_int, [c_void_p, POINTER(c_int), POINTER(c_uint64)]),
    "wf_schedule": (c_int, [c_void_p, POINTER(c_int), POINTER(c_double)]),
    "wf_batch_cells": (c_uint64, [c_void_p]),
    "wf_mlp_param_count": (c_uint32, []),
    "wf_mlp_trainer_create": (c_int, [c_int, c_uint64, POINTER(c_float), POINTER(c_float),
                                      POINTER(c_void_p)]),
    "wf_mlp_trainer_destroy": (None, [c_void_p]),
    "wf_mlp_set_params": (c_int, [c_void_p, POINTER(c_float)]),
    "wf_mlp_get_params": (c_int, [c_void_p, POINTER(c_float)]),
    "wf_mlp_gradient": (c_int, [c_void_p, POINTER(c_float), POINTER(c_double)]),
    "wf_mlp_train": (c_int, [c_void_p, c_uint32, c_float, c_float, c_int, POINTER(c_double)]),
    "wf_mlp_predict": (c_int, [c_void_p, c_uint64, POINTER(c_float), POINTER(c_float)]),
    "wf_mlp_predict_noalloc": (c_int, [c_int, POINTER(c_float), c_int, c_uint64, POINTER(c_float),
                                       POINTER(c_float)]),
    "wf_mlp_predict_noalloc_max": (c_uint64, []),
    "wf_set_learned_faces": (c_int, [c_void_p, c_uint32, POINTER(c_float), c_int]),
    "wf_set_face_limits": (c_int, [c_void_p, c_int, c_float, c_float]),
    "wf_read_face_state": (c_int, [c_void_p, c_uint32, POINTER(c_float)]),
    "wf_set_replay_limits": (c_int, [c_void_p, c_int, c_float, c_float]),
    "wf_mlp_set_head": (c_int, [c_void_p, c_int]),
    "wf_clear_faces": (c_int, [c_void_p]),
    "wf_record_face_samples": (c_int, [c_void_p, c_uint32, POINTER(FaceRegion), c_uint64,
                                       c_uint64]),
    "wf_replay_faces": (c_int, [c_void_p, c_uint32, c_uint64, POINTER(POINTER(c_float))]),
    "wf_face_sample_count": (c_uint64, [c_void_p, c_uint32]),
    "wf_read_face_samples": (c_int, [c_void_p, c_uint32, POINTER(c_float)]),
    "wf_psi_shell_blocks": (c_uint64, [c_void_p]),
    "wf_read_psi": (c_int, [c_void_p, POINTER(c_float)]),
    "wf_read_probe_sums": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_uint64)]),
    "wf_read_trace": (c_int, [c_void_p, POINTER(c_float), c_uint64]),
    "wf_time_steps": (c_int, [c_void_p, c_uint64, POINTER(c_double)]),
    "wf_time_kernels": (c_int, [c_void_p, c_uint64, POINTER(c_double), POINTER(c_double),
                                POINTER(c_double)]),
    "wf_kernel_bytes": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64),
                                POINTER(c_uint64)]),
    "wf_kernel_variant": (c_char_p, [c_void_p]),
    "wf_placement": (c_int, [c_void_p, POINTER(c_double), c_int, POINTER(c_int), POINTER(c_int)]),
    "wf_hbm_copy": (c_int, [c_uint64, c_int, c_int, POINTER(c_double)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads (once) and returns the in-tree shared library."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                "`make -C wavefront_amd/csrc` (there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().wf_last_error()
    return msg.decode() if msg else ""


def check(code: int) -> int:
    """Raises WavefrontError for a negative status."""
    if code < 0:
        raise WavefrontError(code, last_error())
    return code


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array's data."""
    return arr.ctypes.data_as(POINTER(ctype))

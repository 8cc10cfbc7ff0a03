"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, its host-side table builders agree bit for bit with the
oracle's, and its error paths report cleanly."""
import os
import re

import numpy as np
import pytest

from oracle import wfo as W
from wavefront_amd import _lib
from wavefront_amd.layout import (GridDims, MaterialGrid, MaterialTable, PmlConfig, PmlContext,
                                  PmlProfile1D, aosoa_to_xyz, effective_pml_thickness_blocks,
                                  xyz_to_aosoa)
from wavefront_amd.scene import Scene
from wavefront_amd.source import Waveform, source_values, Source, FieldComponent
from wavefront_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(wf_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    declared = _declared_functions()
    assert len(declared) > 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert declared <= set(_lib._SIGNATURES)


def test_create_without_device_fails_cleanly():
    lib = _lib.lib()
    if lib.wf_device_count() > 0:
        pytest.skip("a GPU is present")
    from wavefront_amd.engine import DeviceEngine
    with pytest.raises(_lib.WavefrontError) as e:
        DeviceEngine(GridDims(16, 16, 16), dt=1e-12, dx=1e-3)
    assert e.value.code == _lib.WF_ERR_HIP
    assert "device" in str(e.value)


def test_create_argument_errors():
    from ctypes import byref, c_void_p
    lib = _lib.lib()
    t = np.zeros((256, 4), np.int32)
    h = c_void_p()
    P = lambda a, ct: _lib.ptr(a, ct)
    from ctypes import c_int32, c_uint32, c_float
    assert lib.wf_create(12, 16, 16, None, P(t, c_int32), None, None, 0, 0, byref(h)) == -1
    assert "multiples of 8" in _lib.last_error()
    sb = np.array([3, 1, 1], np.uint32)
    prof = np.zeros(3 * 48, np.float32)
    assert lib.wf_create(16, 16, 16, None, P(t, c_int32), P(sb, c_uint32), P(prof, c_float), 0,
                         0, byref(h)) == -1
    assert "exceeds half" in _lib.last_error()
    # non-identity profile outside the graded layer is rejected
    sb = np.array([1, 0, 0], np.uint32)
    prof = np.concatenate([np.stack([np.ones(n), np.zeros(n), np.ones(n)]).reshape(-1)
                           for n in (32, 16, 16)]).astype(np.float32)
    prof[16] = 0.5  # x profile b at index 16 (interior)
    assert lib.wf_create(32, 16, 16, None, P(t, c_int32), P(sb, c_uint32), P(prof, c_float), 0,
                         0, byref(h)) == -1
    assert "not identity" in _lib.last_error()


def test_run_argument_errors():
    from ctypes import byref, c_double
    lib = _lib.lib()
    secs = c_double()
    assert lib.wf_run(None, 10, 1, b"/tmp/x.bin", byref(secs)) == _lib.WF_ERR_INVALID
    assert "null" in _lib.last_error()
    assert lib.wf_step(None, 1) == _lib.WF_ERR_INVALID


def test_snapshot_async_argument_errors():
    lib = _lib.lib()
    assert lib.wf_snapshot_async(None, None) == _lib.WF_ERR_INVALID
    assert "null" in _lib.last_error()
    assert lib.wf_wait(None) == _lib.WF_ERR_INVALID


def test_placement_argument_errors():
    from ctypes import byref, c_int
    lib = _lib.lib()
    n, kept = c_int(), c_int()
    assert lib.wf_placement(None, None, 0, byref(n), byref(kept)) == _lib.WF_ERR_INVALID
    assert "null" in _lib.last_error()


def test_batch_argument_errors():
    from ctypes import byref, c_void_p
    lib = _lib.lib()
    h = c_void_p()
    assert lib.wf_batch_create(0, None, byref(h)) == _lib.WF_ERR_INVALID
    arr = (c_void_p * 2)(None, None)
    assert lib.wf_batch_create(2, arr, byref(h)) == _lib.WF_ERR_INVALID
    assert "null" in _lib.last_error()
    assert lib.wf_batch_step(None, 1) == _lib.WF_ERR_INVALID
    assert lib.wf_batch_set_tiling(None, 8, 64, 0) == _lib.WF_ERR_INVALID
    assert lib.wf_batch_cells(None) == 0


def test_fault_injection_hook(monkeypatch):
    from ctypes import byref, c_void_p, c_int32
    lib = _lib.lib()
    monkeypatch.setenv("WF_FAULT_INJECT", "wf_create")
    t = np.zeros((256, 4), np.int32)
    h = c_void_p()
    assert lib.wf_create(16, 16, 16, None, _lib.ptr(t, c_int32), None, None, 0, 0,
                         byref(h)) == _lib.WF_ERR_HIP
    assert "fault injected" in _lib.last_error()


# ---------------------------------------------------------------------------
# host tables: product (C++) vs oracle (C), bit for bit
# ---------------------------------------------------------------------------
def test_material_coeffs_match_oracle():
    rng = np.random.default_rng(0)
    for _ in range(500):
        eps, mu, sig = rng.uniform(1, 12), rng.uniform(0.5, 3), rng.uniform(0, 0.2)
        dt, dx = rng.uniform(1e-13, 3e-12), rng.uniform(1e-4, 2e-3)
        a = MaterialTable.coeffs_from_physical(eps, mu, sig, dt, dx)
        b = W.material_coeffs(eps, mu, sig, dt, dx)
        assert a.tolist() == b.tolist()
    assert np.array_equal(MaterialTable.vacuum_filled(1.5e-12, 1e-3).q16,
                          W.vacuum_table(1.5e-12, 1e-3))


def test_pml_profiles_match_oracle():
    for n, T in [(64, 16), (1024, 24), (200, 64), (48, 8), (16, 0)]:
        for dt in (1.9065752e-12, 1.5e-12, 3.3e-13):
            a = PmlProfile1D.build(n, PmlConfig(thickness=T), dt, 1e-3, T)
            b = W.pml_profile(n, T, dt, 1e-3)
            for x, y in zip((a.b, a.a, a.inv_kappa), b):
                assert x.view(np.uint32).tolist() == y.view(np.uint32).tolist()
    for dims in [(64, 64, 64), (1024, 1024, 1024), (16, 32, 48), (32, 40, 192)]:
        for req in (0, 1, 8, 10, 16, 20, 1000, 2**32 - 8, 2**32 - 1):
            d = GridDims(*dims)
            assert effective_pml_thickness_blocks(d, req) == \
                W.lib().wfo_effective_pml_thickness_blocks(*dims, req)
    # the CLI clamps --pml-thickness to UINT32_MAX: div_ceil must not wrap to 1 block
    assert effective_pml_thickness_blocks(GridDims(64, 64, 64), 2**32 - 1) == 4


def test_waveforms_and_source_values_match_oracle():
    rng = np.random.default_rng(1)
    waves = [Waveform.gaussian_pulse(4e-11, 1e-11), Waveform.sinusoid(5e9),
             Waveform.ricker_wavelet(1e10, 1e-10), Waveform.sinusoid(3e10)]
    for w in waves:
        for t in rng.uniform(0, 5e-10, 300).astype(np.float32):
            assert np.float32(w.evaluate(float(t))) == np.float32(
                W.lib().wfo_waveform(w.kind, w.p0, w.p1, float(t)))
    srcs = [Source(1, 2, 3, FieldComponent.Ez, 0.5 + i, w) for i, w in enumerate(waves)]
    vals = source_values(srcs, 1.9065752e-12, 7, 50)
    for k in range(50):
        t = (np.float32(7 + k) + np.float32(1.0)) * np.float32(1.9065752e-12)
        for s, src in enumerate(srcs):
            exp = np.float32(src.amplitude) * np.float32(
                W.lib().wfo_waveform(src.waveform.kind, src.waveform.p0, src.waveform.p1, float(t)))
            assert vals[k, s] == exp


def test_scene_voxelizer_matches_oracle():
    text = workloads.random_fill_text(96, seed=5, nspheres=30, nboxes=30)
    text += "sphere 3.0 1.0 0.01 -10 40 50 30\nbox 2 1 0 90 -4 10 200 20.5 30\n"
    grid = MaterialGrid.create(GridDims(96, 96, 96))
    table = MaterialTable.vacuum_filled(1.9e-12, 1e-3)
    n = Scene.parse(text).voxelize(grid, table, 1.9e-12, 1e-3)
    ids, otable, on = W.scene_voxelize(text, 96, 96, 96, 1.9e-12, 1e-3)
    assert n == on
    assert np.array_equal(grid.ids, ids)
    assert np.array_equal(table.q16, otable)
    with pytest.raises(_lib.WavefrontError):
        Scene.parse("sphere 1 2\n").voxelize(grid, table, 1e-12, 1e-3)


def _full_scan_spheres(text, nx, ny, nz):
    """The reference's sphere rasterizer taken literally (src/scene.rs:197-209):
    every voxel of the grid, the f32 predicate (x + 0.5 - c) per axis, squared,
    summed left to right and compared with r*r, each operation rounded to f32
    (numpy float32 arithmetic, no contraction).  Spheres only; id k for the
    k-th distinct material, later primitives overwrite earlier ones."""
    f = np.float32
    ids = np.zeros((nz, ny, nx), np.uint8)
    mats = {}
    for line in text.splitlines():
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        assert tok[0] == "sphere"
        v = [f(t) for t in tok[1:]]
        key = tuple(np.array(v[:3], f).view(np.uint32).tolist())
        mid = mats.setdefault(key, len(mats) + 1)
        c, r = v[3:6], v[6]
        d = [((np.arange(n, dtype=f) + f(0.5)) - c[a]).astype(f) for a, n in
             enumerate((nx, ny, nz))]
        sq = [(x * x).astype(f) for x in d]
        s = ((sq[0][None, None, :] + sq[1][None, :, None]).astype(f) + sq[2][:, None, None]).astype(f)
        ids[s <= f(r * r)] = mid
    return ids.reshape(-1)


def test_sphere_rasterizer_equals_full_scan():
    """Product and oracle against the literal full scan, on spheres whose f32
    rounding at |c| spans many cells: a half-space written as a huge sphere
    (centre 1e9, radius 999999900: the full scan sets x = 96..125, where a
    range of c +- (|r| + 2) cells would start at 126), its mirror image, tiny
    and zero radii, centres outside the grid and an ordinary sphere."""
    scenes = [
        "sphere 4 1 0 1e9 8 8 999999900\n",
        "sphere 2.5 1 0.01 -999999900 8 8 999999972\n",
        "sphere 3 1 0 8 1e8 8 99999960\nsphere 5 1 0 40 8 8 6.5\n",
        "sphere 2 1 0 60.5 3.5 12.25 0\nsphere 2 1 0 61 4 12 0.4\n",
        "sphere 7 1 0 -30 -30 -30 60\nsphere 1.5 1 0 3e7 -3e7 8 4.2426408e7\n",
    ]
    nx, ny, nz = 128, 16, 24
    for text in scenes:
        exp = _full_scan_spheres(text, nx, ny, nz)
        grid = MaterialGrid.create(GridDims(nx, ny, nz))
        table = MaterialTable.vacuum_filled(1.9e-12, 1e-3)
        Scene.parse(text).voxelize(grid, table, 1.9e-12, 1e-3)
        ids, _, _ = W.scene_voxelize(text, nx, ny, nz, 1.9e-12, 1e-3)
        assert np.array_equal(exp, grid.ids), text
        assert np.array_equal(exp, ids), text
    # the huge sphere of the first scene: x = 96..125 inside on the centre row
    row = _full_scan_spheres(scenes[0], nx, ny, nz).reshape(nz, ny, nx)[8, 8]
    assert np.flatnonzero(row).tolist() == list(range(96, 128))


# Tokens against the documented grammar of Rust's f32 FromStr, which the
# reference parses scene numbers with (src/scene.rs:96): [+-]? then inf,
# infinity, nan (any case) or D+ | D+ . D* | D* . D+ with [eE][+-]?D+.
_NUMBER_TOKENS = {
    "1": True, "+1": True, "-1": True, "1.": True, ".5": True, "-.5e-2": True,
    "1e3": True, "1E+3": True, "0000.0000": True, "1e400": True, "1e-50": True,
    "inf": True, "-INF": True, "Infinity": True, "+infinity": True, "nan": True,
    "NaN": True, "-nan": True,
    "0x10": False, "0x1p3": False, "nan(1)": False, ".": False, "e5": False,
    "1e": False, "1e+": False, "--1": False, "1.2.3": False, "1,5": False,
    "infin": False, "1f": False, "+": False, "1_0": False, "inf1": False,
}


def test_scene_numbers_follow_the_reference_grammar():
    """Product and oracle accept exactly the tokens Rust's f32 parser takes:
    strtof's hex floats and nan(...) are errors, as in the reference."""
    for tok, ok in _NUMBER_TOKENS.items():
        text = f"sphere 2 1 0 {tok} 8 8 3\n"
        grid = MaterialGrid.create(GridDims(16, 16, 16))
        table = MaterialTable.vacuum_filled(1.9e-12, 1e-3)
        if ok:
            Scene.parse(text).voxelize(grid, table, 1.9e-12, 1e-3)
            W.scene_voxelize(text, 16, 16, 16, 1.9e-12, 1e-3)
        else:
            with pytest.raises(_lib.WavefrontError, match="invalid number"):
                Scene.parse(text).voxelize(grid, table, 1.9e-12, 1e-3)
            with pytest.raises(ValueError, match="invalid number"):
                W.scene_voxelize(text, 16, 16, 16, 1.9e-12, 1e-3)


def _nearest_f32(q):
    """The f32 nearest to the rational q, ties to even (exact arithmetic)."""
    from fractions import Fraction
    a = np.float32(float(q))
    best = None
    for c in (np.nextafter(a, np.float32(-np.inf)), a, np.nextafter(a, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - q)
        key = (d, int(np.array([c], np.float32).view(np.uint32)[0]) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def test_strtof_rounds_decimals_correctly():
    """Both parsers convert with the C library's strtof; the reference's
    Rust parser rounds correctly to nearest, ties to even.  strtof does the
    same on decimal strings exactly halfway between two f32 values and a
    hair either side of them (where rounding through f64 first can go
    wrong), checked against exact rational arithmetic."""
    import ctypes
    from fractions import Fraction
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    for _ in range(300):
        x = np.float32(rng.uniform(-1e6, 1e6) * 10.0 ** rng.integers(-20, 20))
        y = np.nextafter(x, np.float32(np.inf))
        mid = (Fraction(float(x)) + Fraction(float(y))) / 2
        # a finite binary fraction has a finite decimal expansion
        k = 0
        while (mid * 10 ** k).denominator != 1:
            k += 1
        digits = str(abs(mid * 10 ** k).numerator)
        sign = "-" if mid < 0 else ""
        for extra, q in (("", mid), ("1", None), ("-1", None)):
            if extra == "":
                text = f"{sign}{digits}e-{k}"
            else:
                # a 1 appended 40 decimals further down moves |q| away from the midpoint;
                # borrowing one there moves it towards zero
                step = Fraction(1, 10 ** (k + 40))
                q = mid + step if (extra == "1") == (mid > 0) else mid - step
                n = abs(q * 10 ** (k + 40)).numerator
                text = f"{sign}{n}e-{k + 40}"
            q = mid if q is None else q
            got = libc.strtof(text.encode(), None)
            assert np.float32(got) == _nearest_f32(q), text


def test_source_columns_share_equal_sources():
    """Sources with the same amplitude and waveform share one value column,
    whose values equal each source's own amplitude * waveform(t)."""
    from wavefront_amd.source import source_columns
    g = Waveform.gaussian_pulse(4e-11, 1e-11)
    r = Waveform.ricker_wavelet(3e10, 3.3333333e-11)
    srcs = ([Source(x, y, 3, FieldComponent.Ez, 1.0, g) for x in range(8) for y in range(8)] +
            [Source(1, 2, 3, FieldComponent.Ex, 0.5, r), Source(2, 2, 3, FieldComponent.Ex, 0.5, r),
             Source(4, 4, 4, FieldComponent.Ey, 0.25, g)])
    cols, vals = source_columns(srcs, 1.5e-12, 7, 300)
    assert vals.shape == (300, 3)
    assert cols.tolist() == [0] * 64 + [1, 1, 2]
    per = source_values(srcs, 1.5e-12, 7, 300)
    assert np.array_equal(vals[:, cols].view(np.uint32), per.view(np.uint32))


def test_demo_sphere_matches_oracle():
    import ctypes
    g = MaterialGrid.create(GridDims(32, 24, 40))
    g.voxelize_demo_sphere()
    o = np.zeros(32 * 24 * 40, np.uint8)
    W.lib().wfo_voxelize_demo_sphere(32, 24, 40, o.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert np.array_equal(g.ids, o) and o.sum() > 0


def test_probe_twiddles_and_spectrum():
    from wavefront_amd.probe import Probe, FieldComponent as PC, probe_tables
    import math
    p = [Probe(1, 1, 1, PC.Hx, [1e9, 2.5e9], 3e-12), Probe(0, 0, 0, PC.Ez, [7e8], 0.0)]
    sites, ct, st, act, nf = probe_tables(p, 1e-12, 2, 10)
    assert nf == 3
    for k in range(10):
        t = (np.float32(2 + k) + np.float32(1.0)) * np.float32(1e-12)
        assert act[k, 0] == (0 if t < np.float32(3e-12) else 1) and act[k, 1] == 1
        for j, f in enumerate([1e9, 2.5e9, 7e8]):
            ph = (2.0 * math.pi * float(np.float32(f))) * float(t)
            assert ct[k, j] == math.cos(ph) and st[k, j] == math.sin(ph)
    q = p[0]
    q.sum_cos[:] = [3.0, -1.0]
    q.sum_sin[:] = [4.0, 0.0]
    q.samples_recorded = 2
    sp = q.spectrum()
    assert sp[0].amplitude == 5.0 and sp[0].phase_rad == math.atan2(3.0, 4.0)


def test_aosoa_layout_helpers_roundtrip():
    d = GridDims(16, 24, 8)
    x = np.random.default_rng(0).standard_normal((6, 8, 24, 16)).astype(np.float32)
    a = xyz_to_aosoa(x, d)
    # block (bx=1, by=2, bz=0), local (lx=3, ly=4, lz=5) -> x=11, y=20, z=5
    b = (0 * 3 + 2) * 2 + 1
    assert a[b, 4, (5 * 8 + 4) * 8 + 3] == x[4, 5, 20, 11]
    assert np.array_equal(aosoa_to_xyz(a, d), x)


def test_workload_definitions():
    w = workloads.plain_64()
    assert np.float32(w.dt).view(np.uint32) == np.float32(1.9065752e-12).view(np.uint32)
    assert w.pml().shell_blocks == (2, 2, 2)
    assert len(w.probes) == 6 and w.num_steps == 200
    w3 = workloads.random_1024(n=128)
    assert w3.notes["materials"] <= 255

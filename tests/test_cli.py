"""The `wavefront` driver (src/main.rs): flags, defaults, messages and exit
codes, checked on the CPU up to the point where the engine needs a device."""
import os
import subprocess

import numpy as np
import pytest

from oracle import wfo
from tests import cli_expect as X

pytestmark = pytest.mark.skipif(not os.path.exists(X.CLI), reason="CLI not built")


def _run(args, cwd, env=None):
    e = dict(os.environ)
    e["HIP_VISIBLE_DEVICES"] = e.get("HIP_VISIBLE_DEVICES", "")
    if env:
        e.update(env)
    p = subprocess.run([X.CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60, env=e)
    return p.returncode, p.stderr.splitlines()


def test_help(tmp_path):
    rc, err = _run(["--help"], tmp_path)
    assert rc == 0
    assert err[0] == "wavefront -- out-of-core 3D FDTD electromagnetic simulator"
    assert any("--pml-thickness <N>    absorbing boundary depth in voxels, 0 disables it [default: 8]"
               in l for l in err)


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "wavefront: unrecognized argument: --bogus"),
    (["--nx"], "wavefront: missing value for --nx"),
    (["--nx", "abc"], "wavefront: invalid integer for --nx: abc"),
    (["--nx", "-8"], "wavefront: invalid integer for --nx: -8"),
    (["--dt", "1e"], "wavefront: invalid number for --dt: 1e"),
    (["--source", "foo"], "wavefront: --source: expected key=value, got 'foo'"),
    (["--source", "x=1,colour=red"], "wavefront: --source: unknown key 'colour'"),
    (["--source", "x=q"], "wavefront: --source: invalid x 'q'"),
    (["--source-component", "hx"], "wavefront: invalid source component: hx (expected ex, ey, or ez)"),
    (["--source-waveform", "square"],
     "wavefront: invalid waveform: square (expected gaussian, sinusoid, or ricker)"),
    (["--probe", "x=1,freq=1e9;zz"], "wavefront: --probe: invalid freq 'zz'"),
    (["--probe-component", "q"],
     "wavefront: invalid probe component: q (expected ex, ey, ez, hx, hy, or hz)"),
    (["--probe-freq", "1e9,x"], "wavefront: invalid number for --probe-freq: x"),
])
def test_parse_errors(tmp_path, args, msg):
    rc, err = _run(args, tmp_path)
    assert rc == 1
    assert err[0] == msg
    assert err[1] == ""
    assert err[2] == "wavefront -- out-of-core 3D FDTD electromagnetic simulator"


def test_grid_must_tile(tmp_path):
    rc, err = _run(["--nx", "12"], tmp_path)
    assert rc == 101
    assert "grid dimensions must be multiples of BLOCK_DIM (8)" in err[-1]


@pytest.mark.parametrize("args,msg", [
    (["--probe-x", "1", "--probe-freq", "1e9"],
     "wavefront: each probe requires x, y, and z (via --probe-x/-y/-z, or x=/y=/z= in --probe)"),
    (["--probe", "x=1,y=1,z=1"],
     "wavefront: each probe requires at least one frequency (via --probe-freq, or freq= in --probe)"),
    (["--probe", "x=64,y=1,z=1,freq=1e9"],
     "wavefront: probe position (64, 1, 1) is outside the 64x64x64 domain"),
])
def test_probe_errors(tmp_path, args, msg):
    rc, err = _run(args, tmp_path)
    assert rc == 1
    assert err[-1] == msg


def test_missing_scene(tmp_path):
    rc, err = _run(["--scene", "/nonexistent/scene.txt"], tmp_path)
    assert rc == 1
    assert err[-1] == ('wavefront: failed to read scene file "/nonexistent/scene.txt": '
                       "No such file or directory (os error 2)")


def test_bad_material_path(tmp_path):
    rc, err = _run(["--materials", str(tmp_path / "no" / "m.grid")], tmp_path)
    assert rc == 1
    assert err[-1].startswith(f'wavefront: failed to create material grid at "{tmp_path}/no/m.grid": '
                              "No such file or directory")


def test_setup_lines_and_material_file(tmp_path):
    """Header, source and probe lines in the reference's format, and the
    material grid file (demo sphere) written before the run starts; without
    a device the run itself then fails with the engine's message."""
    args = ["--nx", "32", "--ny", "24", "--nz", "40", "--steps", "7", "--pml-thickness", "10",
            "--source-waveform", "gaussian", "--source", "x=3,y=4,z=5,component=ex,waveform=sinusoid,freq=2.5e9",
            "--source", "waveform=ricker,amplitude=2",
            "--probe-x", "1", "--probe-y", "2", "--probe-z", "3", "--probe-freq", "1e10,2e10",
            "--probe", "x=4,y=5,z=6,component=hy,freq=3e10;6e10,start=1e-11"]
    rc, err = _run(args, tmp_path)
    dt = 1.5e-12
    assert err[0] == X.header_line(32, 24, 40, 1e-3, dt, 7, 10)
    assert err[1] == X.default_source_line(32, 24, 40, dt, waveform="gaussian")
    assert err[2] == X.default_source_line(32, 24, 40, dt, waveform="sinusoid", freq=2.5e9, comp="Ex",
                                           x=3, y=4, z=5)
    assert err[3] == X.default_source_line(32, 24, 40, dt, waveform="ricker")
    assert err[4] == "wavefront: probe at (1, 2, 3) tracking Ez"
    assert err[5] == "wavefront: probe at (4, 5, 6) tracking Hy"
    ids = np.fromfile(tmp_path / "materials.grid", np.uint8)
    ref = np.zeros(32 * 24 * 40, np.uint8)
    import ctypes
    wfo.lib().wfo_voxelize_demo_sphere(32, 24, 40, ref.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert np.array_equal(ids, ref)
    if rc != 0:
        assert err[-1].startswith("wavefront: simulation run failed: ")


def test_scene_file_materials(tmp_path):
    text = "sphere 4.0 1.0 0.0 8 8 8 5\nbox 2.25 1.0 0.01 0 0 0 6 20 3\n"
    (tmp_path / "s.txt").write_text(text)
    rc, err = _run(["--nx", "16", "--ny", "24", "--nz", "16", "--scene", "s.txt"], tmp_path)
    assert err[1] == 'wavefront: loaded scene "s.txt" (2 distinct materials)'
    ids = np.fromfile(tmp_path / "materials.grid", np.uint8)
    oids, _, n = wfo.scene_voxelize(text, 16, 24, 16, 1.5e-12, 1e-3)
    assert n == 2
    assert np.array_equal(ids, oids.reshape(-1))


def test_rust_float_formatting_examples():
    """The expectation helper itself, on values whose Rust output is known."""
    assert X.rust_debug_f32(3e-11) == "3e-11"
    assert X.rust_debug_f32(1.5e-11) == "1.5e-11"
    assert X.rust_debug_f32(33333334000.0) == "33333334000.0"
    assert X.rust_debug_f32(2.5e9) == "2500000000.0"
    assert X.rust_debug_f32(0.001) == "0.001"
    assert X.rust_debug_f32(1.0) == "1.0"
    assert X.rust_debug_f32(1e16) == "1e16"
    assert X.rust_exp(1e-3, 3) == "1.000e-3"
    assert X.rust_exp(12345.0, 4) == "1.2345e4"

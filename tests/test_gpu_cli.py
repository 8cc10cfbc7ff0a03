"""The `wavefront` driver end to end on the GPU: the snapshot file and the
printed spectra equal what the CPU oracle gives for the same run."""
import os
import subprocess

import numpy as np
import pytest

from tests import cli_expect as X
from tests.helpers import assert_bitwise

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(X.CLI), reason="CLI not built")]


@pytest.mark.parametrize("pml", [8, 0])
def test_cli_demo_run_matches_oracle(tmp_path, pml):
    n, steps, every = 64, 45, 20
    dt, dx = np.float32(1.5e-12), np.float32(1e-3)
    args = ["--steps", str(steps), "--snapshot-every", str(every), "--pml-thickness", str(pml),
            "--probe-x", "20", "--probe-y", "32", "--probe-z", "32", "--probe-freq", "1e10,3e10",
            "--probe", "x=40,y=30,z=33,component=hz,freq=2e10,start=2e-11"]
    print(f"[gpu-case] cli demo 64^3 steps={steps} every={every} pml={pml}", flush=True)
    p = subprocess.run([X.CLI] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    err = p.stderr.splitlines()
    assert p.returncode == 0, p.stderr

    ids, sim = X.demo_oracle(n, n, n, float(dt), float(dx), pml)
    f32 = np.float32
    f0 = f32(1.0) / (f32(20.0) * dt)
    sim.add_source(32, 32, 32, 2, 1.0, 2, float(f0), float(f32(1.0) / f0))
    sim.add_probe(20, 32, 32, 2, [1e10, 3e10], 0.0)
    sim.add_probe(40, 30, 33, 5, [2e10], 2e-11)
    per = n * n * n * 6
    data = np.fromfile(tmp_path / "wave_trajectory.bin", np.float32)
    snap_steps = [k for k in range(steps) if k % every == 0]
    assert data.size == per * len(snap_steps)
    done = 0
    for i, k in enumerate(snap_steps):
        sim.run(k + 1 - done, float(dt), first_step=done, threads=4)
        done = k + 1
        assert_bitwise(data[i * per:(i + 1) * per], sim.read_fields().reshape(-1), f"snapshot {i}")
    sim.run(steps - done, float(dt), first_step=done, threads=4)
    sc, ss, cnt = sim.probe_sums()
    assert np.array_equal(np.fromfile(tmp_path / "materials.grid", np.uint8), ids)

    i = err.index("wavefront: probe (20, 32, 32) Ez frequency response:")
    assert err[i + 1:i + 3] == X.spectrum_lines([1e10, 3e10], sc[:2], ss[:2], cnt[0])
    j = err.index("wavefront: probe (40, 30, 33) Hz frequency response:")
    assert err[j + 1:j + 2] == X.spectrum_lines([2e10], sc[2:], ss[2:], cnt[1])
    run_line = [l for l in err if l.startswith(f"wavefront: {steps} steps over 1 slab(s) in ")]
    assert len(run_line) == 1 and run_line[0].endswith(" steps/s)")

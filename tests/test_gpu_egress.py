"""Snapshot egress (wf_run, the twin of engine::run in src/engine.rs:653-819):
the snapshot file holds, back to back, the reference-layout field state after
every step whose loop index k has k % snapshot_every == 0, bit-identical to
the CPU oracle's state at those steps; write errors surface as WF_ERR_IO."""
import os

import numpy as np
import pytest

from tests.helpers import assert_bitwise, oracle_sim_for
from wavefront_amd import workloads
from wavefront_amd._lib import WF_ERR_INVALID, WF_ERR_IO, WavefrontError
from wavefront_amd.engine import DeviceEngine, EngineConfig, run
from wavefront_amd.layout import CoeffGrid, MaterialGrid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg} kernel={os.environ.get('WF_KERNEL', 'fused')}", flush=True)


def _oracle_snapshots(w, num_steps, every):
    """The oracle's state after each snapshot step, stepping to it."""
    sim = oracle_sim_for(w)
    snaps, done = [], 0
    for k in range(num_steps):
        if k % every == 0:
            sim.run(k + 1 - done, w.dt, first_step=done, threads=4)
            done = k + 1
            snaps.append(sim.read_fields().copy())
    if done < num_steps:
        sim.run(num_steps - done, w.dt, first_step=done, threads=4)
    return snaps, sim


@pytest.mark.parametrize("kernel", ["fused", "split"])
@pytest.mark.parametrize("num_steps,every", [(13, 4), (20, 1), (30, 9), (5, 100)])
def test_snapshot_file_matches_oracle(tmp_path, monkeypatch, kernel, num_steps, every):
    monkeypatch.setenv("WF_KERNEL", kernel)
    w = workloads.plain_64(steps=num_steps)
    _say(f"egress plain_64 steps={num_steps} every={every}")
    snaps, sim = _oracle_snapshots(w, num_steps, every)
    path = tmp_path / "wave_trajectory.bin"
    cfg = EngineConfig(num_steps=num_steps, snapshot_every=every, dt=w.dt, output_path=str(path))
    final = run(None, CoeffGrid(MaterialGrid(w.dims, w.ids), w.table), w.pml(), w.sources, w.probes, cfg)
    data = np.fromfile(path, dtype=np.float32)
    per = w.dims.voxel_count() * 6
    assert data.size == per * len(snaps)
    for i, ref in enumerate(snaps):
        assert_bitwise(data[i * per:(i + 1) * per].reshape(ref.shape), ref, f"snapshot {i}")
    assert_bitwise(final, sim.read_fields(), "final state")
    oc, os_, on = sim.probe_sums()
    sc = np.concatenate([p.sum_cos for p in w.probes])
    assert_bitwise(sc, oc, "sum_cos")
    assert [p.samples_recorded for p in w.probes] == list(on)


def test_run_continues_step_count(tmp_path):
    """wf_run after wf_step: snapshot cadence counts from the call, sources
    and probes continue from the context's step."""
    w = workloads.plain_64(steps=24)
    _say("egress after 7 eager steps")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, 24)
    eng.set_probes(w.probes, w.dt, 24)
    eng.step(7)
    path = tmp_path / "snap.bin"
    secs = eng.run(17, 8, str(path))
    assert secs > 0
    assert eng.step_count() == 24
    sim = oracle_sim_for(w)
    sim.run(8, w.dt, first_step=0, threads=4)
    s0 = sim.read_fields().copy()          # loop index 0 = step 8
    sim.run(8, w.dt, first_step=8, threads=4)
    s1 = sim.read_fields().copy()          # loop index 8 = step 16
    sim.run(8, w.dt, first_step=16, threads=4)
    s2 = sim.read_fields().copy()          # loop index 16 = step 24
    data = np.fromfile(path, dtype=np.float32).reshape(3, *s0.shape)
    assert_bitwise(data[0], s0, "snapshot 0")
    assert_bitwise(data[1], s1, "snapshot 1")
    assert_bitwise(data[2], s2, "snapshot 2")
    assert_bitwise(eng.read_fields(), s2, "final")
    eng.close()


def test_run_reduced_1024_class(tmp_path):
    """A 256^3 random-material run with 20-cell CPML: 40 steps, a snapshot
    every 10 (four 384 MiB snapshots through the pinned chunk ring), checked
    against a second context stepped eagerly to the same steps."""
    w = workloads.random_1024(steps=40, n=256)
    _say("egress random 256^3 steps=40 every=10")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, 40)
    eng.set_probes(w.probes, w.dt, 40)
    path = tmp_path / "big.bin"
    eng.run(40, 10, str(path))
    per = w.dims.voxel_count() * 24
    assert os.path.getsize(path) == 4 * per
    with open(path, "rb") as f:
        f.seek(3 * per)
        last = np.frombuffer(f.read(per), dtype=np.float32)
    final = eng.read_fields().reshape(-1)
    # snapshot 3 is taken after loop index 30, i.e. after 31 steps
    eng2 = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng2.set_sources(w.sources, w.dt, 40)
    eng2.set_probes(w.probes, w.dt, 40)
    eng2.step(31)
    assert_bitwise(last, eng2.read_fields().reshape(-1), "snapshot 3 vs eager step 31")
    eng2.step(9)
    assert_bitwise(final, eng2.read_fields().reshape(-1), "final vs eager")
    eng.close()
    eng2.close()


def test_run_errors(tmp_path, monkeypatch):
    w = workloads.plain_64(steps=4)
    _say("egress error paths")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, 4)
    with pytest.raises(WavefrontError) as e:
        eng.run(2, 0, str(tmp_path / "a.bin"))
    assert e.value.code == WF_ERR_INVALID
    with pytest.raises(WavefrontError) as e:
        eng.run(2, 1, str(tmp_path / "missing_dir" / "a.bin"))
    assert e.value.code == WF_ERR_IO
    assert eng.step_count() == 0
    monkeypatch.setenv("WF_FAULT_INJECT", "short_write")
    with pytest.raises(WavefrontError) as e:
        eng.run(3, 1, str(tmp_path / "b.bin"))
    assert e.value.code == WF_ERR_IO
    assert "short snapshot write" in str(e.value)
    monkeypatch.delenv("WF_FAULT_INJECT")
    # the context stays usable; the table still covers one more step
    eng.run(1, 1, str(tmp_path / "c.bin"))
    assert os.path.getsize(tmp_path / "c.bin") == w.dims.voxel_count() * 24
    eng.close()

"""Batched scenes: many independent scenes advanced by one launch per phase
must give every scene exactly the state of stepping it alone (the CPU
oracle's solo run, and the GPU engine's solo run)."""
import numpy as np
import pytest

from tests.helpers import assert_bitwise, oracle_sim_for
from wavefront_amd import workloads
from wavefront_amd.engine import DeviceBatch, DeviceEngine

pytestmark = pytest.mark.gpu


def _engine(w, steps):
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, steps)
    eng.set_probes(w.probes, w.dt, steps)
    return eng


@pytest.mark.parametrize("schedule", ["twostep", "split"])
def test_batch_matches_oracle_with_tuning(schedule, monkeypatch):
    """The first six scenes of the 96-scene set plus three hand-picked shapes
    (rows of 10 quads and of 66 quads, split in two fitted segments of 33),
    48 steps.  Split steps: 13 tuner candidates x 3 real steps each, then 9
    more.  Two steps per pass (nothing to tune): 5 steps (two passes and a
    split step), then 43 from the other state (21 passes, a split step).
    Scenes 6, 8 (x all shell) and 7 (y shells meeting) take the
    whole-scene every-axis CPML box, the others the region split."""
    monkeypatch.setenv("WF_BATCH_TWOSTEP", "1" if schedule == "twostep" else "0")
    ws = workloads.batch_scenes(nscenes=6, steps=48)
    ws += workloads.batch_scenes(nscenes=3, steps=48, seed=11,
                                 edges=[(32, 48, 40), (200, 32, 64), (40, 24, 32)], max_edge=264)
    ws[-1] = workloads.batch_scenes(nscenes=1, steps=48, seed=12, edges=[(264, 16, 24)],
                                    max_edge=264)[0]
    print(f"[gpu-case] batch of 9 vs oracle, 48 steps, {schedule}: "
          + ", ".join(w.name for w in ws), flush=True)
    engs = [_engine(w, 48) for w in ws]
    with DeviceBatch(engs) as b:
        two, pass_bytes = b.two_step()
        assert two == (schedule == "twostep")
        z, lpr, grid, taken = b.tune(3)
        if two:
            assert taken == 0 and pass_bytes > 49 * sum(w.cells() for w in ws)
            b.step(5)
            b.sync()
            assert [e.step_count() for e in engs] == [5] * len(engs)
            b.step(43)
        else:
            assert pass_bytes == 0
            assert taken == 39 and z >= 1 and lpr in (0, 16, 64) and grid in (0, 2048, 8192)
            b.step(48 - taken)
        b.sync()
    for w, eng in zip(ws, engs):
        assert eng.step_count() == 48
        sim = oracle_sim_for(w)
        sim.run(48, w.dt, threads=8)
        assert_bitwise(eng.read_fields(), sim.read_fields(), f"{w.name} fields")
        assert_bitwise(eng.read_psi(), sim.read_psi(), f"{w.name} psi")
        sc, ss, n = eng.probe_sums()
        oc, os_, on = sim.probe_sums()
        assert_bitwise(sc, oc, f"{w.name} sum_cos")
        assert_bitwise(ss, os_, f"{w.name} sum_sin")
        assert list(n) == list(on)
        eng.close()


def test_batch_96_full_run_matches_oracle():
    """The whole batched workload: all 96 scenes, 500 steps in one batch, the
    schedule autotuned on the first steps as the bench does, then every
    scene's fields, psi and probe sums bit for bit against its own CPU run
    (the oracle's solo engine::run, src/engine.rs:1008-1099 decomposition
    invariance; 71 G cell updates, the scenes spread over 16 host threads,
    largest first)."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    steps = 500
    ws = workloads.batch_scenes(steps=steps)
    print(f"[gpu-case] batch of 96 vs oracle, {steps} steps, "
          f"{sum(w.cells() for w in ws)} cells", flush=True)
    engs = [_engine(w, steps) for w in ws]
    with DeviceBatch(engs) as b:
        _, _, _, taken = b.tune(4)
        b.step(steps - taken)
        b.sync()
    got = []
    for eng in engs:
        assert eng.step_count() == steps
        got.append((eng.read_fields(), eng.read_psi(), eng.probe_sums()))
        eng.close()
    print("[gpu-case] batch of 96: GPU done, oracle runs", flush=True)

    def oracle(i):
        sim = oracle_sim_for(ws[i])
        sim.run(steps, ws[i].dt, threads=1)
        out = (sim.read_fields(), sim.read_psi(), sim.probe_sums())
        sim.close()
        return out

    order = sorted(range(len(ws)), key=lambda i: -ws[i].cells())
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        results = dict(zip(order, pool.map(oracle, order)))
    for i, w in enumerate(ws):
        of, op, (oc, os_, on) = results[i]
        gf, gp, (gc, gs, gn) = got[i]
        assert_bitwise(gf, of, f"{w.name} fields")
        assert_bitwise(gp, op, f"{w.name} psi")
        assert_bitwise(gc, oc, f"{w.name} sum_cos")
        assert_bitwise(gs, os_, f"{w.name} sum_sin")
        assert list(gn) == list(on) == [steps] * len(w.probes)
        assert np.abs(of).max() > 0


def test_batch_96_matches_solo_engines():
    """All 96 scenes, 12 batched steps, against each scene stepped alone."""
    ws = workloads.batch_scenes(steps=12)
    print(f"[gpu-case] batch of 96 vs solo engines, 12 steps, "
          f"{sum(w.cells() for w in ws)} cells", flush=True)
    engs = [_engine(w, 12) for w in ws]
    with DeviceBatch(engs) as b:
        b.set_tiling(8, 64)
        b.step(4)
        b.set_tiling(16, 16)
        b.step(4)
        b.set_tiling(8, 64, 1024)  # persistent: 1024 workgroups stride over the tiles
        b.step(2)
        b.set_tiling(4, 0)  # rows fitted to each scene's width
        b.step(2)
        b.sync()
    for w, eng in zip(ws, engs):
        solo = _engine(w, 12)
        solo.step(12)
        assert_bitwise(eng.read_fields(), solo.read_fields(), f"{w.name} fields")
        a, b_, c = eng.probe_sums()
        d, e, f = solo.probe_sums()
        assert_bitwise(a, d, f"{w.name} sum_cos")
        assert list(c) == list(f)
        solo.close()
        eng.close()


def test_batch_then_solo_steps_continue():
    """Batched steps, then the scene's own wf_step, then batched again."""
    ws = workloads.batch_scenes(nscenes=3, steps=20, seed=5, max_edge=96)
    print("[gpu-case] batch/solo interleave", flush=True)
    engs = [_engine(w, 20) for w in ws]
    with DeviceBatch(engs) as b:
        b.step(6)
        engs[1].step(5)      # scene 1 alone (graph path needs >= 8, eager here)
        engs[0].step(9)      # graph path
        b.set_tiling(4, 32)
        with pytest.raises(Exception):
            b.step(6)        # scene 0's tables end at step 20: 15 + 6 > 20
        b.step(5)
        b.sync()
    steps = [eng.step_count() for eng in engs]
    assert steps == [20, 16, 11]
    for w, eng, k in zip(ws, engs, steps):
        sim = oracle_sim_for(w)
        sim.run(k, w.dt, threads=8)
        assert_bitwise(eng.read_fields(), sim.read_fields(), f"{w.name} fields")
        eng.close()


@pytest.mark.parametrize("schedule", ["twostep", "split"])
def test_batch_many_sources_list_order(schedule, monkeypatch):
    """A batched scene with 3000 sources (an Ez sheet, scattered points and
    voxels named several times, soft and hard mixed, the three E
    components), applied in parallel per (component, voxel) group by the
    batched point kernels and, on two-step passes, inside the two-step
    kernels: 11 steps (five passes and a split step, or split steps only)
    equal the oracle's list-order loop; a CPML scene beside it."""
    from oracle import wfo
    from wavefront_amd.layout import GridDims
    from wavefront_amd.source import FieldComponent as SC, Source, Waveform
    monkeypatch.setenv("WF_BATCH_TWOSTEP", "1" if schedule == "twostep" else "0")
    steps = 11
    rng = np.random.default_rng(21)
    d = GridDims(32, 24, 40)
    dt, dx = 1.5e-12, 1e-3
    comps = (SC.Ex, SC.Ey, SC.Ez)
    forms = (Waveform.sinusoid(3e10), Waveform.ricker_wavelet(2e10, 5e-11),
             Waveform.gaussian_pulse(2e-11, 1e-11))
    srcs = [Source(x, y, 9, SC.Ez, 0.5, forms[0]) for y in range(d.ny) for x in range(d.nx)]
    pts = [(int(rng.integers(d.nx)), int(rng.integers(d.ny)), int(rng.integers(d.nz)))
           for _ in range(300)]
    while len(srcs) < 3000:
        x, y, z = pts[int(rng.integers(len(pts)))] if rng.random() < 0.7 else (
            int(rng.integers(d.nx)), 5, int(rng.integers(d.nz)))
        srcs.append(Source(x, y, z, comps[int(rng.integers(3))], float(rng.uniform(-2, 2)),
                           forms[int(rng.integers(3))], hard=bool(rng.random() < 0.2)))
    srcs = [srcs[i] for i in rng.permutation(len(srcs))]
    print(f"[gpu-case] batch many sources n={len(srcs)} {schedule}", flush=True)
    sheet = DeviceEngine(d, None, None, None, dt=dt, dx=dx)
    sheet.set_sources(srcs, dt, steps)
    w = workloads.batch_scenes(nscenes=1, steps=steps, seed=7, edges=[(48, 40, 56)])[0]
    other = _engine(w, steps)
    with DeviceBatch([sheet, other]) as b:
        assert b.two_step()[0] == (schedule == "twostep")
        b.step(steps)
        b.sync()
    sim = wfo.OracleSim(d.nx, d.ny, d.nz, table=wfo.vacuum_table(dt, dx))
    for s in srcs:
        sim.add_source(s.x, s.y, s.z, int(s.component), s.amplitude, s.waveform.kind,
                       s.waveform.p0, s.waveform.p1, hard=s.hard)
    sim.run(steps, dt)
    assert_bitwise(sheet.read_fields(), sim.read_fields(), "sheet scene fields")
    sim.close()
    sim = oracle_sim_for(w)
    sim.run(steps, w.dt, threads=4)
    assert_bitwise(other.read_fields(), sim.read_fields(), f"{w.name} fields")
    sim.close()
    sheet.close()
    other.close()

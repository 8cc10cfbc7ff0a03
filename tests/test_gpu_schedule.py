"""Schedules of a single engine: the two-step regions, the two-step rows
(the batched two-step kernel over the scene's own full x rows) and single
steps each report their kernels and algorithmic bytes and give the CPU
oracle's results bit for bit; without WF_SCHEDULE a small grid picks one of
them by timing at wf_create, and its results are the same."""
import pytest

from tests.helpers import assert_bitwise, oracle_sim_for
from wavefront_amd import workloads
from wavefront_amd.engine import DeviceEngine

pytestmark = pytest.mark.gpu

# 96 x 72 x 64 with CPML 16: region split and rows both apply (x interior
# 56 cells); 40 x 48 x 32 with CPML 16: x all shell, rows take one box.
SHAPES = [(96, 72, 64), (40, 48, 32)]


def _workload(shape, steps):
    return workloads.batch_scenes(nscenes=1, steps=steps, seed=21, edges=[shape],
                                  max_edge=max(shape))[0]


def _run(w, steps):
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, steps)
    eng.set_probes(w.probes, w.dt, steps)
    return eng


def _check(w, eng, steps):
    sim = oracle_sim_for(w)
    sim.run(steps, w.dt, threads=8)
    assert_bitwise(eng.read_fields(), sim.read_fields(), f"{w.name} fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), f"{w.name} psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, f"{w.name} sum_cos")
    assert_bitwise(ss, os_, f"{w.name} sum_sin")
    assert list(n) == list(on)
    sim.close()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("schedule", ["0", "1", "2"])
def test_forced_schedule(shape, schedule, monkeypatch):
    monkeypatch.setenv("WF_SCHEDULE", schedule)
    steps = 13
    w = _workload(shape, steps)
    print(f"[gpu-case] schedule {schedule} {w.name}", flush=True)
    eng = _run(w, steps)
    assert eng.schedule() == (int(schedule), [-1.0, -1.0, -1.0])
    v = eng.kernel_variant()
    if schedule == "1":
        assert v == "twostep-rows"
    elif schedule == "2":
        assert v == "fused"
    else:
        # the regions where the geometry allows them, else single steps
        assert v == ("twostep" if shape == (96, 72, 64) else "fused")
    kb = eng.kernel_bytes()
    cells = w.cells()
    if v.startswith("twostep"):
        # a pass over every cell: 49 B per cell for two steps, plus psi
        assert sum(kb.values()) > 49 * cells
    km = eng.time_kernels(4)
    assert set(km) == set(kb) and max(km.values()) > 0.0
    eng.step(steps - 4)
    eng.sync()
    assert eng.step_count() == steps
    _check(w, eng, steps)
    eng.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_timed_schedule(shape, monkeypatch):
    monkeypatch.delenv("WF_SCHEDULE", raising=False)
    steps = 11
    w = _workload(shape, steps)
    print(f"[gpu-case] timed schedule {w.name}", flush=True)
    eng = _run(w, steps)
    k, ms = eng.schedule()
    timed = [m for m in ms if m >= 0.0]
    # the regions are timed only where the geometry allows them
    assert (ms[0] >= 0.0) == (shape == (96, 72, 64)) and ms[1] > 0.0 and ms[2] > 0.0
    assert ms[k] == min(timed)
    assert eng.kernel_variant() == ("twostep", "twostep-rows", "fused")[k]
    assert eng.step_count() == 0
    eng.step(steps)
    eng.sync()
    _check(w, eng, steps)
    eng.close()

"""GPU engine vs the CPU oracle: bit-exact field, psi and probe state.

Every comparison is bitwise (FP32 with contraction off on both sides; f64
probe sums).  Each test prints its case before touching the GPU."""
import csv
import os

import numpy as np
import pytest

from oracle import wfo
from tests.helpers import assert_bitwise, oracle_pml, oracle_sim_for
from wavefront_amd.engine import PERIODIC_X, PERIODIC_Y, PERIODIC_Z, DeviceEngine
from wavefront_amd.layout import GridDims, MaterialTable, PmlConfig, PmlContext, PmlProfile1D
from wavefront_amd import workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(autouse=True, params=["fused", "rows", "onestep", "split", "single"])
def kernel_variant(request, monkeypatch):
    """Every case runs on each step-kernel scheme: two steps per pass on the
    regions (the two-step kernel variant of each region's CPML axes; the
    schedule of large grids), two steps per pass in full rows through the
    batched two-step kernel (one launch per CPML variant), one step per
    pass (the fused interior kernel with the CPML shell in separate H and E
    frame kernels; what odd step counts end with), the in-place H and E
    kernels over the whole domain, and the fused kernel over the whole
    domain, shell included (opt-in).  The schedule is set, not timed."""
    monkeypatch.setenv("WF_KERNEL", "split" if request.param == "split" else "fused")
    monkeypatch.setenv("WF_PML_FUSED", "1" if request.param == "single" else "0")
    monkeypatch.setenv("WF_TWOSTEP", "0" if request.param == "onestep" else "1")
    monkeypatch.setenv("WF_SCHEDULE", {"rows": "1", "onestep": "2"}.get(request.param, "0"))
    return request.param


def _say(msg):
    print(f"[gpu-case] {msg} kernel={os.environ.get('WF_KERNEL')} "
          f"single_pass={os.environ.get('WF_PML_FUSED')} "
          f"two_step={os.environ.get('WF_TWOSTEP')} "
          f"schedule={os.environ.get('WF_SCHEDULE')}", flush=True)


def _run_workload_gpu(w, steps):
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, steps)
    eng.set_probes(w.probes, w.dt, steps)
    eng.step(steps)
    eng.sync()
    return eng


def test_plain_64_bit_exact():
    """64^3 vacuum, 16-cell CPML, Gaussian point source, 6 probes, 200 steps:
    fields, psi and probe accumulators identical to the CPU engine."""
    w = workloads.plain_64()
    _say(f"plain_64 steps={w.num_steps}")
    eng = _run_workload_gpu(w, w.num_steps)
    sim = oracle_sim_for(w)
    sim.run(w.num_steps, w.dt, threads=4)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, "sum_cos")
    assert_bitwise(ss, os_, "sum_sin")
    assert list(n) == list(on) == [w.num_steps] * 6
    # the wave reached the probes
    assert np.abs(sc).max() > 0
    eng.close()


def _random_case(dims, shell, periodic, steps, seed, nmat=5, pml=True):
    rng = np.random.default_rng(seed)
    dt, dx = workloads.DT_COURANT_099, 1e-3
    table = MaterialTable.vacuum_filled(dt, dx)
    otable = wfo.vacuum_table(dt, dx)
    for m in range(1, nmat):
        eps, sig = float(rng.uniform(1, 10)), float(rng.uniform(0, 0.05))
        table.set_material(m, eps, 1.0, sig, dt, dx)
        otable[m] = wfo.material_coeffs(eps, 1.0, sig, dt, dx)
    ids = rng.integers(0, nmat, dims.voxel_count(), dtype=np.uint8)
    fields = rng.standard_normal((dims.block_count(), 6, 512)).astype(np.float32)
    # exercise signed zeros and exact cancellations too
    fields[rng.random(fields.shape) < 0.05] = 0.0
    fields[rng.random(fields.shape) < 0.02] = -0.0
    ctx = None
    oshell = oprof = None
    if pml:
        n = (dims.nx, dims.ny, dims.nz)
        profs = [PmlProfile1D.build(n[a], PmlConfig(), dt, dx, 8 * shell[a]) for a in range(3)]
        ctx = PmlContext(profs, shell)
        oshell = shell
        oprof = np.concatenate([wfo.pml_profile(n[a], 8 * shell[a], dt, dx).reshape(-1)
                                for a in range(3)])
    eng = DeviceEngine(dims, ids, table, ctx, periodic=periodic)
    eng.write_fields(fields)
    sim = wfo.OracleSim(dims.nx, dims.ny, dims.nz, ids=ids, table=otable, shell_blocks=oshell,
                        profiles=oprof, periodic=periodic)
    sim.write_fields(fields)
    eng.step(steps)
    for _ in range(steps):
        sim.step_fields(threads=2)
    return eng, sim


@pytest.mark.parametrize("dims,shell,periodic", [
    ((48, 40, 56), (1, 2, 1), 0),
    ((64, 64, 64), (2, 2, 2), 0),
    ((136, 24, 32), (3, 1, 2), 0),
    ((72, 48, 40), (1, 0, 2), PERIODIC_Y),
    ((40, 64, 48), (0, 1, 0), PERIODIC_X | PERIODIC_Z),
    ((16, 16, 24), (1, 1, 1), 0),
    ((296, 16, 16), (1, 0, 0), PERIODIC_Y | PERIODIC_Z),
    # x at least 512 with an x CPML: the fused variant folds the x shell into
    # two 256-cell edge columns of the single pass (no middle columns at 512)
    ((512, 24, 32), (1, 1, 1), 0),
    ((584, 32, 24), (2, 1, 1), 0),
    ((528, 24, 40), (3, 0, 2), PERIODIC_Y),
])
def test_random_fields_pml(dims, shell, periodic):
    """Random fields, random lossy materials, CPML on uneven per-axis shells:
    every voxel class (interior, faces, edges, corners, identity axes inside
    shell blocks, signed zeros) after several steps."""
    d = GridDims(*dims)
    _say(f"random_fields_pml dims={dims} shell={shell} periodic={periodic}")
    eng, sim = _random_case(d, shell, periodic, steps=5, seed=sum(dims))
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    eng.close()


@pytest.mark.parametrize("periodic", [0, PERIODIC_X, PERIODIC_Y, PERIODIC_Z, 7])
def test_random_fields_plain(periodic):
    d = GridDims(80, 32, 24)
    _say(f"random_fields_plain periodic={periodic}")
    eng, sim = _random_case(d, None, periodic, steps=4, seed=11 + periodic, pml=False)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    eng.close()


def test_zchunk_invariance(monkeypatch):
    """The z-chunking of the launch grids (interior and frame) must not change
    any value."""
    d = GridDims(64, 32, 96)
    outs = []
    for zc, fz in (("1", "5"), ("7", "1"), ("96", "96")):
        monkeypatch.setenv("WF_ZCHUNK", zc)
        monkeypatch.setenv("WF_FRAME_ZCHUNK", fz)
        _say(f"zchunk={zc} frame zchunk={fz}")
        eng, sim = _random_case(d, (1, 1, 2), 0, steps=3, seed=5)
        outs.append(eng.read_fields())
        assert_bitwise(outs[-1], sim.read_fields(), f"zchunk {zc}/{fz} vs the CPU engine")
        eng.close()
    assert_bitwise(outs[0], outs[1], "zchunk 1/5 vs 7/1")
    assert_bitwise(outs[0], outs[2], "zchunk 1/5 vs 96/96")


def test_field_placement_probe_leaves_the_initial_state(monkeypatch):
    """The field-placement probe of wf_create, forced on a small grid with
    three extra state buffers: every arrangement is timed (each extra buffer
    as state A and, with two states, as state B), the fastest is kept, and
    the engine then starts from the initial state (zero fields and psi, no
    negative zeros, step 0), so 40 steps of plain_64 with its sources and
    probes stay bit-identical to the CPU engine."""
    monkeypatch.setenv("WF_PLACEMENT", "3")
    monkeypatch.setenv("WF_PLACEMENT_MIN_MIB", "0")
    w = workloads.plain_64()
    steps = 40
    _say(f"placement probe plain_64 steps={steps}")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    pl = eng.placement()
    states = 1 if eng.kernel_variant() == "split" else 2
    assert len(pl["ms_per_step"]) == 1 + 3 * states
    assert all(t > 0 for t in pl["ms_per_step"])
    assert pl["ms_per_step"][pl["kept"]] == min(pl["ms_per_step"])
    f0, p0 = eng.read_fields(), eng.read_psi()
    assert not f0.any() and not np.signbit(f0).any()
    assert not p0.any() and not np.signbit(p0).any()
    eng.set_sources(w.sources, w.dt, steps)
    eng.set_probes(w.probes, w.dt, steps)
    eng.step(steps)
    eng.sync()
    sim = oracle_sim_for(w)
    sim.run(steps, w.dt, threads=4)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, "sum_cos")
    assert_bitwise(ss, os_, "sum_sin")
    assert list(n) == list(on) == [steps] * 6
    eng.close()
    monkeypatch.setenv("WF_PLACEMENT", "0")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    assert eng.placement() == {"ms_per_step": [], "kept": 0}
    eng.close()


def test_sources_order_soft_and_hard():
    """Several sources on one voxel: applied in list order (src/engine.rs:770)."""
    from wavefront_amd.source import FieldComponent as SC, Source, Waveform
    d = GridDims(16, 16, 16)
    dt, dx = 1.5e-12, 1e-3
    srcs = [Source(5, 6, 7, SC.Ey, 0.7, Waveform.sinusoid(3e10)),
            Source(5, 6, 7, SC.Ey, 1.3, Waveform.ricker_wavelet(2e10, 5e-11), hard=True),
            Source(5, 6, 7, SC.Ey, 0.2, Waveform.gaussian_pulse(2e-11, 1e-11)),
            Source(9, 1, 2, SC.Ex, 2.0, Waveform.sinusoid(1e10))]
    _say("sources_order")
    eng = DeviceEngine(d, None, None, None, dt=dt, dx=dx)
    eng.set_sources(srcs, dt, 12)
    eng.step(12)
    sim = wfo.OracleSim(16, 16, 16, table=wfo.vacuum_table(dt, dx))
    for s in srcs:
        sim.add_source(s.x, s.y, s.z, int(s.component), s.amplitude, s.waveform.kind,
                       s.waveform.p0, s.waveform.p1, hard=s.hard)
    sim.run(12, dt)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    eng.close()


def test_sources_replaced_mid_run():
    """plain_64's source table replaced after 10 of 40 steps (the same
    sources, their values from step 10 on) with the probe table kept: every
    schedule picks up the new table, and the run stays bit-identical to the
    CPU engine's uninterrupted one."""
    w = workloads.plain_64()
    steps, first = 40, 10
    _say(f"sources replaced mid-run plain_64 steps={steps}")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, steps)
    eng.set_probes(w.probes, w.dt, steps)
    eng.step(first)
    eng.sync()
    eng.set_sources(w.sources, w.dt, steps - first)
    eng.step(steps - first)
    eng.sync()
    sim = oracle_sim_for(w)
    sim.run(steps, w.dt, threads=4)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, "sum_cos")
    assert_bitwise(ss, os_, "sum_sin")
    assert list(n) == list(on)
    eng.close()


def test_many_sources_with_repeated_voxels():
    """3000 sources (sheets, scattered points and voxels named several times
    in the list, soft and hard mixed, the three E components) applied in
    parallel per voxel: identical to the oracle's list-order loop."""
    from wavefront_amd.source import FieldComponent as SC, Source, Waveform
    rng = np.random.default_rng(21)
    d = GridDims(32, 24, 40)
    dt, dx = 1.5e-12, 1e-3
    comps = (SC.Ex, SC.Ey, SC.Ez)
    forms = (Waveform.sinusoid(3e10), Waveform.ricker_wavelet(2e10, 5e-11),
             Waveform.gaussian_pulse(2e-11, 1e-11))
    srcs = []
    for y in range(d.ny):  # an Ez sheet at z = 9
        for x in range(d.nx):
            srcs.append(Source(x, y, 9, SC.Ez, 0.5, forms[0]))
    pts = [(int(rng.integers(d.nx)), int(rng.integers(d.ny)), int(rng.integers(d.nz)))
           for _ in range(300)]
    while len(srcs) < 3000:
        x, y, z = pts[int(rng.integers(len(pts)))] if rng.random() < 0.7 else (
            int(rng.integers(d.nx)), 5, int(rng.integers(d.nz)))
        srcs.append(Source(x, y, z, comps[int(rng.integers(3))], float(rng.uniform(-2, 2)),
                           forms[int(rng.integers(3))], hard=bool(rng.random() < 0.2)))
    order = rng.permutation(len(srcs))  # interleave the sheet with the rest
    srcs = [srcs[i] for i in order]
    _say(f"many_sources n={len(srcs)}")
    eng = DeviceEngine(d, None, None, None, dt=dt, dx=dx)
    eng.set_sources(srcs, dt, 10)
    eng.step(10)
    sim = wfo.OracleSim(d.nx, d.ny, d.nz, table=wfo.vacuum_table(dt, dx))
    for s in srcs:
        sim.add_source(s.x, s.y, s.z, int(s.component), s.amplitude, s.waveform.kind,
                       s.waveform.p0, s.waveform.p1, hard=s.hard)
    sim.run(10, dt)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    eng.close()


def test_sources_on_interior_tile_boundaries():
    """The single-pass kernel's tiles meet at inner boundaries where an edge
    lane reads (and recomputes) its neighbour tile's cells.  A 552 x 40 x 40
    domain with CPML 8 (interior box x in [12, 544): three 256-wide tiles,
    inner boundaries at x = 268 and 524), soft and hard sources of every E
    component on the cells at x = 267, 268, 523, 524, random fields written
    before the first step, 20 steps (graph replays included): fields, psi and
    probes against the oracle bit for bit."""
    from wavefront_amd.source import FieldComponent as SC, Source, Waveform
    from wavefront_amd.probe import FieldComponent as PC, Probe
    rng = np.random.default_rng(5)
    d = GridDims(552, 40, 40)
    dt, dx = 1.9065752e-12, 1e-3
    forms = (Waveform.sinusoid(3e10), Waveform.ricker_wavelet(2e10, 5e-11),
             Waveform.gaussian_pulse(2e-11, 1e-11))
    srcs = []
    for x in (267, 268, 523, 524):
        for k in range(12):
            srcs.append(Source(x, int(rng.integers(0, 40)), int(rng.integers(0, 40)),
                               (SC.Ex, SC.Ey, SC.Ez)[k % 3], float(rng.uniform(-2, 2)),
                               forms[k % 3], hard=(k % 4 == 3)))
    probes = [Probe(x, 20, 20, PC.Ez, [2e10], 0.0) for x in (267, 268, 300)]
    cfg = PmlConfig(thickness=8)
    prof = [PmlProfile1D.build(n, cfg, dt, dx, 8) for n in (d.nx, d.ny, d.nz)]
    pml = PmlContext(prof, (1, 1, 1))
    f0 = rng.standard_normal((d.block_count(), 6, 512)).astype(np.float32)
    _say("sources on tile boundaries 552x40x40")
    eng = DeviceEngine(d, None, MaterialTable.vacuum_filled(dt, dx), pml)
    eng.write_fields(f0)
    eng.set_sources(srcs, dt, 20)
    eng.set_probes(probes, dt, 20)
    eng.step(3)
    eng.step(17)
    shell, profs = oracle_pml(d, cfg, dt, dx)
    sim = wfo.OracleSim(d.nx, d.ny, d.nz, table=wfo.vacuum_table(dt, dx), shell_blocks=shell,
                        profiles=profs)
    for s_ in srcs:
        sim.add_source(s_.x, s_.y, s_.z, int(s_.component), s_.amplitude, s_.waveform.kind,
                       s_.waveform.p0, s_.waveform.p1, hard=s_.hard)
    for p_ in probes:
        sim.add_probe(p_.x, p_.y, p_.z, int(p_.component), p_.frequencies_hz, 0.0)
    sim.write_fields(f0)
    sim.run(20, dt, threads=4)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, "sum_cos")
    eng.close()


def test_points_on_two_step_margins():
    """Two steps per pass: the two-step kernel's tiles recompute a margin
    (one quad on each side in x, two rows below and one above in y, two
    planes below and one above each z slice) whose values feed the owned
    cells, and the first step's sources are added and its probes recorded
    inside the kernel; the box's boundary layers are also computed by the
    frame's first step.  A 560 x 48 x 64 domain with CPML 8: two-step box
    x in [12, 540) (tiles of 248 cells: inner boundaries at x = 260, 508),
    y in [10, 39) (9-row tiles: boundaries at 19, 28, 37), z in [10, 55)
    (slices of 2-3 planes).  Soft and hard sources of every E component and
    probes of all six components on the cells around those boundaries, on
    the box's boundary layers and just outside them; random fields first,
    then 3 + 17 steps (a pass and a single step, graph replays, a trailing
    single step): fields, psi, probe sums and traces against the oracle."""
    from wavefront_amd.source import FieldComponent as SC, Source, Waveform
    from wavefront_amd.probe import FieldComponent as PC, Probe
    rng = np.random.default_rng(17)
    d = GridDims(560, 48, 64)
    dt, dx = 1.9065752e-12, 1e-3
    forms = (Waveform.sinusoid(3e10), Waveform.ricker_wavelet(2e10, 5e-11),
             Waveform.gaussian_pulse(2e-11, 1e-11))
    xs = (11, 12, 15, 16, 255, 256, 259, 260, 263, 264, 503, 504, 507, 508, 511, 512, 535, 536,
          539, 540)
    ys = (9, 10, 11, 17, 18, 19, 20, 27, 28, 29, 36, 37, 38, 39)
    zs = (9, 10, 11, 12, 13, 20, 33, 53, 54, 55)
    srcs, probes = [], []
    for k in range(160):
        x, y, z = (int(rng.choice(xs)), int(rng.choice(ys)), int(rng.choice(zs)))
        srcs.append(Source(x, y, z, (SC.Ex, SC.Ey, SC.Ez)[k % 3], float(rng.uniform(-2, 2)),
                           forms[k % 3], hard=(k % 5 == 4)))
        if k % 4 == 0:  # the same voxel named again, later in the list
            srcs.append(Source(x, y, z, (SC.Ex, SC.Ey, SC.Ez)[k % 3], 0.5, forms[0]))
    comps = (PC.Ex, PC.Ey, PC.Ez, PC.Hx, PC.Hy, PC.Hz)
    for k in range(48):
        x, y, z = (int(rng.choice(xs)), int(rng.choice(ys)), int(rng.choice(zs)))
        probes.append(Probe(x, y, z, comps[k % 6], [2e10, 3.5e10], 0.0))
    cfg = PmlConfig(thickness=8)
    prof = [PmlProfile1D.build(n, cfg, dt, dx, 8) for n in (d.nx, d.ny, d.nz)]
    pml = PmlContext(prof, (1, 1, 1))
    table = MaterialTable.vacuum_filled(dt, dx)
    otable = wfo.vacuum_table(dt, dx)
    for m in range(1, 4):
        eps, sig = float(rng.uniform(1, 10)), float(rng.uniform(0, 0.05))
        table.set_material(m, eps, 1.0, sig, dt, dx)
        otable[m] = wfo.material_coeffs(eps, 1.0, sig, dt, dx)
    ids = rng.integers(0, 4, d.voxel_count(), dtype=np.uint8)
    f0 = rng.standard_normal((d.block_count(), 6, 512)).astype(np.float32)
    steps = 20
    _say("points on two-step margins 560x48x64")
    eng = DeviceEngine(d, ids, table, pml)
    eng.write_fields(f0)
    eng.set_sources(srcs, dt, steps)
    eng.set_probes(probes, dt, steps, trace=True)
    eng.step(3)
    eng.step(17)
    shell, profs = oracle_pml(d, cfg, dt, dx)
    sim = wfo.OracleSim(d.nx, d.ny, d.nz, ids=ids, table=otable, shell_blocks=shell,
                        profiles=profs)
    for s_ in srcs:
        sim.add_source(s_.x, s_.y, s_.z, int(s_.component), s_.amplitude, s_.waveform.kind,
                       s_.waveform.p0, s_.waveform.p1, hard=s_.hard)
    for p_ in probes:
        sim.add_probe(p_.x, p_.y, p_.z, int(p_.component), p_.frequencies_hz, 0.0)
    sim.write_fields(f0)
    traces = sim.run(steps, dt, threads=4, trace=True)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, "sum_cos")
    assert_bitwise(ss, os_, "sum_sin")
    assert_bitwise(eng.read_trace(steps), traces, "probe traces")
    eng.close()


def test_scene_spheres_128():
    """The two-sphere scene scaled to 128^3 (lossy sphere, CPML 16, Ricker)."""
    w = workloads.spheres_512(steps=60, n=128)
    _say("spheres_128 steps=60")
    eng = _run_workload_gpu(w, 60)
    sim = oracle_sim_for(w)
    sim.run(60, w.dt, threads=4)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    sc, ss, n = eng.probe_sums()
    oc, os_, on = sim.probe_sums()
    assert_bitwise(sc, oc, "sum_cos")
    assert_bitwise(ss, os_, "sum_sin")
    eng.close()


def test_random_fill_128():
    """The random sphere/box fill (255-entry palette) at 128^3."""
    w = workloads.random_1024(steps=40, n=128)
    _say("random_fill_128 steps=40")
    eng = _run_workload_gpu(w, 40)
    sim = oracle_sim_for(w)
    sim.run(40, w.dt, threads=4)
    assert_bitwise(eng.read_fields(), sim.read_fields(), "fields")
    assert_bitwise(eng.read_psi(), sim.read_psi(), "psi")
    eng.close()


def test_graph_replay_matches_eager(monkeypatch):
    """The captured step graph (8 steps per replay, device step counter) gives
    the eager launches' bits after every replay, with eager work, odd step
    counts and host reads between replays."""
    w = workloads.plain_64(steps=120)
    engines = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("WF_GRAPH", mode)
        _say(f"graph_replay WF_GRAPH={mode}")
        eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
        eng.set_sources(w.sources, w.dt, w.num_steps)
        eng.set_probes(w.probes, w.dt, w.num_steps, trace=True)
        engines[mode] = eng
    done = 0
    for chunk in (8, 16, 3, 8, 9, 1, 24, 51):
        for eng in engines.values():
            eng.step(chunk)
        done += chunk
        g, e = engines["1"], engines["0"]
        assert g.step_count() == e.step_count() == done
        assert_bitwise(g.read_fields(), e.read_fields(), f"fields after {done} steps")
        sg, se = g.probe_sums(), e.probe_sums()
        assert_bitwise(sg[0], se[0], "sum_cos")
        assert list(sg[2]) == list(se[2]) == [done] * 6
    assert_bitwise(engines["1"].read_trace(done), engines["0"].read_trace(done), "trace")
    sim = oracle_sim_for(w)
    sim.run(done, w.dt, threads=4)
    assert_bitwise(engines["1"].read_fields(), sim.read_fields(), "graph vs oracle")
    for eng in engines.values():
        eng.close()


def test_write_read_roundtrip():
    d = GridDims(24, 32, 40)
    f = np.random.default_rng(3).standard_normal((d.block_count(), 6, 512)).astype(np.float32)
    _say("write_read_roundtrip")
    eng = DeviceEngine(d, None, None, None, dt=1e-12, dx=1e-3)
    eng.write_fields(f)
    assert_bitwise(eng.read_fields(), f, "roundtrip")
    eng.close()


def test_snapshot_async_holds_the_state_at_the_call():
    """wf_snapshot_async: the buffer receives the state after the steps
    queued before the call, even though more steps are queued before
    wf_wait (a second snapshot reuses the device buffer only after the first
    copy has left it); both equal wf_read_fields at those steps."""
    w = workloads.plain_64()
    _say("snapshot async")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_sources(w.sources, w.dt, 60)
    eng.step(20)
    ref20 = eng.read_fields()
    a = np.empty_like(ref20)
    b = np.empty_like(ref20)
    eng.snapshot_async(a)
    eng.step(17)
    eng.snapshot_async(b)
    eng.step(5)
    eng.wait()
    assert_bitwise(a, ref20, "snapshot at 20")
    ref42 = eng.read_fields()
    assert not np.array_equal(ref42, ref20)
    sim = oracle_sim_for(w)
    sim.run(37, w.dt, threads=8)
    assert_bitwise(b, sim.read_fields(), "snapshot at 37 vs oracle")
    eng.close()


def test_wave_arrives_at_c():
    """src/engine.rs:841-932: Ricker 30 GHz at the centre of a 64^3 vacuum
    grid, no PML; |Ez| at +20 cells first exceeds 5e-4 within 20% of 20*dx/c."""
    from wavefront_amd.probe import FieldComponent as PC, Probe
    from wavefront_amd.source import FieldComponent as SC, Source, Waveform
    d = GridDims(64, 64, 64)
    dx, dt = np.float32(1e-3), np.float32(1.5e-12)
    src = Source(32, 32, 32, SC.Ez, 1.0, Waveform.ricker_wavelet(3e10, float(np.float32(1.0) / np.float32(3e10))))
    _say("wave_arrives_at_c")
    eng = DeviceEngine(d, None, None, None, dt=float(dt), dx=float(dx))
    eng.set_sources([src], float(dt), 200)
    eng.set_probes([Probe(52, 32, 32, PC.Ez, [], 0.0)], float(dt), 200, trace=True)
    eng.step(200)
    tr = eng.read_trace(200)[:, 0]
    hits = np.flatnonzero(np.abs(tr) > 5e-4)
    assert hits.size, "wave never reached the probe point within max_steps"
    measured = (np.float32(hits[0]) + np.float32(1.0)) * dt
    expected = (np.float32(20.0) * dx) / np.float32(299792458.0)
    assert abs(measured - expected) / expected < 0.20
    eng.close()


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return list(csv.reader(f))[1:]


def test_convergence_study_on_gpu():
    """examples/convergence_study.rs on the device: measured phase velocity
    and error columns equal validation/convergence_data.csv bit for bit."""
    from wavefront_amd.studies import convergence_rows
    _say("convergence_study")
    rows = convergence_rows()
    gold = _golden("convergence_data.csv")
    for (dx, ppw, v, err), g in zip(rows, gold):
        assert np.float32(dx) == np.float32(float(g[0]))
        assert np.float32(v) == np.float32(float(g[2])), (v, g[2])
        assert np.float32(err) == np.float32(float(g[4])), (err, g[4])


def test_pml_reflection_study_on_gpu():
    """examples/pml_reflection_study.rs on the device: reflection coefficients
    equal validation/pml_reflection_data.csv exactly (f64)."""
    from wavefront_amd.studies import pml_reflection_rows
    _say("pml_reflection_study")
    got = pml_reflection_rows()
    gold = [float(r[2]) for r in _golden("pml_reflection_data.csv")]
    assert got == gold
    assert all(b < a for a, b in zip(got, got[1:]))
    assert got[-1] < 1e-2


@pytest.mark.parametrize("case", ["plain_64", "random_fill_128"])
def test_contracted_kernels_within_tolerance(case):
    """The contracted kernel set (a*b+c as fused multiply-adds): relative L2
    field error against the CPU engine <= 1e-6 after the whole run, and not
    bit-identical (the contracted kernels did run)."""
    w = workloads.plain_64() if case == "plain_64" else workloads.random_1024(steps=40, n=128)
    steps = w.num_steps if case == "plain_64" else 40
    _say(f"contracted {case} steps={steps}")
    eng = DeviceEngine(w.dims, w.ids, w.table, w.pml())
    eng.set_fp_contract(True)
    eng.set_sources(w.sources, w.dt, steps)
    eng.set_probes(w.probes, w.dt, steps)
    eng.step(steps)
    got = eng.read_fields().astype(np.float64)
    eng.close()
    sim = oracle_sim_for(w)
    sim.run(steps, w.dt, threads=4)
    ref = sim.read_fields().astype(np.float64)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"contracted {case}: relative L2 error {rel:.3e}")
    assert rel <= 1e-6
    assert not np.array_equal(got, ref)


def test_sheet_source_shares_one_value_column():
    """A 1024 x 1024 sheet of identical soft Gaussian Ez sources (1 048 576
    sources) with a 10 000-step value table: one shared column, 40 KB of
    values on the device (one column per source would be 42 GB).  20 steps
    against the oracle's per-source loop, bit for bit."""
    from wavefront_amd.source import FieldComponent as SC
    from wavefront_amd.source import Source, Waveform, source_columns
    w = workloads.Workload(name="sheet_1024", dims=GridDims(1024, 1024, 8), ids=None,
                           table=MaterialTable.vacuum_filled(1.9065752e-12, 1e-3),
                           pml_config=None, dt=1.9065752e-12, dx=1e-3, num_steps=20,
                           sources=[Source(x, y, 4, SC.Ez, 1.0,
                                           Waveform.gaussian_pulse(4.0e-11, 1.0e-11))
                                    for y in range(1024) for x in range(1024)],
                           probes=[], scene_text="")
    cols, vals = source_columns(w.sources, w.dt, 0, 10000)
    assert vals.shape == (10000, 1) and not cols.any()
    _say("sheet source 1024^2, 20 of 10000 steps")
    with DeviceEngine(w.dims, None, w.table, None) as eng:
        eng.set_sources(w.sources, w.dt, 10000)
        eng.step(20)
        got = eng.read_fields()
    sim = oracle_sim_for(w)
    sim.run(20, w.dt, threads=8)
    assert_bitwise(got, sim.read_fields(), "sheet fields")
    assert np.abs(got).max() > 0

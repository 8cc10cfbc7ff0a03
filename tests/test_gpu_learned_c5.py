"""The learned thin boundary at its 256^3 workload, and the passive limits that
keep any network's ghosts bounded.

The workload: a 256^3 FP32 vacuum domain (and the same domain with random
dielectric spheres) whose six faces end in the 4-cell learned layer: per face
the second-order Higdon update whose adaptive factor is set by a 16-64-64-1
MLP (BF16 weights on MFMA, FP32 accumulate) trained on the GPU on samples
recorded from 64-cell-CPML runs, evaluated by the reference's two-run
reflection measurement (examples/pml_reflection_study.rs:17-41, 204-224,
252-313) against 32-cell CPML and an FP64 CPU run, and by the reflected
energy of a 3D Ricker point source against 8- and 32-cell CPML and the linear
3D predictor over 3000 steps."""
import numpy as np
import pytest

from oracle import fdtd1d
from wavefront_amd import learned as L
from wavefront_amd._lib import check, lib
from wavefront_amd.layout import GridDims, MaterialGrid, MaterialTable
from wavefront_amd.scene import Scene
from wavefront_amd.source import FieldComponent as SC
from wavefront_amd.source import Source, Waveform

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _spheres(n, seed, dt, dx):
    """Random dielectric spheres clear of the faces and of the centre source."""
    rng = np.random.default_rng(seed)
    lines = []
    while len(lines) < 12:
        r = float(rng.uniform(8, 30))
        c = rng.uniform(12 + r, n - 12 - r, 3)
        if np.linalg.norm(c - n / 2) < r + 6:
            continue
        lines.append(f"sphere {rng.uniform(1, 10):.3f} 1.0 {rng.uniform(0, 0.05):.4f} "
                     f"{c[0]:.2f} {c[1]:.2f} {c[2]:.2f} {r:.2f}\n")
    grid = MaterialGrid.create(GridDims(n, n, n))
    table = MaterialTable.vacuum_filled(dt, dx)
    Scene.parse("".join(lines)).voxelize(grid, table, dt, dx)
    return grid.ids, table


def test_replayed_exact_ghosts_pass_the_limits():
    """The exact ghosts of a 64-cell-CPML run, replayed through the passive
    limits, change the thin run's fields by less than 1e-3 (relative L2; a
    per-position credit instead of the face's pooled one changed them by
    3.3x, since reactive energy leaves at one position and returns at
    another), and a limited learned run repeats bit for bit."""
    _say("replay through the limits; determinism")
    dx, dt, _ = L.study_grid()
    n, steps = 64, 700
    thin = GridDims(n, n, n)
    w = Waveform.ricker_wavelet(3e10, float(np.float32(1) / np.float32(3e10)))
    srcs = [Source(30, 34, 27, SC.Ez, 1.0, w), Source(20, 40, 33, SC.Ex, 0.7, w)]
    rec = L.record_run(thin, srcs, steps, dt, dx)
    out = []
    for limited in (False, True):
        with L.thin_engine(thin, None, dt, dx, mode="replay", replay=rec) as eng:
            if limited:
                check(lib().wf_set_replay_limits(eng.handle, 1, 0.0, 0.0))
            eng.set_sources(srcs, dt, steps)
            eng.step(steps)
            out.append(eng.read_fields().copy())
    rel = np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])
    print(f"exact ghosts through the limits: relative field change {rel:.3e}")
    assert np.abs(out[0]).max() > 0
    assert rel < 1e-3
    S = L.collect_3d(thin, runs=2, steps=400, every=10, per_face=50000, seed=5, dt=dt, dx=dx,
                     plane_runs=4)
    nets = L.train_all(S, dt, dx)
    runs = []
    for _ in range(2):
        with L.thin_engine(thin, nets, dt, dx) as eng:
            eng.set_sources(srcs, dt, steps)
            eng.step(steps)
            runs.append(eng.read_fields().copy())
    assert np.isfinite(runs[0]).all()
    assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))


def test_learned_layer_256():
    """The C5 workload with the shipped layer: per face the second-order
    Higdon update whose adaptive factor comes from the MLP, trained on the
    GPU on plane-wave samples of the 64-cell-CPML twin (32 runs per axis)
    and early-stopped on closed-loop validation.  Checks: (1) normal
    incidence: the two-run reflection is <= 1e-4 and agrees with the same
    measurement against an FP64 CPU run of the incident wave; 32-cell CPML
    gives its golden 5.38e-5.  (2) 3D, 256^3 (248^3 interior) vacuum, Ricker
    point source, 3000 steps: the reflected energy is below 8-cell CPML's
    and below the linear 3D predictor's (fitted to 3D point-source runs);
    the remaining gap to 32-cell CPML is printed.  (3) The field energy of
    the 256^3 vacuum domain and of 12 random dielectric spheres decays from
    step 300 to step 3000 and stays bounded to step 6000."""
    _say("C5: data, training with closed-loop validation")
    dx, dt, freq = L.study_grid()
    S1 = L.collect_3d(GridDims(96, 96, 96), runs=0, steps=0, every=1, per_face=200000, seed=11,
                      dt=dt, dx=dx, plane_runs=32)
    nets, table = L.train_higdon(S1, dt, dx, log=print)
    assert nets.head == L.HEAD_HIGDON and len(table) > 1
    S3 = L.collect_3d(GridDims(128, 128, 128), runs=8, steps=700, every=10, per_face=300000,
                      seed=23, dt=dt, dx=dx, plane_runs=24)
    lin3 = L.linear_all(S3, dt, dx)

    _say("C5: normal incidence")
    r_learned, test, ref = L.reflection("learned", nets)
    r_cpml32, _, _ = L.reflection("cpml", cpml_voxels=32, reference=ref)
    nx, xs, xp, steps, skip, nxr, xsr, xpr = L.reflection_geometry(L.LAYER)
    ref64 = fdtd1d.plane_wave_trace_f64(nxr, MaterialTable.vacuum_filled(dt, dx).q16[0], dt, freq,
                                        xsr, xpr, steps)
    r_fp64 = (L.quadrature_amplitude(test - ref64, dt, freq, skip) /
              L.quadrature_amplitude(ref64, dt, freq, skip))
    floor = (L.quadrature_amplitude(ref - ref64, dt, freq, skip) /
             L.quadrature_amplitude(ref64, dt, freq, skip))
    print(f"normal incidence: R learned {r_learned:.3e}, vs FP64 incident {r_fp64:.3e} "
          f"(FP32/FP64 floor {floor:.3e}), CPML-32 {r_cpml32:.3e}")
    assert abs(r_cpml32 - 5.3755640258903854e-05) < 1e-12
    assert r_learned <= 1e-4
    assert abs(r_fp64 - r_learned) < 0.1 * r_learned + 2 * floor

    _say("C5: 3D reflected energy at 248^3 interior, 3000 steps")
    refs = {}
    r3 = L.reflection_3d(nets, interior=248, steps=3000, kinds=("learned", "cpml8", "cpml32"),
                         refs=refs)
    r3["linear_3d"] = L.reflection_3d(lin3, interior=248, steps=3000, kinds=("learned",),
                                      refs=refs)["learned"]
    print(f"3D reflected / incident energy over 3000 steps: {r3}; gap to CPML-32: "
          f"{r3['learned'] / r3['cpml32']:.1f}x")
    assert r3["learned"] < r3["cpml8"]
    assert r3["learned"] < r3["linear_3d"]

    _say("C5: field energy, 256^3 vacuum and random spheres")
    dims = GridDims(256, 256, 256)
    e_vac = L.energy_history(dims, nets, dt, dx, steps=(300, 3000, 6000))
    ids, table = _spheres(256, 7, dt, dx)
    e_sph = L.energy_history(dims, nets, dt, dx, ids=ids, table=table, steps=(300, 3000, 6000))
    print(f"energy after 300 / 3000 / 6000 steps: vacuum {e_vac}, spheres {e_sph}")
    for e in (e_vac, e_sph):
        assert np.isfinite(e).all() and e[1] < e[0]
        assert e[2] < 2 * e[1] + 1e-3 * e[0]

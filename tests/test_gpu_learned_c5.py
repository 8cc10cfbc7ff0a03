"""The learned thin boundary at its 256^3 workload, and the passive limits that
keep any network's ghosts bounded.

The workload: a 256^3 FP32 vacuum domain (and the same domain with random
dielectric spheres) whose six faces end in the 4-cell learned layer, per-face
MLPs 16-64-64-1 (BF16 weights on MFMA, FP32 accumulate) trained on the GPU on
samples recorded from 64-cell-CPML runs, evaluated by the reference's two-run
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
    """The C5 workload.  Networks: per face an MLP trained on plane-wave runs
    of the 64-cell-CPML twin (normal incidence, 32 runs per axis), and the
    linear 3D predictor fitted to 3D point-source runs beside it.
    Checks: (1) normal incidence: the two-run reflection of the MLP layer is
    <= 1e-4 and agrees with the same measurement against an FP64 CPU run of
    the incident wave; 32-cell CPML gives its golden 5.38e-5.  (2) 3D, 256^3
    (248^3 interior) vacuum, Ricker point source, 3000 steps: with the passive
    limits the MLP layer's fields stay finite and its energy decays; without
    them the same networks grow without bound.  (3) The same with 12 random
    dielectric spheres: bounded and decaying.  The 3D reflected energies of
    the MLP layer, the linear predictor and 8-/32-cell CPML (against a
    64-cell CPML run) are printed: the MLP layer's is far above the linear
    predictor's and CPML-8's, the remaining gap of this boundary."""
    _say("C5: data and training")
    dx, dt, freq = L.study_grid()
    S1 = L.collect_3d(GridDims(96, 96, 96), runs=0, steps=0, every=1, per_face=200000, seed=11,
                      dt=dt, dx=dx, plane_runs=32)
    nets = L.train_all(S1, dt, dx)
    assert np.any(nets.params[:, :L.param_count() - 17])  # MLP weights in use
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

    _say("C5: 256^3 vacuum, 3000 steps, limited and unlimited")
    dims = GridDims(256, 256, 256)
    e_lim = L.energy_history(dims, nets, dt, dx)
    e_free = L.energy_history(dims, nets, dt, dx, passive=False)
    print(f"energy after 300 / 3000 steps: limited {e_lim}, unlimited {e_free}")
    assert np.isfinite(e_lim).all() and e_lim[0] > 0
    assert e_lim[1] < e_lim[0]
    assert not np.isfinite(e_free[1]) or e_free[1] > 10 * e_free[0]

    _say("C5: 256^3 random spheres, 3000 steps")
    ids, table = _spheres(256, 7, dt, dx)
    e_sph = L.energy_history(dims, nets, dt, dx, ids=ids, table=table)
    print(f"spheres: energy after 300 / 3000 steps {e_sph}")
    assert np.isfinite(e_sph).all() and e_sph[1] < e_sph[0]

    _say("C5: 3D reflected energy at 248^3 interior")
    r3 = L.reflection_3d(nets, interior=248, steps=3000, kinds=("learned", "cpml8", "cpml32"))
    r3["linear_3d"] = L.reflection_3d(lin3, interior=248, steps=3000,
                                      kinds=("learned", "cpml64"))["learned"]
    print(f"3D reflected / incident energy over 3000 steps: {r3}")
    assert np.isfinite(r3["learned"])

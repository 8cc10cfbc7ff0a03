"""Pins the parts of the oracle the golden CSVs leave open.

The CSV studies exercise only the X-axis CPML path with zero transverse
derivatives.  Here the blocked oracle is checked against an independent,
unblocked numpy restatement of the same update (global x-fastest arrays,
zero padding for the boundary) on small grids with random fields, random
lossy materials and CPML on all three axes -- so a wrong block index, a
swapped neighbour, a dropped psi term or a wrong axis window fails.  Plus
the reference's own unit tests (src/fdtd.rs:539-628, src/layout.rs:793-897,
src/source.rs:103-156, src/probe.rs:188-260) and decomposition invariance
(src/engine.rs:1008-1099)."""
import math

import numpy as np
import pytest

from oracle import wfo as W

f32 = np.float32


# ---------------------------------------------------------------------------
# independent scalar restatement (numpy, whole-grid arrays indexed [z, y, x])
# ---------------------------------------------------------------------------
def _shift(a, axis, d, periodic):
    """Value of `a` at index+d along axis (d = +1 / -1); zero outside."""
    if periodic:
        return np.roll(a, -d, axis=axis)
    out = np.zeros_like(a)
    sl_dst = [slice(None)] * 3
    sl_src = [slice(None)] * 3
    if d == 1:
        sl_dst[axis] = slice(0, -1)
        sl_src[axis] = slice(1, None)
    else:
        sl_dst[axis] = slice(1, None)
        sl_src[axis] = slice(0, -1)
    out[tuple(sl_dst)] = a[tuple(sl_src)]
    return out


class Brute:
    """f[c][z, y, x]; psi kept for every voxel and term (12 x grid)."""

    AX = {"x": 2, "y": 1, "z": 0}

    def __init__(self, dims, ids, table, shell, prof, periodic):
        self.nx, self.ny, self.nz = dims
        self.per = periodic
        q = table.astype(np.float32) * f32(1.0 / 65536.0)
        idg = ids.reshape(self.nz, self.ny, self.nx)
        self.ca, self.cb, self.da, self.db = (q[idg, k].astype(np.float32) for k in range(4))
        self.pml = prof is not None
        n = (self.nx, self.ny, self.nz)
        if self.pml:
            p = np.asarray(prof, np.float32)
            off = 0
            self.prof = []
            for a in range(3):
                self.prof.append(p[off:off + 3 * n[a]].reshape(3, n[a]))
                off += 3 * n[a]
            z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx),
                                  indexing="ij")
            sh = np.zeros(idg.shape, bool)
            for a, c in enumerate((x, y, z)):
                t = shell[a]
                nb = n[a] // 8
                if t > 0:
                    sh |= (c // 8 < t) | (c // 8 >= nb - t)
            self.shell = sh
            self.psi = {k: np.zeros(idg.shape, np.float32) for k in range(12)}

    def coef(self, axis_name):
        """(b, a, ik) broadcast over the grid for one axis."""
        a = "xyz".index(axis_name)
        shape = [1, 1, 1]
        shape[2 - a] = -1
        return tuple(self.prof[a][j].reshape(shape) for j in range(3))

    def corr(self, raw, axis_name, k):
        if not self.pml:
            return raw
        b, a, ik = self.coef(axis_name)
        pn = b * self.psi[k] + a * raw
        d = raw * ik + pn
        self.psi[k] = np.where(self.shell, pn, self.psi[k]).astype(np.float32)
        return np.where(self.shell, d, raw).astype(np.float32)

    def step(self, f):
        ex, ey, ez, hx, hy, hz = f
        p = lambda a, ax: _shift(a, self.AX[ax], +1, bool(self.per & (1 << "xyz".index(ax))))
        m = lambda a, ax: _shift(a, self.AX[ax], -1, bool(self.per & (1 << "xyz".index(ax))))
        d_ez_dy = self.corr(p(ez, "y") - ez, "y", 2)
        d_ex_dy = self.corr(p(ex, "y") - ex, "y", 3)
        d_ey_dz = self.corr(p(ey, "z") - ey, "z", 4)
        d_ex_dz = self.corr(p(ex, "z") - ex, "z", 5)
        d_ez_dx = self.corr(p(ez, "x") - ez, "x", 0)
        d_ey_dx = self.corr(p(ey, "x") - ey, "x", 1)
        hx = self.da * hx - self.db * (d_ez_dy - d_ey_dz)
        hy = self.da * hy - self.db * (d_ex_dz - d_ez_dx)
        hz = self.da * hz - self.db * (d_ey_dx - d_ex_dy)
        d_hz_dy = self.corr(hz - m(hz, "y"), "y", 9)
        d_hx_dy = self.corr(hx - m(hx, "y"), "y", 8)
        d_hy_dz = self.corr(hy - m(hy, "z"), "z", 10)
        d_hx_dz = self.corr(hx - m(hx, "z"), "z", 11)
        d_hz_dx = self.corr(hz - m(hz, "x"), "x", 6)
        d_hy_dx = self.corr(hy - m(hy, "x"), "x", 7)
        ex = self.ca * ex + self.cb * (d_hz_dy - d_hy_dz)
        ey = self.ca * ey + self.cb * (d_hx_dz - d_hz_dx)
        ez = self.ca * ez + self.cb * (d_hy_dx - d_hx_dy)
        return [ex, ey, ez, hx, hy, hz]


def _to_xyz(aosoa, dims):
    nx, ny, nz = dims
    f = aosoa.reshape(nz // 8, ny // 8, nx // 8, 6, 8, 8, 8)
    return [f[:, :, :, c].transpose(0, 3, 1, 4, 2, 5).reshape(nz, ny, nx) for c in range(6)]


def _to_aosoa(xyz, dims):
    nx, ny, nz = dims
    a = np.stack(xyz).reshape(6, nz // 8, 8, ny // 8, 8, nx // 8, 8)
    return a.transpose(1, 3, 5, 0, 2, 4, 6).reshape(-1, 6, 512).astype(np.float32)


@pytest.mark.parametrize("dims,shell,periodic", [
    ((16, 16, 16), (1, 1, 1), 0),
    ((24, 16, 32), (1, 1, 2), 0),
    ((32, 24, 16), (2, 0, 1), 0),
    ((16, 24, 16), (0, 1, 0), 5),
    ((24, 16, 24), None, 7),
    ((40, 16, 8), None, 0),
])
def test_blocked_oracle_matches_brute_force(dims, shell, periodic):
    rng = np.random.default_rng(sum(dims) + periodic)
    dt, dx = 1.9065752e-12, 1e-3
    table = W.vacuum_table(dt, dx)
    for m in range(1, 4):
        table[m] = W.material_coeffs(float(rng.uniform(1, 8)), 1.0, float(rng.uniform(0, 0.04)),
                                     dt, dx)
    nx, ny, nz = dims
    ids = rng.integers(0, 4, nx * ny * nz, dtype=np.uint8)
    prof = None
    if shell is not None:
        prof = np.concatenate([W.pml_profile(n, 8 * s, dt, dx).reshape(-1)
                               for n, s in zip(dims, shell)])
    f0 = rng.standard_normal((nx * ny * nz // 512, 6, 512)).astype(np.float32)
    f0[rng.random(f0.shape) < 0.05] = -0.0
    sim = W.OracleSim(nx, ny, nz, ids=ids, table=table, shell_blocks=shell, profiles=prof,
                      periodic=periodic)
    sim.write_fields(f0)
    br = Brute(dims, ids, table, shell or (0, 0, 0), prof, periodic)
    f = _to_xyz(f0, dims)
    for _ in range(4):
        sim.step_fields(threads=1)
        f = br.step(f)
    got = sim.read_fields()
    exp = _to_aosoa(f, dims)
    assert got.view(np.uint32).tolist() == exp.view(np.uint32).tolist()


def test_decomposition_invariance(wfo):
    """src/engine.rs:1008-1099: 1 slab vs many slabs, bit for bit."""
    dims = (64, 64, 64)
    dt, dx = 1.5e-12, 1e-3
    outs = []
    for threads in (1, 4, 8):
        sim = W.OracleSim(*dims, table=W.vacuum_table(dt, dx), shell_blocks=(2, 2, 2),
                          profiles=np.concatenate([W.pml_profile(64, 16, dt, dx).reshape(-1)] * 3))
        sim.add_source(32, 32, 32, 2, 1.0, 2, 3e10, 1.0 / 3e10)
        sim.run(30, dt, threads=threads)
        outs.append(sim.read_fields())
    assert outs[0].tobytes() == outs[1].tobytes() == outs[2].tobytes()


def test_uniform_field_has_zero_curl():
    """src/fdtd.rs:539-584 analogue: a uniform field in a periodic grid stays
    unchanged (every raw derivative is exactly zero)."""
    dims = (16, 16, 16)
    sim = W.OracleSim(*dims, table=W.vacuum_table(1e-12, 1e-3), periodic=7)
    f = np.empty((8, 6, 512), np.float32)
    for c, v in enumerate([1.5, -2.0, 3.25, 0.5, -0.75, 4.0]):
        f[:, c, :] = v
    sim.write_fields(f)
    sim.step_fields()
    assert sim.read_fields().tobytes() == f.tobytes()


def test_pml_identity_equals_plain():
    """src/fdtd.rs:590-628: PML kernels with identity profiles == plain
    kernels, except -0 derivatives become +0 (raw + 0.0f)."""
    dims = (16, 16, 16)
    rng = np.random.default_rng(1)
    f0 = rng.standard_normal((8, 6, 512)).astype(np.float32)
    ident = np.concatenate([np.stack([np.ones(16), np.zeros(16), np.ones(16)]).astype(
        np.float32).reshape(-1)] * 3)
    t = W.vacuum_table(1e-12, 1e-3)
    a = W.OracleSim(*dims, table=t, shell_blocks=(1, 1, 1), profiles=ident)
    b = W.OracleSim(*dims, table=t)
    a.write_fields(f0)
    b.write_fields(f0)
    for _ in range(3):
        a.step_fields()
        b.step_fields()
    fa, fb = a.read_fields(), b.read_fields()
    assert np.array_equal(fa, fb)  # equal as numbers (+0 == -0)
    assert np.all(a.read_psi() == 0)


def test_q16_and_vacuum_coefficients():
    """src/layout.rs:793-830."""
    lib = W.lib()
    for v in [0.0, 1.0, -1.0, 0.5, -0.5, 123.456, -999.999, 3.0e-4]:
        q = lib.wfo_q16_from_f32(v)
        assert abs(lib.wfo_q16_to_f32(q) - v) < 2e-4 * max(abs(v), 1.0)
    assert lib.wfo_q16_from_f32(0.5 / 65536) == 1  # round half away from zero
    assert lib.wfo_q16_from_f32(-0.5 / 65536) == -1
    assert lib.wfo_q16_from_f32(1e30) == 2**31 - 1  # saturating cast
    assert lib.wfo_q16_from_f32(float("nan")) == 0
    dt, dx = 1.5e-12, 1e-3
    ca, cb, da, db = (lib.wfo_q16_to_f32(int(q)) for q in W.material_coeffs(1, 1, 0, dt, dx))
    assert abs(ca - 1) < 1e-4 and abs(da - 1) < 1e-4
    eps0, mu0 = 8.854188e-12, 1.2566371e-6
    assert abs(cb - dt / (eps0 * dx)) < 0.05 * dt / (eps0 * dx)
    assert abs(db - dt / (mu0 * dx)) < 0.05 * dt / (mu0 * dx)
    lossy = W.material_coeffs(2.0, 1.0, 0.02, dt, dx)
    assert 0 < lib.wfo_q16_to_f32(int(lossy[0])) < 1


def test_pml_profile_properties():
    """src/layout.rs:843-889."""
    lib = W.lib()
    p0 = W.pml_profile(64, 0, 1.5e-12, 1e-3)
    assert np.all(p0[0] == 1) and np.all(p0[1] == 0) and np.all(p0[2] == 1)
    p = W.pml_profile(64, 8, 1.5e-12, 1e-3)
    assert p[0, 32] == 1 and p[1, 32] == 0 and p[2, 32] == 1
    assert p[2, 0] < 1 and p[1, 0] != 0 and 0 < p[0, 0] < 1
    assert np.array_equal(p[:, :8], p[:, ::-1][:, :8])
    assert lib.wfo_effective_pml_thickness_blocks(16, 16, 16, 1000) == 1
    assert lib.wfo_effective_pml_thickness_blocks(16, 16, 16, 0) == 0
    assert lib.wfo_effective_pml_thickness_blocks(64, 64, 64, 10) == 2
    assert lib.wfo_effective_pml_thickness_blocks(1024, 1024, 1024, 20) == 3
    # graded: sigma (and so 1-b) decreases inward
    assert np.all(np.diff(p[0, :8]) > 0)


def test_pml_profile_against_f64_formula():
    """The f32 profile against an f64 evaluation of the same CPML formulas
    (Roden & Gedney), to f32 accuracy: catches a wrong sign or term."""
    T, dt, dx = 16, 1.9e-12, 1e-3
    p = W.pml_profile(64, T, dt, dx)
    eps0, mu0 = 8.854188e-12, 1.2566371e-6
    eta0 = math.sqrt(mu0 / eps0)
    smax = -(3 + 1) * math.log(1e-6) / (2 * eta0 * T * dx)
    for g in range(T):
        rho = (T - g) / T
        s = smax * rho ** 3
        k = 1 + 6 * rho ** 3
        al = 0.05 * (1 - rho)
        b = math.exp(-(s / k + al) * dt / eps0)
        a = s * (b - 1) / (k * (s + k * al))
        assert abs(p[0, g] - b) < 1e-5 * max(b, 1e-3) + 1e-7
        assert abs(p[1, g] - a) < 1e-4 * abs(a) + 1e-9
        assert abs(p[2, g] - 1 / k) < 1e-6


def test_waveforms():
    """src/source.rs:103-129."""
    lib = W.lib()
    dt = 4.0 / 4000
    s = sum(lib.wfo_waveform(2, 1.0, 2.0, float(np.float32(i) * np.float32(dt)))
            for i in range(4000))
    assert abs(s * dt) < 0.05
    assert abs(lib.wfo_waveform(1, 1.0, 0.0, 0.0)) < 1e-6
    assert abs(lib.wfo_waveform(1, 1.0, 0.0, 0.25) - 1.0) < 1e-5
    assert lib.wfo_waveform(0, 2.0, 0.5, 2.0) == 1.0
    assert abs(lib.wfo_waveform(0, 2.0, 0.5, 2.5) - math.exp(-0.5)) < 1e-6
    assert abs(lib.wfo_waveform(2, 1.0, 2.0, 2.0) - 1.0) < 1e-6


def _ulps(a, b):
    """Distance between two finite f32 values in units in the last place."""
    ia, ib = (np.array([v], f32).view(np.int32)[0].astype(np.int64) for v in (a, b))
    ia = ia if ia >= 0 else -(ia & 0x7FFFFFFF)
    ib = ib if ib >= 0 else -(ib & 0x7FFFFFFF)
    return abs(int(ia) - int(ib))


def test_waveforms_against_f64_closed_forms():
    """src/source.rs:45-60 evaluated in f32 in the reference's operation order
    (Rust evaluates left to right: (-0.5 * arg) * arg, ((2 PI) f) t,
    (PI f) (t - t0), (1 - 2 sq) * exp(-sq)), with exp and sin taken in f64
    from the rounded f32 argument and rounded once.  The oracle's value is
    within one ulp of that everywhere and equal to it for at least 98% of
    5000 random (parameters, t) per waveform: a libm exp/sin more than an ulp
    off, a changed operation order (the sinusoid's argument reaches 157 rad,
    where one ulp of it moves sin by hundreds of ulps), swapped parameters or
    a dropped factor fail."""
    lib = W.lib()
    rng = np.random.default_rng(3)
    pi = f32(math.pi)
    for kind in range(3):
        exact = 0
        n = 5000
        for _ in range(n):
            if kind == 0:
                p0, p1, t = f32(rng.uniform(1e-11, 1e-10)), f32(rng.uniform(5e-12, 3e-11)), \
                    f32(rng.uniform(0, 3e-10))
                arg = f32(f32(t - p0) / p1)
                exp = f32(math.exp(float(f32(f32(f32(-0.5) * arg) * arg))))
            elif kind == 1:
                p0, p1, t = f32(rng.uniform(1e9, 5e10)), f32(0), f32(rng.uniform(0, 5e-10))
                exp = f32(math.sin(float(f32(f32(f32(f32(2.0) * pi) * p0) * t))))
            else:
                p0, p1, t = f32(rng.uniform(5e9, 5e10)), f32(rng.uniform(1e-11, 1e-10)), \
                    f32(rng.uniform(0, 3e-10))
                arg = f32(f32(pi * p0) * f32(t - p1))
                sq = f32(arg * arg)
                exp = f32(f32(f32(1.0) - f32(f32(2.0) * sq)) * f32(math.exp(-float(sq))))
            got = f32(lib.wfo_waveform(kind, float(p0), float(p1), float(t)))
            assert _ulps(got, exp) <= 1, (kind, p0, p1, t, got, exp)
            exact += got == exp
        assert exact >= 0.98 * n, (kind, exact)


def test_soft_source_adds():
    """src/source.rs:131-156."""
    sim = W.OracleSim(16, 16, 16, table=W.vacuum_table(1e-12, 1e-3))
    sim.add_source(8, 8, 8, 2, 2.0, 1, 1.0)
    f = sim.read_fields()
    # inject via a zero-length field step: run with dt so that t = 0.25
    W.lib().wfo_sim_run  # noqa: B018
    sim2 = W.OracleSim(16, 16, 16, table=W.vacuum_table(1e-12, 1e-3))
    sim2.add_source(8, 8, 8, 2, 2.0, 1, 1.0)
    sim2.add_source(8, 8, 8, 2, 2.0, 1, 1.0)
    sim2.run(1, 0.25)
    v = sim2.read_fields()[(1 * 2 + 1) * 2 + 1, 2, 0]
    assert abs(v - 4.0) < 1e-4
    assert f.sum() == 0


def test_probe_recovers_sinusoid_and_start_gate():
    """src/probe.rs:188-260 through the oracle's own Probe::accumulate: the
    probe voxel (8,8,8) is driven with A sin(omega t + phi) in f32 (A = 3.5,
    phi = 0.7, 1 GHz, dt = 1 ps, 4000 steps) by writing the fields, the
    oracle accumulates once per step, and its spectrum recovers the amplitude
    and phase within 5% (0.05 rad) and a near-zero amplitude at the undriven
    3 GHz.  Then the start gate: a huge value before start is never recorded.
    Plants: a dropped sin term (phase error pi/2), the wrong sign of the
    phase (error 1.4 rad) or accumulating the f32 frequency twice would each
    fail one of these bounds."""
    import ctypes
    f32 = np.float32
    freq, A, ph, dt = f32(1e9), f32(3.5), 0.7, f32(1e-12)
    sim = W.OracleSim(16, 16, 16, table=W.vacuum_table(1e-12, 1e-3))
    sim.add_probe(8, 8, 8, 2, [freq, f32(3) * freq], 0.0)
    fields = np.zeros((8, 6, 512), np.float32)
    blk, loc = (1 * 2 + 1) * 2 + 1, 0  # voxel (8,8,8): block (1,1,1), local (0,0,0)
    omega = f32(2.0) * f32(math.pi) * freq
    n = 4000
    for step in range(n):
        t = (f32(step) + f32(1.0)) * dt
        fields[blk, 2, loc] = A * f32(math.sin(float(omega * t + f32(ph))))
        sim.write_fields(fields)
        sim.accumulate_probes(t)
    sc, ss, cnt = sim.probe_sums()
    assert cnt.tolist() == [n]
    amp = np.zeros(2)
    pha = np.zeros(2)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    W.lib().wfo_probe_spectrum(2, P(sc), P(ss), n, P(amp), P(pha))
    assert abs(amp[0] - A) < 0.05 * A, amp
    err = (pha[0] - ph + math.pi) % (2 * math.pi) - math.pi
    assert abs(err) < 0.05, pha
    assert amp[1] < 0.05 * A, amp
    sim.close()
    # start gate (src/probe.rs:244-259), again through the oracle's accumulate
    sim = W.OracleSim(16, 16, 16, table=W.vacuum_table(1e-12, 1e-3))
    sim.add_probe(0, 0, 0, 0, [1e9], 5e-9)
    fields = np.zeros((8, 6, 512), np.float32)
    fields[0, 0, 0] = 1.0e6
    sim.write_fields(fields)
    sim.accumulate_probes(f32(1.0e-9))
    sim.accumulate_probes(f32(4.9e-9))
    c, s, cnt = sim.probe_sums()
    assert cnt.tolist() == [0] and c[0] == 0 and s[0] == 0
    sim.accumulate_probes(f32(5.0e-9))  # t == start records
    c, s, cnt = sim.probe_sums()
    assert cnt.tolist() == [1] and c[0] != 0


def test_scene_voxelize_reference_example(wfo):
    """scenes/two_spheres.scene-style text: dedup, ids in first-seen order,
    box clamping, later primitives overwrite earlier ones."""
    text = ("# comment\n"
            "sphere 4.0 1.0 0.0   16 16 10   6\n"
            "sphere 2.1 1.0 0.0   16 16 22   4\n"
            "box    1.0 1.0 0.05   4 4 26   28 28 30\n"
            "sphere 4.0 1.0 0.0   2 2 2   3\n"
            "box 2.1 1.0 0.0 -5 -5 -5 1.5 1.2 0.3\n")
    ids, table, n = W.scene_voxelize(text, 32, 32, 32, 1e-12, 1e-3)
    g = ids.reshape(32, 32, 32)
    assert n == 3
    assert g[10, 16, 16] == 1 and g[22, 16, 16] == 2 and g[27, 10, 10] == 3
    assert g[0, 0, 0] == 2  # box [0,2)x[0,2)x[0,1) overwrote the sphere at (2,2,2)
    assert g[1, 1, 1] == 1
    assert np.array_equal(table[1], W.material_coeffs(4.0, 1.0, 0.0, 1e-12, 1e-3))
    with pytest.raises(ValueError):
        W.scene_voxelize("cone 1 2 3\n", 8, 8, 8, 1e-12, 1e-3)
    many = "".join(f"box {1 + i * 0.01} 1 0 0 0 0 1 1 1\n" for i in range(256))
    with pytest.raises(ValueError, match="255"):
        W.scene_voxelize(many, 8, 8, 8, 1e-12, 1e-3)

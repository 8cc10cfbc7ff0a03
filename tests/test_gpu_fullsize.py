"""Workloads at their full size against the CPU oracle, bit for bit.

* The 1024^3 random fill (1 073 741 824 cells, CPML 24, 200 materials,
  sinusoid source, probe): two steps from random fields on the GPU, in the
  default fused variant and in the in-place split variant -- fields, psi and
  probe sums.  This covers, at the real shape, what the small parity cases
  cannot: the interior box with its z slices, the frame boxes, the chosen
  field placement and 64-bit indexing past 2^31 elements.
* The 512^3 two-sphere run through the snapshot stream (wf_run): every
  snapshot written, the final state and the probe sums.

Host memory: about 90 GB at 1024^3 (initial fields, the oracle's state and
result, one read-back at a time); the oracle runs on up to 16 threads."""
import os

import numpy as np
import pytest

from tests.helpers import oracle_sim_for
from wavefront_amd import workloads
from wavefront_amd.engine import DeviceEngine

pytestmark = pytest.mark.gpu
STEPS = 2


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _random_fields(dims):
    """A 64 MiB block of random values (5% exact zeros) repeated over the grid."""
    rng = np.random.default_rng(7)
    pat = rng.standard_normal(1 << 24).astype(np.float32)
    pat[rng.random(pat.size) < 0.05] = 0.0
    out = np.empty((dims.block_count(), 6, 512), np.float32)
    flat = out.reshape(-1)
    for i in range(0, flat.size, pat.size):
        n = min(pat.size, flat.size - i)
        flat[i:i + n] = pat[:n]
    return out


def _equal(a, b, what, chunk=1 << 28):
    """Bitwise equality, compared in chunks to bound the temporary memory."""
    assert a.shape == b.shape and a.dtype == b.dtype, f"{what}: {a.shape} vs {b.shape}"
    it = np.uint32 if a.dtype == np.float32 else np.uint64
    va, vb = a.reshape(-1).view(it), b.reshape(-1).view(it)
    for i in range(0, va.size, chunk):
        x, y = va[i:i + chunk], vb[i:i + chunk]
        if not np.array_equal(x, y):
            j = i + int(np.flatnonzero(x != y)[0])
            raise AssertionError(f"{what}: first difference at flat index {j} of {va.size}: "
                                 f"{a.reshape(-1)[j]!r} vs {b.reshape(-1)[j]!r}")


@pytest.fixture(scope="module")
def oracle_run():
    w = workloads.random_1024(steps=STEPS)
    _say(f"random_1024 full size: oracle, {STEPS} steps from random fields")
    f0 = _random_fields(w.dims)
    sim = oracle_sim_for(w)
    sim.write_fields(f0)
    sim.run(STEPS, w.dt, threads=min(os.cpu_count() or 1, 16))
    out = (sim.read_fields(), sim.read_psi(), sim.probe_sums())
    sim.close()
    return w, f0, out


@pytest.mark.parametrize("variant", ["fused", "onestep", "split"])
def test_random_1024_full_size_bit_exact(oracle_run, variant, monkeypatch):
    """fused: two steps per pass (the two-step kernel and the frame through a
    third state); onestep: the fused interior kernel and frame kernels once
    per step; split: the in-place H and E kernels."""
    monkeypatch.setenv("WF_KERNEL", "split" if variant == "split" else "fused")
    monkeypatch.setenv("WF_TWOSTEP", "0" if variant == "onestep" else "1")
    w, f0, (of, op, (oc, os_, on)) = oracle_run
    _say(f"random_1024 full size: GPU {variant}, {STEPS} steps")
    with DeviceEngine(w.dims, w.ids, w.table, w.pml()) as eng:
        eng.write_fields(f0)
        eng.set_sources(w.sources, w.dt, STEPS)
        eng.set_probes(w.probes, w.dt, STEPS)
        eng.step(STEPS)
        eng.sync()
        _equal(eng.read_fields(), of, "fields")
        _equal(eng.read_psi(), op, "psi")
        sc, ss, n = eng.probe_sums()
    _equal(sc, oc, "sum_cos")
    _equal(ss, os_, "sum_sin")
    assert list(n) == list(on) == [STEPS] * len(w.probes)
    # the steps changed the fields everywhere, not just near the source
    assert np.count_nonzero(of.reshape(-1)[::4099] != f0.reshape(-1)[::4099]) > 0.9 * (of.size // 4099)


def _rel_l2(a, b, chunk=1 << 26):
    """||a - b|| / ||b|| in float64, chunked to bound the temporary memory."""
    va, vb = a.reshape(-1), b.reshape(-1)
    num = den = 0.0
    for i in range(0, va.size, chunk):
        x = va[i:i + chunk].astype(np.float64)
        y = vb[i:i + chunk].astype(np.float64)
        num += float(((x - y) ** 2).sum())
        den += float((y ** 2).sum())
    return (num / den) ** 0.5


def test_random_1024_full_size_contracted(oracle_run):
    """The contracted kernel set (a*b+c as fused multiply-adds) at the full
    1024^3 size, two steps from random fields: relative L2 field error
    against the CPU engine <= 1e-6, and not bit-identical (the contracted
    kernels ran)."""
    w, f0, (of, op, _) = oracle_run
    _say(f"random_1024 full size: GPU contracted, {STEPS} steps")
    with DeviceEngine(w.dims, w.ids, w.table, w.pml()) as eng:
        eng.set_fp_contract(True)
        eng.write_fields(f0)
        eng.set_sources(w.sources, w.dt, STEPS)
        eng.set_probes(w.probes, w.dt, STEPS)
        eng.step(STEPS)
        eng.sync()
        got = eng.read_fields()
    rel = _rel_l2(got, of)
    print(f"contracted 1024^3: relative L2 error {rel:.3e}")
    assert rel <= 1e-6
    assert not np.array_equal(got.reshape(-1)[::4097], of.reshape(-1)[::4097])


def test_spheres_512_full_size_contracted():
    """The contracted kernel set on the 512^3 two-sphere workload, 200 steps of
    the Ricker source from zero fields (the pulse well past the spheres):
    relative L2 field error against the CPU engine <= 1e-6."""
    steps = 200
    w = workloads.spheres_512(steps=steps)
    _say(f"spheres_512 full size: GPU contracted, {steps} steps")
    with DeviceEngine(w.dims, w.ids, w.table, w.pml()) as eng:
        eng.set_fp_contract(True)
        eng.set_sources(w.sources, w.dt, steps)
        eng.set_probes(w.probes, w.dt, steps)
        eng.step(steps)
        eng.sync()
        got = eng.read_fields()
    sim = oracle_sim_for(w)
    sim.run(steps, w.dt, threads=min(os.cpu_count() or 1, 16))
    ref = sim.read_fields()
    sim.close()
    rel = _rel_l2(got, ref)
    print(f"contracted 512^3 spheres, {steps} steps: relative L2 error {rel:.3e}")
    assert np.abs(ref).max() > 0
    assert rel <= 1e-6
    assert not np.array_equal(got, ref)


def test_spheres_512_full_size_snapshots(tmp_path):
    """The 512^3 two-sphere workload at its full size (134 M cells, CPML 16,
    Ricker source, 32 probes) from random fields through the snapshot stream:
    30 steps with a snapshot every 10 (three 3.22 GB states written by
    wf_run), each snapshot and the final state bit for bit against the
    oracle's state after the same step, and the probe sums."""
    steps, every = 30, 10
    w = workloads.spheres_512(steps=steps)
    _say(f"spheres_512 full size: {steps} steps, a snapshot every {every}")
    f0 = _random_fields(w.dims)
    per = w.dims.voxel_count() * 24
    path = tmp_path / "spheres_512.bin"
    with DeviceEngine(w.dims, w.ids, w.table, w.pml()) as eng:
        eng.write_fields(f0)
        eng.set_sources(w.sources, w.dt, steps)
        eng.set_probes(w.probes, w.dt, steps)
        eng.run(steps, every, str(path))
        final = eng.read_fields()
        sc, ss, n = eng.probe_sums()
    assert os.path.getsize(path) == 3 * per
    sim = oracle_sim_for(w)
    sim.write_fields(f0)
    threads = min(os.cpu_count() or 1, 16)
    done = 0
    with open(path, "rb") as fh:
        for k in range(0, steps, every):  # snapshot after loop index k = after k + 1 steps
            sim.run(k + 1 - done, w.dt, first_step=done, threads=threads)
            done = k + 1
            snap = np.frombuffer(fh.read(per), dtype=np.float32)
            _equal(snap, sim.read_fields().reshape(-1), f"snapshot after step {k + 1}")
    sim.run(steps - done, w.dt, first_step=done, threads=threads)
    _equal(final, sim.read_fields(), "final state")
    oc, os_, on = sim.probe_sums()
    sim.close()
    _equal(sc, oc, "sum_cos")
    _equal(ss, os_, "sum_sin")
    assert list(n) == list(on) == [steps] * len(w.probes)
    os.remove(path)

"""The 1024^3 random-fill workload at its full size (1 073 741 824 cells, CPML 24,
200 materials, sinusoid source, probe): two steps from random fields on the
GPU, in the default fused variant and in the in-place split variant, against
the CPU oracle -- fields, psi and probe sums bit for bit.  This covers, at the
real shape, what the small parity cases cannot: the interior box with its
z slices, the frame boxes, the chosen field placement and 64-bit indexing past
2^31 elements.

Host memory: about 90 GB (initial fields, the oracle's state and result, one
read-back at a time); the oracle runs on up to 16 threads."""
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


@pytest.mark.parametrize("variant", ["fused", "split"])
def test_random_1024_full_size_bit_exact(oracle_run, variant, monkeypatch):
    monkeypatch.setenv("WF_KERNEL", variant)
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

"""The allocation-free network evaluation (wf_mlp_predict_noalloc) against
the allocating one (wf_mlp_predict, through a trainer): the same kernels on
the same packed network, so the outputs must agree bit for bit -- on two
calls in a row, the second smaller and with another network and head, so a
second call that read anything the first left in the shared static array
would show."""
from ctypes import c_float

import numpy as np
import pytest

from tests.test_learned_cpu import _params
from wavefront_amd import learned as L
from wavefront_amd._lib import WF_ERR_INVALID, WavefrontError, check, lib, ptr


def _noalloc(params, head, u, device=0):
    p = np.ascontiguousarray(params, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    out = np.full(len(u), np.nan, np.float32)
    check(lib().wf_mlp_predict_noalloc(device, ptr(p, c_float), head, len(u), ptr(u, c_float),
                                       ptr(out, c_float)))
    return out


def _with_trainer(params, head, u):
    tr = L.Trainer(u, np.zeros(len(u), np.float32), head=head)
    tr.set_params(np.asarray(params, np.float32))
    y = tr.predict(u)
    tr.close()
    return y


def test_noalloc_rejects_more_samples_than_the_scratch_holds():
    """Needs no device: the size check comes before any device work."""
    cap = int(lib().wf_mlp_predict_noalloc_max())
    assert cap == 65536
    p = np.zeros(L.param_count(), np.float32)
    u = np.zeros((1, 16), np.float32)
    y = np.zeros(1, np.float32)
    rc = lib().wf_mlp_predict_noalloc(0, ptr(p, c_float), L.HEAD_DIRECT, cap + 1, ptr(u, c_float),
                                      ptr(y, c_float))
    assert rc == WF_ERR_INVALID
    assert "maximum" in lib().wf_last_error().decode()
    with pytest.raises(WavefrontError):
        check(lib().wf_mlp_predict_noalloc(0, ptr(p, c_float), 7, 1, ptr(u, c_float),
                                           ptr(y, c_float)))


@pytest.mark.gpu
def test_noalloc_predict_equals_trainer_predict_twice():
    print("[gpu-case] wf_mlp_predict_noalloc vs wf_mlp_predict: n=65536 direct, n=333 delta",
          flush=True)
    rng = np.random.default_rng(21)
    cases = ((65536, L.HEAD_DIRECT, 1.0), (333, L.HEAD_DELTA, 3.0))
    for n, head, amp in cases:
        p = _params(rng).astype(np.float32)
        u = (amp * rng.normal(0, 1, (n, 16))).astype(np.float32)
        ref = _with_trainer(p, head, u)
        got = _noalloc(p, head, u)
        assert np.isfinite(ref).all() and np.abs(ref).max() > 0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (n, head)
    # an empty call stages the network only and succeeds
    assert _noalloc(p, head, np.zeros((0, 16), np.float32)).size == 0

"""The allocation-free forward pass (tc_ws_net_forward_noalloc) against tc_ws_net_forward_backward:
two calls in a row, the second on a smaller input, bit for bit; inputs beyond its limits rejected."""
import ctypes as C

import numpy as np
import pytest

from oracle import warmstart_oracle as W
from taremincloth_amd import _lib

pytestmark = pytest.mark.gpu

MAX_ROWS = 16384


def _case(rows, seed):
    rng = np.random.default_rng(seed)
    feat = rng.normal(0, 1, (rows, W.F)).astype(np.float32)
    feat[:, 16:] = 0.0
    pairs = {(i, (i + 1) % rows) for i in range(rows)}
    for _ in range(rows // 2):
        a, b = rng.integers(0, rows, 2)
        if a != b:
            pairs.add((int(a), int(b)))
    a = np.array([p[0] for p in pairs])
    b = np.array([p[1] for p in pairs])
    off, nbr, _ = W.csr_pairs(rows, a, b)
    return feat, off, nbr


def _noalloc(net, feat, off, nbr):
    L = _lib.load()
    rows = len(feat)
    y = np.zeros((rows, 16), np.float32)
    vp = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
    st = L.tc_ws_net_forward_noalloc(net.handle, vp(feat), rows, vp(off), vp(nbr), vp(y))
    return st, y


def test_gpu_noalloc_forward_equals_the_allocating_pass(require_gpu):
    from taremincloth_amd.warmstart import WarmStartNet
    net = WarmStartNet(3)
    net.set_params(W.init_params(3, 21))
    for rows, seed in ((MAX_ROWS, 1), (3001, 2)):   # the largest input, then a smaller one
        feat, off, nbr = _case(rows, seed)
        want, _, _ = net.forward_backward(feat, off, nbr)
        st, got = _noalloc(net, feat, off, nbr)
        assert st == _lib.TC_OK, _lib.load().tc_last_error()
        assert np.array_equal(got, want), (rows, float(np.abs(got - want).max()))


def test_gpu_noalloc_forward_rejects_inputs_beyond_its_limits(require_gpu):
    from taremincloth_amd.warmstart import WarmStartNet
    net = WarmStartNet(1)
    net.set_params(W.init_params(1, 3))
    feat, off, nbr = _case(MAX_ROWS + 1, 4)
    st, _ = _noalloc(net, feat, off, nbr)
    assert st == _lib.TC_ERR_VALUE and b"limit" in _lib.load().tc_last_error()
    feat, off, _ = _case(64, 5)
    dense_off = (np.arange(65, dtype=np.uint32) * 2100).astype(np.uint32)   # 134,400 > 131,072 entries
    st, _ = _noalloc(net, feat, dense_off, np.zeros(int(dense_off[-1]), np.uint32))
    assert st == _lib.TC_ERR_VALUE and b"limit" in _lib.load().tc_last_error()

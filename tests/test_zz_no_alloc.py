"""mt_exclusive_scan_i32_noalloc (no workspace, no device allocation) against
mt_exclusive_scan_i32 with a caller workspace, through the C ABI: bit-identical on two
calls in a row, the second smaller than the first (the shared state array keeps the first
call's look-back words, which the second must not see), plus the size limit."""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import _native

pytestmark = pytest.mark.gpu

LIMIT = 1 << 28  # MT_SCAN_NOALLOC_MAX_ELEMENTS


def _scan_with_workspace(x):
    n = x.numel()
    ws = _native.workspace(_native.query("mt_scan_workspace_bytes", n), x.device)
    _native.call("mt_exclusive_scan_i32", x.data_ptr(), n, ws.data_ptr(), ws.numel(),
                 _native.stream(x))


def _expected(host):
    out = np.zeros_like(host)
    out[1:] = np.cumsum(host[:-1].astype(np.int64)).astype(np.int32)  # int32 wrap-around
    return out


def test_noalloc_scan_matches_workspace_scan_twice():
    _native.load_library()
    rng = np.random.default_rng(0)
    # The second call is smaller: its tiles reuse the first call's look-back words.
    for n in (5_000_003, 70_001, 1):
        host = rng.integers(-(1 << 20), 1 << 20, n).astype(np.int32)
        a = torch.from_numpy(host).cuda()
        b = a.clone()
        _native.call("mt_exclusive_scan_i32_noalloc", a.data_ptr(), n, _native.stream(a))
        _scan_with_workspace(b)
        torch.cuda.synchronize()
        assert torch.equal(a, b), n
        np.testing.assert_array_equal(a.cpu().numpy(), _expected(host))


def test_noalloc_scan_wraps_like_int32():
    n = 3 * 4096 + 5
    host = np.full(n, 1 << 20, np.int32)  # the running sum passes 2^31
    a = torch.from_numpy(host).cuda()
    _native.call("mt_exclusive_scan_i32_noalloc", a.data_ptr(), n, _native.stream(a))
    np.testing.assert_array_equal(a.cpu().numpy(), _expected(host))


def test_noalloc_scan_size_limit_and_empty():
    x = torch.zeros(4, dtype=torch.int32, device="cuda")
    # Rejected on the host before any launch.
    with pytest.raises(RuntimeError, match="no-allocation limit"):
        _native.call("mt_exclusive_scan_i32_noalloc", x.data_ptr(), LIMIT + 1, _native.stream(x))
    with pytest.raises(RuntimeError, match="negative"):
        _native.call("mt_exclusive_scan_i32_noalloc", x.data_ptr(), -1, _native.stream(x))
    _native.call("mt_exclusive_scan_i32_noalloc", x.data_ptr(), 0, _native.stream(x))
    torch.cuda.synchronize()
    assert torch.equal(x, torch.zeros_like(x))

"""Device primitives checked on their own: the wave reductions of the blend
backward (lane layout of v_permlane32_swap / v_permlane16_swap / DPP).

Bar: each total equals the float64 sum of its 64 inputs within 1e-5 relative
to the sum of magnitudes (float32 reassociation of 64 terms).
"""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

ORDER = [0, 4, 2, 6, 1, 5, 3, 7]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wave_sum8_lane_layout(seed):
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((9, 64)).astype(np.float32)
    # distinct magnitudes per value so a misplaced total cannot pass
    x *= (np.arange(9, dtype=np.float32) + 1.0)[:, None] * 10.0
    dev = torch.device("cuda:0")
    inp = torch.tensor(x.reshape(-1), device=dev)
    out = torch.zeros(128, device=dev)
    assert lib.dgmr_selftest_wave_sum(ctypes.c_void_p(inp.data_ptr()), ctypes.c_void_p(out.data_ptr()), None) == N.DGMR_OK
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    want = x.astype(np.float64).sum(axis=1)
    scale = np.abs(x.astype(np.float64)).sum(axis=1)
    for m in range(8):
        q = ORDER[m]
        assert abs(o[8 * m] - want[q]) <= 1e-5 * scale[q], (m, q, o[8 * m], want[q])
    assert abs(o[64 + 63] - want[8]) <= 1e-5 * scale[8]


@pytest.mark.parametrize("n,begin,end", [(0, 0, 13), (1, 0, 13), (4095, 0, 13), (70001, 0, 13), (70001, 0, 32), (12345, 3, 20),
                                         (50000, 0, 6), (100003, 0, 17), ((48 << 20) + 5, 0, 13)])
def test_sort_pairs_matches_stable_argsort(n, begin, end):
    """Bit-exact against numpy's stable sort on the key bits [begin, end); keys
    carry random bits outside that range, which must not take part.  The last
    case takes the large-key path (8 192 keys per block)."""
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    rng = np.random.default_rng(n + 7 * end)
    # few distinct digits so stability is exercised
    keys = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 4:
        keys[: n // 2] &= np.uint32(0xFFFFFFFF ^ (((1 << (end - begin)) - 1) << begin) | (7 << begin))
    vals = np.arange(n, dtype=np.uint32)
    field = (keys.astype(np.uint64) >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1)
    order = np.argsort(field, kind="stable")
    dev = torch.device("cuda:0")
    as_dev = lambda a: torch.tensor(a.view(np.int32), device=dev)
    k, v = as_dev(keys), as_dev(vals)
    kt, vt = torch.empty_like(k), torch.empty_like(v)
    nbytes = ctypes.c_size_t(0)
    assert lib.dgmr_sort_workspace_size(n, ctypes.byref(nbytes)) == N.DGMR_OK
    ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.dgmr_sort_pairs(p(k), p(v), p(kt), p(vt), n, begin, end, p(ws), nbytes.value, None) == N.DGMR_OK
    torch.cuda.synchronize()
    np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32), order.astype(np.uint32))
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint32), keys[order])


@pytest.mark.parametrize("n,inclusive", [(1, 0), (2048, 1), (2049, 0), (131072, 1), (1000003, 0), (1000003, 1), (262143, 1), (262144, 0), (12_000_001, 0)])
def test_scan_matches_cumsum(n, inclusive):
    """Bit-exact against numpy's cumulative sum modulo 2^32.  2 049 items take
    two tiles of 2 048; from 262 144 items the tiles hold 8 192; 12M take
    1 465 tiles, more than the chip runs at once, so later tiles look back
    across blocks that have already finished.  Values
    up to 2^31 make the running sum wrap.  The workspace is filled with 0xff
    first: the scan must not depend on its contents."""
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    rng = np.random.default_rng(n + inclusive)
    x = rng.integers(0, 2**31, size=n, dtype=np.uint64).astype(np.uint32)
    x[rng.random(n) < 0.3] = 0
    c = np.cumsum(x.astype(np.uint64)) & np.uint64(0xFFFFFFFF)
    want = c if inclusive else np.concatenate([[0], c[:-1]]).astype(np.uint64)
    dev = torch.device("cuda:0")
    inp = torch.tensor(x.view(np.int32), device=dev)
    out = torch.full_like(inp, -1)
    total = torch.full((1,), -1, dtype=torch.int32, device=dev)
    nbytes = ctypes.c_size_t(0)
    assert lib.dgmr_scan_workspace_size(n, ctypes.byref(nbytes)) == N.DGMR_OK
    ws = torch.full((max(nbytes.value, 1),), 255, dtype=torch.uint8, device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    for rep in range(2):  # a second call reuses the workspace as the first left it
        assert lib.dgmr_scan_u32(p(inp), p(out), p(total), n, inclusive, p(ws), nbytes.value, None) == N.DGMR_OK
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), want.astype(np.uint32), err_msg=f"call {rep}")
        assert int(total.cpu().numpy().view(np.uint32)[0]) == int(c[-1])
    # in place
    assert lib.dgmr_scan_u32(p(inp), p(inp), None, n, inclusive, p(ws), nbytes.value, None) == N.DGMR_OK
    torch.cuda.synchronize()
    np.testing.assert_array_equal(inp.cpu().numpy().view(np.uint32), want.astype(np.uint32))

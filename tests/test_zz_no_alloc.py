"""dgmr_sort_pairs_noalloc (the sort with its scratch in a library-owned
device array) against dgmr_sort_pairs with a caller workspace, through the C
ABI: two calls in a row, the second with fewer keys (the array keeps the
first call's contents), both bit-exact against the workspace entry point and
numpy's stable argsort; a size above the maximum is rejected."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


def _sort_both(lib, N, torch, keys, begin, end):
    n = keys.size
    dev = torch.device("cuda:0")
    as_dev = lambda a: torch.tensor(a.view(np.int32), device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    vals = np.arange(n, dtype=np.uint32)
    # the entry point with a caller workspace
    k1, v1 = as_dev(keys), as_dev(vals)
    kt1, vt1 = torch.empty_like(k1), torch.empty_like(v1)
    nbytes = ctypes.c_size_t(0)
    assert lib.dgmr_sort_workspace_size(n, ctypes.byref(nbytes)) == N.DGMR_OK
    ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=dev)
    assert lib.dgmr_sort_pairs(p(k1), p(v1), p(kt1), p(vt1), n, begin, end, p(ws), nbytes.value, None) == N.DGMR_OK
    # the one without
    k2, v2 = as_dev(keys), as_dev(vals)
    kt2, vt2 = torch.empty_like(k2), torch.empty_like(v2)
    rc = lib.dgmr_sort_pairs_noalloc(p(k2), p(v2), p(kt2), p(vt2), n, begin, end, None)
    assert rc == N.DGMR_OK, lib.dgmr_last_error()
    torch.cuda.synchronize()
    return (k1.cpu().numpy().view(np.uint32), v1.cpu().numpy().view(np.uint32), k2.cpu().numpy().view(np.uint32),
            v2.cpu().numpy().view(np.uint32))


def test_sort_without_allocation_matches_the_workspace_sort():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    for n, begin, end, seed in ((5_000_003, 0, 24, 1), (1_000_001, 3, 16, 2)):
        print(f"[gpu no-alloc sort] n={n} bits [{begin},{end})", flush=True)
        rng = np.random.default_rng(seed)
        keys = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] &= np.uint32(~(((1 << (end - begin)) - 1) << begin) & 0xFFFFFFFF) | np.uint32(5 << begin)  # repeated digits
        k_ws, v_ws, k_na, v_na = _sort_both(lib, N, torch, keys, begin, end)
        np.testing.assert_array_equal(v_na, v_ws)
        np.testing.assert_array_equal(k_na, k_ws)
        field = (keys.astype(np.uint64) >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1)
        order = np.argsort(field, kind="stable").astype(np.uint32)
        np.testing.assert_array_equal(v_na, order)
    # above the maximum: rejected before any device work
    dummy = ctypes.c_void_p(256)
    rc = lib.dgmr_sort_pairs_noalloc(dummy, dummy, dummy, dummy, (1 << 25) + 1, 0, 32, None)
    assert rc == N.DGMR_ERR_ARGUMENT
    assert b"DGMR_SORT_NOALLOC_MAX_ITEMS" in lib.dgmr_last_error()

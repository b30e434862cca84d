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

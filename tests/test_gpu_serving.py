"""Continuous-batching server: every request's result is the same alone and in
the mix.

Requests over three resident scenes arrive faster than they are served, so
the scheduler batches several cameras of a scene per launch while other
batches run on other worker streams.  Each request is then rendered alone
through the one-camera path.  Bars: image max |mix - alone| <= 1e-6; for
forward+backward requests, per parameter tensor
||g_mix - g_alone|| <= 1e-4 ||g_alone|| (float atomics sum in another order).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import serving
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter
    resident = serving.make_resident_scenes(3, seed=7, n_min=5_000, n_max=40_000)
    trace = serving.make_trace(n_requests=24, rate_hz=5_000.0, n_scenes=3, seed=3, backward_fraction=0.3, min_res=(64, 64),
                               max_res=(256, 192))
    print(f"[gpu serving] {len(trace)} requests over 3 scenes", flush=True)
    server = serving.ContinuousBatchingServer(resident, hbm_budget_bytes=2 << 30, workers=2, max_batch_views=8)
    stats = server.run(trace)
    assert len(stats.latencies) == len(trace)
    assert stats.batches < len(trace), "no request was batched with another"
    assert any(r.backward for r in trace)
    bg = torch.zeros(3)
    for r in trace:
        sc = resident[r.scene]
        cam = r.camera
        args = (*sc.params, torch.tensor(cam.world_view_transform), torch.tensor(cam.full_proj_transform),
                torch.tensor(cam.camera_center), bg, cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, sc.sh_degree)
        cache = splatter.cpp_binding.forward(*args)
        alone = cache[0]
        assert float((r.image - alone).abs().max()) <= 1e-6, r.rid
        if r.backward:
            g_alone = splatter.cpp_binding.backward(*args, serving.l1_grad(alone, r.target_seed), cache)
            for a, b in zip(r.grads, g_alone):
                nb = float(b.norm())
                assert float((a - b).norm()) <= 1e-4 * nb + 1e-12, (r.rid, float((a - b).norm()), nb)
        else:
            assert r.grads is None
    torch.cuda.synchronize()


def test_servers_share_worker_streams():
    """A second server reuses the first one's worker streams (the caching
    allocator reuses freed blocks only on their own stream), so repeated
    servers do not keep reserving new memory."""
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd import serving
    resident = serving.make_resident_scenes(2, seed=9, n_min=5_000, n_max=20_000)
    print("[gpu serving] two servers in a row", flush=True)
    reserved = []
    for _ in range(2):
        trace = serving.make_trace(n_requests=12, rate_hz=5_000.0, n_scenes=2, seed=4, backward_fraction=0.3, min_res=(64, 64),
                                   max_res=(128, 128))
        for r in trace:
            r.arrival = 0.0
        stats = serving.ContinuousBatchingServer(resident, hbm_budget_bytes=1 << 30, workers=2, max_batch_views=8).run(trace)
        torch.cuda.synchronize()
        assert len(stats.latencies) == len(trace)
        reserved.append(torch.cuda.memory_reserved())
    dev = torch.device("cuda:0")
    assert serving._worker_stream(dev, 0) is serving._worker_stream(dev, 0)
    assert sum(1 for key in serving._STREAMS if key[0] == str(dev)) >= 2
    assert reserved[1] <= reserved[0] * 1.5 + (64 << 20)

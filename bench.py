"""Benchmark: tile splat forward+backward at 1M Gaussians, 1920x1080.

A step is one training iteration of the hot path on one frame: parameter
activations (exp of log-scales, quaternion normalisation, sigmoid of opacity
logits; fused into the splatter's preprocess unless --torch-activations),
splat forward, L1 loss against a seeded uniform-noise target, and the
backward to every parameter (means, log-scales, quaternions, opacity logits,
SH).  Inputs are resident in HBM before the timed region.

Prints one JSON line: throughput, per-stage durations from HIP events, the
roofline of the dominant kernel and a host baseline of the same algorithm.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+bwd Mpixel/s @1M Gaussians 1080p; served p50/p99 latency vs plain path"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--tile", default="16x16")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baseline")
    ap.add_argument("--torch-activations", action="store_true",
                    help="activate the parameters with torch ops before the splatter instead of inside its kernels")
    ap.add_argument("--cpu-rows", type=int, default=0, help="rows of the frame in the host sample (0: the whole frame)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-camera batched workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--torch-loss", action="store_true", help="L1 loss as torch's composite (sub, abs, mean) instead of the fused kernel")
    ap.add_argument("--no-marcher", action="store_true", help="skip the volumetric marcher workload")
    ap.add_argument("--no-raster", action="store_true", help="skip the volumetric rasteriser workload")
    ap.add_argument("--no-tight", action="store_true", help="skip the tight-culling run of the training step")
    ap.add_argument("--batched-cams", type=int, default=64)
    ap.add_argument("--serving-requests", type=int, default=400)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy and an elementwise multiply of MIB MiB, three times each (calibrates HBM byte counters), and exit")
    return ap.parse_args()


def copy_probe(mib: int):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)  # the runtime's copy kernel
    for _ in range(3):
        torch.mul(src, 2.0, out=dst)  # an elementwise kernel, 16 B per lane access
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches_per_kernel": 3}))


def stage_times(lib):
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    arr = (N.StageTime * 64)()
    count = ctypes.c_int32(0)
    lib.dgmr_profile_read(arr, 64, ctypes.byref(count))
    return {arr[i].name.decode(): (arr[i].total_ms, int(arr[i].launches), arr[i].bytes) for i in range(count.value)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus must be 1 (multi-GPU sharding is a design note, not a run)"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import numpy as np
    import torch

    from differentiable_gaussian_mixture_renderer_amd import _native as N
    from differentiable_gaussian_mixture_renderer_amd import losses, scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter

    tile = tuple(int(v) for v in args.tile.split("x"))
    W, H = args.width, args.height
    scene, cam = scenes.training_frame(n=args.n, width=W, height=H)
    dev = torch.device("cuda:0")
    lib = N.load()
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    means = t(scene.means).requires_grad_()
    log_scales = t(scene.log_scales).requires_grad_()
    quats = t(scene.rotations).requires_grad_()
    logits = t(scene.opacity_logits).requires_grad_()
    sh = t(scene.sh).requires_grad_()
    params = [means, log_scales, quats, logits, sh]
    target = torch.tensor(scenes.Stream(2, 9).uniform(3 * H * W).reshape(3, H, W).astype(np.float32), device=dev)
    # camera tensors stay on the host: no device->host copies inside the step
    view = torch.tensor(cam.world_view_transform)
    proj = torch.tensor(cam.full_proj_transform)
    campos = torch.tensor(cam.camera_center)
    bg = torch.zeros(3)

    def render():
        if args.torch_activations:
            return splatter.Splatter.apply(sh, torch.sigmoid(logits), means, torch.exp(log_scales),
                                           torch.nn.functional.normalize(quats, dim=-1), view, proj, campos, bg, W, H, cam.tan_fovx,
                                           cam.tan_fovy, 3)
        # exp / normalise / sigmoid inside the preprocess and its backward
        return splatter.SplatterActivated.apply(sh, logits, means, log_scales, quats, view, proj, campos, bg, W, H, cam.tan_fovx,
                                                cam.tan_fovy, 3)

    def step():
        img = render()
        # mean |img - target| and its gradient, one fused pass each way (torch's composite with --torch-loss)
        loss = (img - target).abs().mean() if args.torch_loss else losses.l1_loss(img, target)
        return torch.autograd.grad(loss, params)

    if tile != (16, 16) or args.batch != 64:
        splatter.DEFAULT_TILING = (tile, args.batch)

    def timed() -> float:
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    elapsed = timed()
    ms_per_step = 1e3 * elapsed / args.steps
    mpx = W * H * args.steps / elapsed / 1e6

    # per-stage durations from hipEvents on the library's stream, over extra steps
    prof_steps = max(3, min(10, args.steps))
    lib.dgmr_profile_enable(1)
    for _ in range(prof_steps):
        step()
    torch.cuda.synchronize()
    lib.dgmr_profile_enable(0)
    stages = stage_times(lib)

    # pair count of this frame (for the byte model)
    with torch.no_grad():
        cache = splatter.cpp_binding.forward(sh, torch.sigmoid(logits), means, torch.exp(log_scales),
                                             torch.nn.functional.normalize(quats, dim=-1), view, proj, campos, bg, W, H, cam.tan_fovx,
                                             cam.tan_fovy, 3)
    R = int(cache[-1][0])
    n_visible = int((cache[7] > 0).sum())
    del cache

    tight = None
    if not args.no_tight:
        tight = tight_cull_run(args, lib, splatter, step, timed, prof_steps,
                               (sh, logits, means, log_scales, quats, view, proj, campos, bg, W, H, cam))

    dominant = max(stages.items(), key=lambda kv: kv[1][0])
    d_name, (d_ms, d_launches, d_bytes) = dominant
    d_avg_ms = d_ms / max(d_launches, 1)
    achieved = (d_bytes / max(d_launches, 1)) / (d_avg_ms * 1e-3) / 1e9
    p_in = 44 + 12 * 16
    # the shared frame byte model: per pair emission 12, sort 24, ranges 8 and
    # both blend reads 40 + 40 (124 B); with 16-bit pair keys (every view here
    # has fewer than 65 536 tiles) the emission and each of the two sort
    # passes move 2 + 6 + 6 bytes less (110 B)
    frame_rest = args.n * (3 * p_in + 160) + 44 * W * H + 8 * ((W + 15) // 16) * ((H + 15) // 16)
    frame_hbm_frac = (frame_rest + 124 * R) / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9)
    frame_hbm_frac_16bit = (frame_rest + 110 * R) / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9)

    traffic = None
    tr_path = os.path.join(ROOT, "profiles", "traffic_splat_1M_1080p.json")
    if os.path.exists(tr_path):
        with open(tr_path) as f:
            tr = json.load(f)
        traffic = tr.get(d_name)

    adam = adam_workload(params)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, cam, args.cpu_rows)
    roofline = blend_roofline(d_name, d_avg_ms, d_bytes / max(d_launches, 1), achieved, traffic, cpu)
    del means, log_scales, quats, logits, sh, params, target
    torch.cuda.empty_cache()
    batched = None if args.no_batched else batched_workload(args.n, args.batched_cams)
    # each section starts from an empty allocator cache: blocks cached on one
    # stream cannot serve another stream's requests, and a full cache makes
    # the allocator free everything and retry in the middle of a timed run
    torch.cuda.empty_cache()
    served = None if args.no_serving else serving_workload(args.serving_requests)
    torch.cuda.empty_cache()
    marcher = None if args.no_marcher else marcher_workload(lib, with_cpu=not args.no_cpu)
    raster = None if args.no_raster else raster_workload(lib, with_cpu=not args.no_cpu)

    line = {
        "metric": METRIC,
        "value": round(mpx, 3),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene: means U[-5,5]^3, log-scales N(-3,0.5), SH degree 3; uniform L1 target)",
        "config": {"workload": f"tile splat fwd+bwd, {args.n} Gaussians SH3, {W}x{H}, L1 loss", "tile": f"{tile[0]}x{tile[1]}",
                   "lds_batch": args.batch, "pairs": R, "visible": n_visible,
                   "activations": "torch" if args.torch_activations else "fused",
                   "l1_loss": "torch" if args.torch_loss else "fused"},
        "roofline": roofline,
        "frame_hbm_fraction": round(frame_hbm_frac, 4),
        "frame_hbm_fraction_16bit_keys": round(frame_hbm_frac_16bit, 4),
        "stages_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in sorted(stages.items(), key=lambda kv: -kv[1][0])},
        "cpu_baseline": cpu,
        "tight_cull": tight,
        "optimizer_step": adam,
        "batched": batched,
        "serving": served,
        "marcher": marcher,
        "raster": raster,
    }
    print(json.dumps(line))


# FLOP per (pixel, pair) evaluation of the blend: forward 12 + one exp,
# backward 36 + one exp (the exp counted as one operation)
BLEND_FLOP_PER_EVAL = {"splat.blend_forward": 13, "splat.blend_backward": 37}
VALU_PEAK_TFLOPS = 157.3  # MI355X FP32 vector


# chip VALU issue capacity: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 FP32
# instruction per SIMD every 4 cycles (16 lanes; packed FP32 is not used)
VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4


def blend_roofline(name, avg_ms, bytes_per_launch, achieved_gbs, traffic, cpu):
    """Roofline of the dominant stage against both ceilings: algorithmic
    HBM bytes per launch / duration against 8 TB/s, and algorithmic FLOP per
    launch (the oracle's count of (pixel, pair) evaluations of the frame times
    the per-evaluation FLOP) / duration against the FP32 vector peak.
    `traffic` is the measured HBM bytes per launch (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, scaled; profiles/traffic_splat_1M_1080p.json) and
    `valu.issue_utilisation` the measured VALU instructions per launch
    (rocprofv3 SQ_INSTS_VALU; profiles/sq_splat_1M_1080p.json) over the
    chip's issue capacity in the launch's duration.  `bound` names the
    resource the measurements show closer to saturation: the measured HBM
    traffic fraction against the VALU issue utilisation (the algorithmic
    fractions when no measurement is at hand)."""
    hbm = {"achieved_gbs": round(achieved_gbs, 2), "peak_gbs": HBM_PEAK_GBS, "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
           "algorithmic_bytes_per_launch": bytes_per_launch}
    if traffic:
        hbm["measured_traffic_frac"] = round(traffic / (avg_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
    valu = None
    work = (cpu or {}).get("blend_work")
    if work and name in BLEND_FLOP_PER_EVAL:
        flop = work["evaluations"] * BLEND_FLOP_PER_EVAL[name]
        tflops = flop / (avg_ms * 1e-3) / 1e12
        valu = {"achieved_tflops": round(tflops, 3), "peak_tflops": VALU_PEAK_TFLOPS, "frac": round(tflops / VALU_PEAK_TFLOPS, 4),
                "evaluations_per_launch": work["evaluations"], "flop_per_evaluation": BLEND_FLOP_PER_EVAL[name],
                "tile_evaluations_per_launch": work["tile_evaluations"]}
    sq_path = os.path.join(ROOT, "profiles", "sq_splat_1M_1080p.json")
    if os.path.exists(sq_path):
        with open(sq_path) as f:
            sq = json.load(f).get(name)
        if sq:
            valu = valu or {}
            valu["measured_valu_insts_per_launch"] = sq["sq_insts_valu"]
            valu["issue_utilisation"] = round(sq["sq_insts_valu"] / (avg_ms * 1e-3) / VALU_WAVE_INSTS_PER_S, 4)
    hbm_level = hbm.get("measured_traffic_frac", hbm["frac"])
    valu_level = (valu or {}).get("issue_utilisation", (valu or {}).get("frac", 0.0))
    if valu and "frac" in valu and valu_level > hbm_level:
        head = {"bound": "valu", "achieved": valu["achieved_tflops"], "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": valu["frac"]}
    else:
        head = {"bound": "hbm", "achieved": hbm["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm["frac"]}
    return {**head, "kernel": name, "traffic": traffic, "avg_launch_ms": round(avg_ms, 4), "hbm": hbm, "valu": valu}


def tight_cull_run(args, lib, splatter, step, timed, prof_steps, frame):
    """The training step with tight culling (dgmr_tiling.cull = 1): the same
    image and gradients from fewer pairs.  The reference binning stays the
    reported value, since the drop-in returns the reference's point list."""
    import torch
    sh, logits, means, log_scales, quats, view, proj, campos, bg, W, H, cam = frame
    splatter.DEFAULT_CULL = True
    el_cull = timed()
    lib.dgmr_profile_enable(1)
    for _ in range(prof_steps):
        step()
    torch.cuda.synchronize()
    lib.dgmr_profile_enable(0)
    stages_cull = stage_times(lib)
    with torch.no_grad():
        cache = splatter.cpp_binding.forward(sh, torch.sigmoid(logits), means, torch.exp(log_scales),
                                             torch.nn.functional.normalize(quats, dim=-1), view, proj, campos, bg, W, H, cam.tan_fovx,
                                             cam.tan_fovy, 3)
    tight = {"value": round(W * H * args.steps / el_cull / 1e6, 3), "ms_per_step": round(1e3 * el_cull / args.steps, 4),
             "pairs": int(cache[-1][0]),
             "stages_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in sorted(stages_cull.items(), key=lambda kv: -kv[1][0])}}
    del cache
    splatter.DEFAULT_CULL = False
    return tight


def adam_workload(params, reps: int = 20):
    """The optimiser step after the backward: Adam over the five parameter
    tensors of the frame (59 floats per Gaussian), fused kernels against
    torch.optim.Adam (foreach), same gradients."""
    import torch

    from differentiable_gaussian_mixture_renderer_amd.optim import FusedAdam

    def timed(make):
        ps = [p.detach().clone().requires_grad_() for p in params]
        for p in ps:
            p.grad = torch.full_like(p, 1e-3)
        opt = make(ps)
        opt.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            opt.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    fused = timed(lambda ps: FusedAdam(ps, lr=1e-3))
    ref = timed(lambda ps: torch.optim.Adam(ps, lr=1e-3, foreach=True))
    n_el = sum(p.numel() for p in params)
    return {"elements": n_el, "fused_ms": round(fused, 4), "torch_adam_ms": round(ref, 4),
            "fused_gb_s": round(28.0 * n_el / fused / 1e6, 1)}


def batched_workload(n: int, n_cams: int, reps: int = 20):
    """One scene seen by n_cams cameras of uneven resolution (256^2 .. 1080p) on
    the radius-6 sphere: forward + L1 backward of all views through the batched
    path (autotuned tile shape and LDS batch) against the one-camera-per-launch
    plain path, same inputs, same per-view gradients."""
    import numpy as np
    import torch

    from differentiable_gaussian_mixture_renderer_amd import autotune, losses, scenes
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter, render_plain

    scene, cams = scenes.multiview_batch(n=n, n_cams=n_cams)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    gen = torch.Generator(device=dev).manual_seed(4)
    targets = [torch.rand((3, c.height, c.width), generator=gen, device=dev) for c in cams]
    l1 = lambda img, v: losses.l1_grad(img, targets[v])  # sign(img - target) / n in one pass
    px = sum(c.width * c.height for c in cams)

    (tile, lds), table = autotune.autotune(params, cams, reps=1, use_cache=False)
    bs = BatchSplatter(tile=tile, batch=lds)

    def batched_step():
        # forward, then the backward with the L1 gradient formed inside the blend backward
        return bs.step_l1(*params, cams, targets)

    batched_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fbs, gb = batched_step()
    torch.cuda.synchronize()
    ms_b = (time.perf_counter() - t0) / reps * 1e3
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    lib.dgmr_profile_read((N.StageTime * 64)(), 64, ctypes.byref(ctypes.c_int32(0)))  # drop older records
    lib.dgmr_profile_enable(1)
    batched_step()
    torch.cuda.synchronize()
    lib.dgmr_profile_enable(0)
    stages = {k: round(v[0] / max(v[1], 1), 3) for k, v in stage_times(lib).items()}

    def plain_step(cull=False):
        it = iter(range(len(cams)))
        return render_plain(*params, cams, grad_fn=lambda img: l1(img, next(it)), cull=cull)

    def time_plain(cull):
        plain_step(cull)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = plain_step(cull)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, out

    ms_p, (imgs, gp) = time_plain(False)
    img_diff = max(float((a - b).abs().max()) for a, b in zip(fbs, imgs))
    grad_rel = max(float((a - b).norm() / b.norm()) for a, b in zip(gb, gp))
    return {"workload": f"{n_cams} cameras of one {n}-Gaussian SH3 scene, resolutions 256^2..1920x1080, fwd+L1 bwd",
            "mpixels_per_batch": round(px / 1e6, 3), "pairs": bs.n_pairs, "tile": f"{tile[0]}x{tile[1]}", "lds_batch": lds,
            "cull": "tight (alpha 1/255 reach)" if bs.tiling.cull else "rectangles",
            "autotune_ms": table, "batched_ms": round(ms_b, 3), "batched_stages_ms": stages, "plain_ms": round(ms_p, 3),
            "batched_mpx_s": round(px / ms_b / 1e3, 1), "plain_mpx_s": round(px / ms_p / 1e3, 1),
            "speedup_vs_plain": round(ms_p / ms_b, 3),
 "max_image_diff_vs_plain": img_diff, "max_grad_rel_l2_vs_plain": grad_rel}


def marcher_workload(lib, with_cpu: bool, steps: int = 3, max_large_steps: int = 2):
    """Volumetric ray marcher: 200k Gaussians, 1024x1024 rays, at most
    max_large_steps x 127 bins per ray (256 marching steps), fp32 forward +
    L1 backward through VolMarcher.apply, with the default backward (every
    bin recomputed, as the reference's) and with the opt-in replay store (the
    forward's bins read back).  Per-stage durations from the stage profile;
    the host baseline is the float restatement on a band of rows."""
    import numpy as np
    import torch

    from differentiable_gaussian_mixture_renderer_amd import losses, scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher

    scene, cam = scenes.marcher_frame()
    W, H = cam.width, cam.height
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh).requires_grad_(), t(scene.opacity).requires_grad_(), t(scene.means).requires_grad_(),
              t(scene.scales).requires_grad_(), t(scene.rotations).requires_grad_()]
    target = torch.rand((3, H, W), generator=torch.Generator(device=dev).manual_seed(3), device=dev)
    cam_args = (torch.tensor(cam.world_view_transform), torch.tensor(cam.full_proj_transform), torch.tensor(cam.camera_center),
                torch.zeros(3), W, H, cam.tan_fovx, cam.tan_fovy, scene.sh_degree)
    saved = dict(vol_marcher.DEFAULT_OPTIONS)

    def step():
        img = vol_marcher.VolMarcher.apply(*params, *cam_args)
        return torch.autograd.grad(losses.l1_loss(img, target), params)

    def measure(replay: bool):
        vol_marcher.DEFAULT_OPTIONS["max_large_steps"] = max_large_steps
        vol_marcher.DEFAULT_OPTIONS["replay"] = replay
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        lib.dgmr_profile_enable(1)
        step()
        torch.cuda.synchronize()
        lib.dgmr_profile_enable(0)
        stages = {k: round(v[0] / max(v[1], 1), 3) for k, v in stage_times(lib).items() if k.startswith("march")}
        return {"ms_per_step": round(ms, 3), "mrays_per_s": round(W * H / ms / 1e3, 2), "stages_ms": stages}

    try:
        plain = measure(False)
        stored = measure(True)
        replay = vol_marcher.make_replay(W, H, torch.float32, dev, steps=max_large_steps)
        stored["replay_store"] = None if replay is None else {"large_steps": replay.steps, "bytes": replay.buffer.numel()}
        del replay
    finally:
        vol_marcher.DEFAULT_OPTIONS.clear()
        vol_marcher.DEFAULT_OPTIONS.update(saved)
    out = {"workload": f"volumetric marcher fwd+bwd, {scene.n} Gaussians SH3, {W}x{H} rays, <= {max_large_steps * 127} bins per ray, L1",
           **plain, "with_replay_store": dict(stored, note="opt-in: the forward keeps each pixel's bins for the backward")}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        rows = 16
        y0 = (H - rows) // 2
        bg = np.zeros(3, np.float32)
        grad = np.full((3, H, W), 1.0 / (3 * H * W), np.float32)
        t0 = time.perf_counter()
        h = O.MarchOracle(O.scene_arrays(scene), cam, dbl=False)
        fb, T = h.render(bg, crop=(0, y0, W, y0 + rows), max_large_steps=max_large_steps)[:2]
        h.backward(bg, fb, T, grad, crop=(0, y0, W, y0 + rows), max_large_steps=max_large_steps)
        secs = time.perf_counter() - t0
        threads = os.environ.get("OMP_NUM_THREADS", str(os.cpu_count()))
        out["cpu_baseline"] = {"value": round(W * rows / secs / 1e6, 5), "unit": "Mrays/s", "kind": "port", "cores": int(threads),
                               "sample": f"float restatement: preprocess+binning of all Gaussians, forward and backward of {W}x{rows} "
                                         f"rays, on {threads} OpenMP threads, {secs:.1f} s"}
    return out


def raster_workload(lib, with_cpu: bool, steps: int = 3):
    """Volumetric rasteriser (forward only, as the reference): 200k Gaussians,
    1024x1024 pixels, 32 bins per pixel.  The host baseline is the float
    restatement on a band of rows."""
    import numpy as np
    import torch

    from differentiable_gaussian_mixture_renderer_amd import scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings.vol_raster import vol_raster_forward

    scene, cam = scenes.marcher_frame()
    W, H = cam.width, cam.height
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
    args = (t(scene.sh), t(np.reshape(scene.opacity, (-1, 1))), t(scene.means), t(scene.scales), t(scene.rotations),
            t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(np.zeros(3)), W, H, cam.tan_fovx,
            cam.tan_fovy, scene.sh_degree)
    vol_raster_forward(*args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        vol_raster_forward(*args)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    lib.dgmr_profile_enable(1)
    vol_raster_forward(*args)
    torch.cuda.synchronize()
    lib.dgmr_profile_enable(0)
    stages = {k: round(v[0] / max(v[1], 1), 3) for k, v in stage_times(lib).items() if k.startswith(("raster", "bin."))}
    out = {"workload": f"volumetric rasteriser forward, {scene.n} Gaussians SH3, {W}x{H}, 32 bins per pixel",
           "ms_per_frame": round(ms, 3), "mpixel_per_s": round(W * H / ms / 1e3, 2), "stages_ms": stages}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        rows = 32
        y0 = (H - rows) // 2
        t0 = time.perf_counter()
        O.raster_render(O.scene_arrays(scene), cam, np.zeros(3, np.float32), crop=(0, y0, W, y0 + rows))
        secs = time.perf_counter() - t0
        threads = os.environ.get("OMP_NUM_THREADS", str(os.cpu_count()))
        out["cpu_baseline"] = {"value": round(W * rows / secs / 1e6, 5), "unit": "Mpixel/s", "kind": "port",
                               "sample": f"float restatement: preprocess+binning of all Gaussians and {W}x{rows} pixels on "
                                         f"{threads} OpenMP threads, {secs:.1f} s"}
    return out


def serving_workload(n_requests: int, rate_hz: float = 200.0):
    """Poisson arrivals at rate_hz over 8 resident scenes (50k..3M Gaussians,
    log-spaced), cameras 256^2..1080p, 30 % forward+backward: continuous
    batching against a FIFO server that renders one request at a time through
    the plain path, at rate_hz, at 1000/s and 2000/s, and with every request
    present at t=0 (saturated, median of three runs), which measures
    sustainable throughput."""
    import copy

    import numpy as np
    import torch

    from differentiable_gaussian_mixture_renderer_amd import serving

    resident = serving.make_resident_scenes(8, seed=5)
    trace = serving.make_trace(n_requests, rate_hz, 8, seed=5)
    out = {"workload": f"{n_requests} requests, Poisson {rate_hz:g}/s, 8 resident scenes of 50k..3M Gaussians (log-spaced), "
                       "256^2..1080p, 30% fwd+bwd", "scenes_gaussians": [sc.n for sc in resident]}
    budget = 48 << 30
    retries0 = torch.cuda.memory_stats().get("num_alloc_retries", 0)

    def run(server_cls, reqs, **kw):
        srv = server_cls(resident, **kw)
        st = srv.run(reqs)
        torch.cuda.synchronize()
        return st.summary()

    def measured(server_cls, reqs, **kw):
        # an untimed pass over the same trace first grows the server's cached
        # workspaces (and capacities) to the trace's largest requests
        run(server_cls, copy.deepcopy(reqs), **kw)
        return run(server_cls, copy.deepcopy(reqs), **kw)

    out["continuous"] = measured(serving.ContinuousBatchingServer, trace, hbm_budget_bytes=budget)
    out["fifo_plain"] = measured(serving.FifoServer, trace)
    # latency and throughput against the arrival rate: the same scenes and
    # request mix at higher Poisson rates
    sweep = {}
    for rate in (1000.0, 2000.0):
        tr = serving.make_trace(n_requests, rate, 8, seed=5)
        sweep[f"{rate:g}/s"] = {"continuous": measured(serving.ContinuousBatchingServer, tr, hbm_budget_bytes=budget),
                                "fifo_plain": measured(serving.FifoServer, tr)}
    out["rate_sweep"] = sweep
    sat = copy.deepcopy(trace)
    for r in sat:
        r.arrival = 0.0

    def saturated(server_cls, **kw):
        # the saturated batches are the largest: their workspaces start from an
        # empty cache (blocks cached on other streams by the runs above cannot
        # serve them and would push the allocator into a free-everything retry),
        # and one untimed run grows every worker's workspace to them; a
        # saturated run lasts a fraction of a second: the median of three
        torch.cuda.empty_cache()
        run(server_cls, copy.deepcopy(sat), **kw)
        runs = sorted((run(server_cls, copy.deepcopy(sat), **kw) for _ in range(3)), key=lambda st: st["mpixel_per_s"])
        med = dict(runs[1])
        med["mpixel_per_s_runs"] = [round(st["mpixel_per_s"], 1) for st in runs]
        return med

    out["continuous_saturated"] = saturated(serving.ContinuousBatchingServer, hbm_budget_bytes=budget)
    out["fifo_plain_saturated"] = saturated(serving.FifoServer)
    out["saturated_throughput_ratio"] = round(out["continuous_saturated"]["mpixel_per_s"] / out["fifo_plain_saturated"]["mpixel_per_s"], 3)
    # times the caching allocator had to free its cache and retry (each stalls every stream)
    out["allocator_retries"] = int(torch.cuda.memory_stats().get("num_alloc_retries", 0) - retries0)
    return out


def cpu_baseline(scene, cam, rows: int):
    """The float restatement of the same algorithm (oracle/) timed on the host:
    preprocess and binning of every Gaussian, blend forward and backward of
    the frame's pixels (or a band of `rows` rows), preprocess backward.  Run
    with all host threads the GPU's share allows (OpenMP, at most 16) and with
    one thread."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    H, W = cam.height, cam.width
    rows = min(rows or H, H)
    y0 = (H - rows) // 2
    h = O.SplatOracle(O.scene_arrays(scene), cam, dbl=False)
    grad = np.full((3, H, W), 1.0 / (3 * H * W), np.float32)
    bg = np.zeros(3, np.float32)
    threads = max(1, min(16, os.cpu_count() or 1))
    secs_mt = h.time_fwd_bwd(bg, grad, crop=(0, y0, W, y0 + rows), threads=threads)
    secs_st = h.time_fwd_bwd(bg, grad, crop=(0, y0, W, y0 + rows), threads=1)
    # the blend's algorithmic work on the whole frame: (pixel, pair)
    # evaluations up to each pixel's stop, and list entries a tile traverses
    # until all its pixels have stopped (times the tile's pixels)
    _, _, _, n_evals = h.render(bg)
    traversed, _ = h.tile_work()
    tile_px = h.tile[0] * h.tile[1]
    px = W * rows
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    sample = (f"float restatement (g++ -O3, OpenMP): preprocess+binning of {scene.n} Gaussians, blend fwd+bwd of {W}x{rows} "
              f"pixels, preprocess bwd; {secs_mt:.1f} s on {threads} threads, {secs_st:.1f} s on 1 thread; host: {model}")
    return {"value": round(px / secs_mt / 1e6, 4), "unit": "Mpixel/s", "cores": threads, "kind": "port", "sample": sample,
            "host_cpus": os.cpu_count(), "host_model": model,
            "single_thread": {"value": round(px / secs_st / 1e6, 4), "unit": "Mpixel/s", "cores": 1},
            "blend_work": {"evaluations": int(n_evals), "tile_evaluations": int(traversed.astype(np.int64).sum()) * tile_px}}


if __name__ == "__main__":
    main()

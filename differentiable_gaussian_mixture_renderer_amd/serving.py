"""Continuous-batching render server over scenes resident in HBM.

Requests (scene id, camera, forward-only or forward+backward) arrive on an
open-loop trace.  A scheduler thread admits waiting requests while the
estimated HBM footprint of everything in flight stays under a budget, groups
the admitted requests of one scene into a batch, and hands the batch to the
next free worker.  Each worker owns a HIP stream and runs the batch through the
multi-view path (BatchSplatter: one preprocess, one binning, one persistent
blend launch); workers run concurrently, so batches of different scenes, and
new arrivals admitted while others are on the GPU, overlap on the device.
Latency is completion time minus arrival time.

`FifoServer` is the plain baseline: one request at a time through the
one-camera-per-launch Splatter path, in arrival order.
"""
from __future__ import annotations

import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import losses, scenes
from .batching import BatchSplatter

# bytes of device memory per (view, Gaussian) slot of the batched path and per
# (tile, Gaussian) pair, used for admission: geometry 91 (rects, tiles, xy,
# conic, rgb, clamp mask, depth, screen gradients) + binning 40 (depth keys
# and ids twice, bin and owner rectangles, counts, offsets) + sort scratch
BYTES_PER_VIEW_GAUSSIAN = 136
BYTES_PER_PAIR = 24
PAIRS_PER_PIXEL = 8
# the gradients a backward request returns: SH 48 + means 3 + scales 3 +
# rotations 4 + opacity 1 floats per Gaussian
BYTES_PER_GRADIENT_GAUSSIAN = 236


@dataclass
class Request:
    rid: int
    scene: int
    camera: scenes.Camera
    backward: bool
    arrival: float            # seconds on the server clock
    target_seed: int = 0
    done_at: float = -1.0
    image: Optional[torch.Tensor] = None
    grads: Optional[list] = None

    @property
    def pixels(self) -> int:
        return self.camera.width * self.camera.height


@dataclass
class ResidentScene:
    params: list               # sh, opacity, means, scales, rotations on the device
    sh_degree: int
    # the densest (tile, Gaussian) pairs per pixel a batch of this scene has
    # had, for admission (at least PAIRS_PER_PIXEL)
    pairs_per_pixel: float = PAIRS_PER_PIXEL

    @property
    def n(self) -> int:
        return self.params[2].shape[0]


def resident_sizes(count: int, n_min: int = 50_000, n_max: int = 3_000_000) -> List[int]:
    """Gaussian counts of `count` resident scenes spread log-uniformly over
    [n_min, n_max]: the quantiles k / (count - 1) of the log-uniform
    distribution, so the smallest scene has n_min Gaussians and the largest
    n_max."""
    if count == 1:
        return [int(round(math.sqrt(n_min * n_max)))]
    lo, hi = math.log(n_min), math.log(n_max)
    return [int(round(math.exp(lo + k * (hi - lo) / (count - 1)))) for k in range(count)]


def make_resident_scenes(count: int, seed: int, n_min: int = 50_000, n_max: int = 3_000_000, device="cuda:0") -> List[ResidentScene]:
    """Scenes whose Gaussian counts span [n_min, n_max] log-uniformly
    (resident_sizes), each with the 1M-Gaussian training-frame distributions."""
    out = []
    for k, n in enumerate(resident_sizes(count, n_min, n_max)):
        sc = scenes.make_scene(n, seed=1000 * seed + k, extent=5.0, sh_degree=3)
        t = lambda a: torch.tensor(np.ascontiguousarray(a), device=device)
        out.append(ResidentScene([t(sc.sh), t(sc.opacity), t(sc.means), t(sc.scales), t(sc.rotations)], sc.sh_degree))
    return out


def make_trace(n_requests: int, rate_hz: float, n_scenes: int, seed: int, backward_fraction: float = 0.3,
               min_res=(256, 256), max_res=(1920, 1080)) -> List[Request]:
    """Poisson arrivals; random camera on the radius-6 sphere; width and height
    uniform between min_res and max_res; forward+backward with the given
    probability."""
    st = scenes.Stream(seed, 12)
    gaps = -np.log1p(-st.uniform(n_requests)) / rate_hz
    arrivals = np.cumsum(gaps) - gaps[0]
    pick = (st.uniform(n_requests) * n_scenes).astype(np.int64)
    bwd = st.uniform(n_requests) < backward_fraction
    ws = (min_res[0] + st.uniform(n_requests) * (max_res[0] - min_res[0])).astype(np.int64)
    hs = (min_res[1] + st.uniform(n_requests) * (max_res[1] - min_res[1])).astype(np.int64)
    z = st.uniform(n_requests) * 2 - 1
    phi = st.uniform(n_requests) * 2 * math.pi
    ty = math.tan(math.radians(30.0))
    reqs = []
    for i in range(n_requests):
        r = math.sqrt(max(0.0, 1 - z[i] * z[i]))
        pos = 6.0 * np.array([r * math.cos(phi[i]), r * math.sin(phi[i]), z[i]])
        w, h = int(ws[i]), int(hs[i])
        cam = scenes.look_at_camera(pos, [0, 0, 0], w, h, ty * w / h, ty)
        reqs.append(Request(i, int(pick[i]), cam, bool(bwd[i]), float(arrivals[i]), target_seed=i))
    return reqs


def l1_target(img: torch.Tensor, seed: int) -> torch.Tensor:
    """The seeded uniform-noise target image of a gradient request."""
    g = torch.Generator(device=img.device)
    g.manual_seed(seed)
    return torch.rand(img.shape, generator=g, device=img.device, dtype=img.dtype)


def l1_grad(img: torch.Tensor, seed: int) -> torch.Tensor:
    """dL/dimg of L = mean |img - target| for a seeded uniform-noise target."""
    return losses.l1_grad(img, l1_target(img, seed))


def estimate_bytes(scene: ResidentScene, reqs: List[Request]) -> int:
    px = sum(r.pixels for r in reqs)
    n_bwd = sum(r.backward for r in reqs)
    return int(len(reqs) * scene.n * BYTES_PER_VIEW_GAUSSIAN + px * (scene.pairs_per_pixel * BYTES_PER_PAIR + 32)
               + n_bwd * scene.n * BYTES_PER_GRADIENT_GAUSSIAN)


@dataclass
class ServerStats:
    latencies: List[float] = field(default_factory=list)
    pixels: int = 0
    wall: float = 0.0
    batches: int = 0
    max_in_flight_bytes: int = 0

    def summary(self) -> Dict[str, float]:
        lat = np.array(self.latencies) * 1e3
        return {"p50_ms": float(np.percentile(lat, 50)) if lat.size else float("nan"),
                "p99_ms": float(np.percentile(lat, 99)) if lat.size else float("nan"),
                "mpixel_per_s": self.pixels / self.wall / 1e6 if self.wall > 0 else 0.0,
                "requests": len(self.latencies), "batches": self.batches, "wall_s": self.wall}


# Worker streams, one per (device, worker index), kept for the process: the
# caching allocator reuses a freed block only on the stream that allocated
# it, so fresh streams per server would leave every earlier server's
# workspaces unusable (reserved memory grew by 60 GB per saturated run until
# an allocation retry freed everything, a 5.8 s stall).
_STREAMS = {}
_STREAMS_LOCK = threading.Lock()


def _worker_stream(device: torch.device, k: int) -> torch.cuda.Stream:
    with _STREAMS_LOCK:
        key = (str(device), k)
        st = _STREAMS.get(key)
        if st is None:
            st = _STREAMS[key] = torch.cuda.Stream(device=device)  # default priority
        return st


def _inline_stream(device: torch.device):
    return torch.cuda.current_stream(device)


# Each worker's multi-view splatter, kept with its stream for the process:
# its workspace only grows, so servers after the first reuse it instead of
# allocating (and caching) another one per run.
_SPLATTERS = {}


def _worker_splatter(device: torch.device, k: int, tile, lds_batch: int) -> BatchSplatter:
    with _STREAMS_LOCK:
        key = (str(device), k, tuple(tile), int(lds_batch))
        sp = _SPLATTERS.get(key)
        if sp is None:
            sp = _SPLATTERS[key] = BatchSplatter(tile=tile, batch=lds_batch, device=device)
        return sp


class ContinuousBatchingServer:
    def __init__(self, resident: List[ResidentScene], hbm_budget_bytes: int, workers: int = 3, max_batch_views: int = 64,
                 tile=(16, 16), lds_batch: int = 64, device="cuda:0", overlap_backlog: int = 8, min_batched_views: int = 3):
        self.resident = resident
        self.budget = int(hbm_budget_bytes)
        self.max_views = max_batch_views
        # batches form and overlap on the worker streams only while at least
        # this many requests wait; below it requests run one at a time, as on
        # the plain path, so a request neither waits for a batch to finish
        # nor shares the GPU with another
        self.overlap_backlog = int(overlap_backlog)
        self.min_batched_views = max(1, int(min_batched_views))
        self._batches = 0  # batches queued or running
        self.device = torch.device(device)
        self.tile, self.lds_batch = tile, lds_batch
        self.n_workers = workers
        self.stats = ServerStats()
        # one mutex, two conditions: the scheduler waits on _lock (arrivals,
        # completions), the workers on _work_ready, so finishing a request
        # wakes the scheduler alone instead of every idle worker
        mutex = threading.Lock()
        self._lock = threading.Condition(mutex)
        self._work_ready = threading.Condition(mutex)
        self._waiting: deque = deque()
        self._in_flight = 0
        self._work: deque = deque()
        self._stop = False
        self._error: Optional[BaseException] = None  # the first exception a worker raised
        self._t0 = 0.0

    # ------------------------------------------------------------------ clock
    def now(self) -> float:
        return time.perf_counter() - self._t0

    # ---------------------------------------------------------------- workers
    def _worker(self, k: int):
        """Runs batches until stopped; the first exception (a DgmrError, an
        out-of-memory error, a failure to set up the stream) is recorded, the
        queued work is dropped and run() re-raises it."""
        try:
            self._worker_loop(k)
        except BaseException as exc:  # noqa: BLE001 -- handed to run(), which re-raises it
            with self._lock:
                if self._error is None:
                    self._error = exc
                self._stop = True
                for _, _, cost in self._work:
                    self._in_flight -= cost
                    self._batches -= 1
                self._work.clear()
                self._lock.notify_all()
                self._work_ready.notify_all()

    def _worker_loop(self, k: int):
        stream = _worker_stream(self.device, k)
        splatter = _worker_splatter(self.device, k, self.tile, self.lds_batch)
        while True:
            with self._lock:
                while not self._work and not self._stop:
                    self._work_ready.wait()
                if not self._work and self._stop:
                    return
                scene_id, batch, cost = self._work.popleft()
            self._run_batch(stream, splatter, scene_id, batch, cost)

    def _run_batch(self, stream, splatter, scene_id, batch, cost):
        """Renders one admitted batch on `stream`, records its results and
        latencies, and releases its admission cost (also when it fails).
        Batches of fewer than min_batched_views requests (low load) take the
        one-camera path request by request, each completing on its own: the
        multi-view path's fixed cost only pays off over more views."""
        # gradient requests first, one gradient group each; forward-only views after them
        batch = [r for r in batch if r.backward] + [r for r in batch if not r.backward]
        n_bwd = sum(r.backward for r in batch)
        try:
            with torch.cuda.stream(stream):
                sc = self.resident[scene_id]
                if len(batch) < self.min_batched_views:
                    for r in sorted(batch, key=lambda q: q.arrival):
                        imgs, grads = self._single(sc, r)
                        stream.synchronize()
                        self._complete([r], imgs, grads)
                else:
                    imgs, grads = self._batched(splatter, sc, batch, n_bwd)
                    stream.synchronize()
                    self._complete(batch, imgs, grads)
        except BaseException:
            with self._lock:
                self._in_flight -= cost  # the failed batch's admission is released
                self._batches -= 1
            raise
        with self._lock:
            self._batches -= 1
            self.stats.batches += 1
            self._in_flight -= cost
            self._lock.notify_all()

    def _complete(self, reqs, imgs, grads):
        t = self.now()
        with self._lock:
            for r, img, g in zip(reqs, imgs, grads):
                r.done_at = t
                r.image = img
                r.grads = g
                self.stats.latencies.append(t - r.arrival)
                self.stats.pixels += r.pixels

    def _single(self, sc, r):
        from .python_bindings import splatter
        cam = r.camera
        args = (*sc.params[:5], torch.tensor(cam.world_view_transform), torch.tensor(cam.full_proj_transform),
                torch.tensor(cam.camera_center), torch.zeros(3), cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, sc.sh_degree)
        cache = splatter.cpp_binding.forward(*args, tile=self.tile, batch=self.lds_batch)
        grads = None
        if r.backward:
            grads = tuple(splatter.cpp_binding.backward(*args, l1_grad(cache[0], r.target_seed), cache, tile=self.tile, batch=self.lds_batch))
        return [cache[0]], [grads]

    @staticmethod
    def _batched(splatter, sc, batch, n_bwd):
        # the spatial slot order saves per-(view, Gaussian) work and costs
        # scattered gradient rows per gradient group: over the saturated
        # serving trace (three workers) batches took 88.7 ms of wall time
        # with it from four views per group, 90.6 from eight, 89.6 never,
        # 95.1 from two and 97.3 always
        splatter.spatial_order = len(batch) >= 4 * max(n_bwd, 1)
        imgs, _ = splatter.forward(*sc.params, [r.camera for r in batch], sh_degree=sc.sh_degree)
        # later admissions of this scene count its measured pair density
        sc.pairs_per_pixel = max(sc.pairs_per_pixel, splatter.n_pairs / max(1, sum(r.pixels for r in batch)))
        grads = [None] * len(batch)
        if n_bwd:
            # the L1 gradient is formed inside the blend backward
            targets = [l1_target(img, r.target_seed) for img, r in zip(imgs, batch[:n_bwd])] + [None] * (len(batch) - n_bwd)
            groups = [k if k < n_bwd else -1 for k in range(len(batch))]
            grads[:n_bwd] = splatter.backward_l1(targets, groups=groups)
        return imgs, grads

    # -------------------------------------------------------------- scheduler
    def _schedule(self):
        """Admits the waiting requests of the scene with the oldest waiting
        request, as many as fit the budget and the batch limit, and queues
        them for the workers."""
        taken = self._take_batch()
        if taken is None:
            return False
        self._work.append(taken)
        self._work_ready.notify()
        return True

    def _take_oldest(self):
        """(scene id, [the oldest waiting request], admitted bytes), removed
        from the waiting queue and counted in flight (nothing else is)."""
        r = self._waiting.popleft()
        cost = estimate_bytes(self.resident[r.scene], [r])
        self._in_flight += cost
        self._batches += 1
        self.stats.max_in_flight_bytes = max(self.stats.max_in_flight_bytes, self._in_flight)
        return r.scene, [r], cost

    def _take_batch(self):
        """(scene id, requests, admitted bytes) of the next batch, removed
        from the waiting queue and counted in flight; None if nothing fits."""
        if not self._waiting:
            return None
        scene_id = self._waiting[0].scene
        picked, rest = [], deque()
        while self._waiting:
            r = self._waiting.popleft()
            if r.scene == scene_id and len(picked) < self.max_views:
                cost = estimate_bytes(self.resident[scene_id], picked + [r])
                if self._in_flight + cost <= self.budget or (not picked and self._in_flight == 0):
                    picked.append(r)
                    continue
            rest.append(r)
        self._waiting = rest
        if not picked:
            return None
        cost = estimate_bytes(self.resident[scene_id], picked)
        self._in_flight += cost
        self._batches += 1
        self.stats.max_in_flight_bytes = max(self.stats.max_in_flight_bytes, self._in_flight)
        return scene_id, picked, cost

    def run(self, trace: List[Request]) -> ServerStats:
        """Serves the trace: arrivals are admitted on their arrival time.
        While fewer than overlap_backlog requests wait, the oldest runs alone
        on this thread through the one-camera path, without a hand-off to a
        worker, so at low load a request waits no longer than on the plain
        path; once a backlog builds, batches of one scene's requests go to the
        worker streams and overlap on the GPU.  Re-raises the first error a
        batch raised."""
        self._t0 = time.perf_counter()
        workers = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in range(self.n_workers)]
        for w in workers:
            w.start()
        # inline batches run on the calling thread's current stream (one
        # request there took 0.07 ms less on average than on a worker's
        # stream, 0.2 ms less than the plain server's polling loop); they run
        # only while no worker holds a batch, so they may borrow worker 0's
        # splatter: its workspace is idle, and each batch ends with a stream
        # synchronisation before the next is scheduled
        inline_stream = _inline_stream(self.device)
        pending = deque(sorted(trace, key=lambda r: r.arrival))
        total = len(pending)
        while True:
            inline = None
            with self._lock:
                if self._error is not None:
                    break
                t = self.now()
                while pending and pending[0].arrival <= t:
                    self._waiting.append(pending.popleft())
                # below the backlog threshold requests run one at a time, the
                # oldest first, on the one-camera path (its latency beats a
                # small batch's, whose requests all finish with the last);
                # with a backlog, batches go to every worker and the rest keep
                # waiting and join later batches
                backlog = len(self._waiting) >= self.overlap_backlog
                if self._batches == 0 and self._waiting and not backlog:
                    inline = self._take_oldest()
                elif backlog:
                    while self._batches < self.n_workers and self._schedule():
                        pass
                if inline is None:
                    if not pending and not self._waiting and not self._work and len(self.stats.latencies) == total:
                        break
                    timeout = (pending[0].arrival - t) if pending else 0.005
                    self._lock.wait(timeout=max(0.0002, min(timeout, 0.005)))
            if inline is not None:
                try:
                    self._run_batch(inline_stream, _worker_splatter(self.device, 0, self.tile, self.lds_batch), *inline)
                except BaseException as exc:  # noqa: BLE001 -- re-raised below, as a worker's error
                    with self._lock:
                        if self._error is None:
                            self._error = exc
                    break
        with self._lock:
            self._stop = True
            self._work_ready.notify_all()
        for w in workers:
            w.join()
        self.stats.wall = self.now()
        if self._error is not None:
            # a failed batch (e.g. a DgmrError or an out-of-memory error) ends the run instead of leaving it waiting
            raise RuntimeError(f"serving worker failed: {self._error!r}") from self._error
        return self.stats


class FifoServer:
    """One request at a time, in arrival order, one camera per launch."""

    def __init__(self, resident: List[ResidentScene], tile=(16, 16), lds_batch: int = 64):
        self.resident = resident
        self.tile, self.lds_batch = tile, lds_batch
        self.stats = ServerStats()

    def run(self, trace: List[Request]) -> ServerStats:
        from .python_bindings import splatter
        t0 = time.perf_counter()
        bg = torch.zeros(3)
        for r in sorted(trace, key=lambda q: q.arrival):
            while time.perf_counter() - t0 < r.arrival:
                time.sleep(0.0002)
            sc = self.resident[r.scene]
            cam = r.camera
            args = (*sc.params[:5], torch.tensor(cam.world_view_transform), torch.tensor(cam.full_proj_transform),
                    torch.tensor(cam.camera_center), bg, cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, sc.sh_degree)
            cache = splatter.cpp_binding.forward(*args, tile=self.tile, batch=self.lds_batch)
            r.image = cache[0]
            if r.backward:
                r.grads = splatter.cpp_binding.backward(*args, l1_grad(cache[0], r.target_seed), cache, tile=self.tile,
                                                        batch=self.lds_batch)
            torch.cuda.synchronize()
            r.done_at = time.perf_counter() - t0
            self.stats.latencies.append(r.done_at - r.arrival)
            self.stats.pixels += r.pixels
            self.stats.batches += 1
        self.stats.wall = time.perf_counter() - t0
        return self.stats

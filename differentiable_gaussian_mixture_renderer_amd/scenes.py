"""Deterministic synthetic scenes and INRIA-convention cameras.

Random numbers come from a counter-based SplitMix64 stream (Steele, Lea and
Flood, "Fast splittable pseudorandom number generators", 2014) with
Box-Muller normals, computed in numpy uint64 arithmetic, so every machine and
numpy version produces the same scene for a seed.

Camera convention (as the reference's callers, python_bindings/vol_marcher.cu:32-40):
world_view_transform = W2C^T and full_proj_transform = (P W2C)^T, where W2C maps
world to a camera frame with x right, y down, z forward, and P is the
projection with w = z_view.  The reference reads tensor row i as glm column i.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SH_C0 = 0.28209479177387814
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64: draw k is mix(seed_state + (k + 1) * golden)."""

    def __init__(self, seed: int, stream: int = 0):
        with np.errstate(over="ignore"):
            self.state = _mix(np.array([np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream)], dtype=np.uint64))[0]
        self.counter = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
            out = _mix(self.state + k * _GOLDEN)
        self.counter += n
        return out

    def uniform(self, n: int) -> np.ndarray:
        """doubles in [0, 1)"""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        th = 2.0 * math.pi * u2
        return np.concatenate([r * np.cos(th), r * np.sin(th)])[:n]


@dataclass
class Scene:
    means: np.ndarray      # [n,3] float32
    scales: np.ndarray     # [n,3] float32 (activated: exp of the log-scales)
    rotations: np.ndarray  # [n,4] float32 unit quaternions (w,x,y,z)
    opacity: np.ndarray    # [n,1] float32 (activated: sigmoid)
    sh: np.ndarray         # [n,K,3] float32
    sh_degree: int
    log_scales: np.ndarray
    opacity_logits: np.ndarray

    @property
    def n(self) -> int:
        return self.means.shape[0]


def make_scene(n: int, seed: int, extent: float = 5.0, sh_degree: int = 3, sh_mode: str = "sh3") -> Scene:
    """Gaussians with means U[-extent, extent]^3, log-scales N(-3, 0.5),
    rotations normalize(N(0,1)^4), opacity sigmoid(N(0,1)).  sh_mode "rgb":
    degree-0 colour U[0,1] stored as DC = (rgb - 0.5) / SH_C0; "sh3": DC U[0,1],
    the other 45 coefficients N(0, 0.2)."""
    means = (Stream(seed, 1).uniform(3 * n) * 2.0 - 1.0) * extent
    log_scales = Stream(seed, 2).normal(3 * n) * 0.5 - 3.0
    q = Stream(seed, 3).normal(4 * n).reshape(n, 4)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    logits = Stream(seed, 4).normal(n)
    if sh_mode == "rgb":
        rgb = Stream(seed, 5).uniform(3 * n).reshape(n, 1, 3)
        sh = (rgb - 0.5) / SH_C0
        sh_degree = 0
    else:
        k = (sh_degree + 1) ** 2
        sh = np.empty((n, k, 3))
        sh[:, 0, :] = Stream(seed, 5).uniform(3 * n).reshape(n, 3)
        if k > 1:
            sh[:, 1:, :] = (Stream(seed, 6).normal(3 * n * (k - 1)) * 0.2).reshape(n, k - 1, 3)
    f32 = np.float32
    return Scene(
        means=means.reshape(n, 3).astype(f32),
        scales=np.exp(log_scales).reshape(n, 3).astype(f32),
        rotations=q.astype(f32),
        opacity=(1.0 / (1.0 + np.exp(-logits))).reshape(n, 1).astype(f32),
        sh=sh.astype(f32),
        sh_degree=sh_degree,
        log_scales=log_scales.reshape(n, 3).astype(f32),
        opacity_logits=logits.reshape(n, 1).astype(f32),
    )


@dataclass
class Camera:
    world_view_transform: np.ndarray  # [4,4] float32, tensor layout
    full_proj_transform: np.ndarray   # [4,4] float32
    camera_center: np.ndarray         # [3] float32
    tan_fovx: float
    tan_fovy: float
    width: int
    height: int


def projection_matrix(znear: float, zfar: float, tan_fovx: float, tan_fovy: float) -> np.ndarray:
    """Perspective projection with w = z_view and x_ndc = x / (z tan_fovx)."""
    top = tan_fovy * znear
    right = tan_fovx * znear
    P = np.zeros((4, 4))
    P[0, 0] = 2.0 * znear / (2 * right)
    P[1, 1] = 2.0 * znear / (2 * top)
    P[3, 2] = 1.0
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


def look_at_camera(position, target, width: int, height: int, tan_fovx: float, tan_fovy: float,
                   znear: float = 0.01, zfar: float = 100.0) -> Camera:
    c = np.asarray(position, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - c
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0])
    if abs(float(np.dot(up, f))) > 0.999:
        up = np.array([1.0, 0.0, 0.0])
    down = -(up - np.dot(up, f) * f)
    down /= np.linalg.norm(down)
    right = np.cross(down, f)
    R = np.stack([right, down, f])   # rows: camera axes in world coordinates
    W2C = np.eye(4)
    W2C[:3, :3] = R
    W2C[:3, 3] = -R @ c
    P = projection_matrix(znear, zfar, tan_fovx, tan_fovy)
    return Camera(
        world_view_transform=W2C.T.astype(np.float32),
        full_proj_transform=(P @ W2C).T.astype(np.float32),
        camera_center=c.astype(np.float32),
        tan_fovx=float(tan_fovx),
        tan_fovy=float(tan_fovy),
        width=int(width),
        height=int(height),
    )


# ----------------------------------------------------------------- workloads
def small_rgb_frame():
    """Plain forward splat: 10k Gaussians, U[-1,1]^3, degree-0 colour, camera at
    distance 4, fov 60 degrees both axes, 256x256."""
    scene = make_scene(10_000, seed=1, extent=1.0, sh_mode="rgb")
    t = math.tan(math.radians(30.0))
    cam = look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], 256, 256, t, t)
    return scene, cam


def training_frame(n: int = 1_000_000, width: int = 1920, height: int = 1080):
    """Tile splat forward+backward: 1M Gaussians in U[-5,5]^3, SH degree 3,
    camera at (0,0,-6) looking at the origin, fov_y 60 degrees, 1920x1080."""
    scene = make_scene(n, seed=2, extent=5.0, sh_degree=3)
    ty = math.tan(math.radians(30.0))
    cam = look_at_camera([0.0, 0.0, -6.0], [0.0, 0.0, 0.0], width, height, ty * width / height, ty)
    return scene, cam


def marcher_frame(n: int = 200_000, width: int = 1024, height: int = 1024):
    """Volumetric marcher: 200k Gaussians in U[-5,5]^3, SH degree 3, 1024x1024 rays."""
    scene = make_scene(n, seed=3, extent=5.0, sh_degree=3)
    t = math.tan(math.radians(30.0))
    cam = look_at_camera([0.0, 0.0, -6.0], [0.0, 0.0, 0.0], width, height, t * width / height, t)
    return scene, cam


RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]


def sphere_cameras(count: int, seed: int, radius: float = 6.0, resolutions=RESOLUTIONS):
    """Cameras uniform on a sphere looking at the origin, resolution drawn
    uniformly from `resolutions` (fov_y 60 degrees, square pixels)."""
    st = Stream(seed, 7)
    z = st.uniform(count) * 2.0 - 1.0
    phi = st.uniform(count) * 2.0 * math.pi
    pick = (st.uniform(count) * len(resolutions)).astype(np.int64)
    ty = math.tan(math.radians(30.0))
    cams = []
    for i in range(count):
        r = math.sqrt(max(0.0, 1.0 - z[i] * z[i]))
        pos = radius * np.array([r * math.cos(phi[i]), r * math.sin(phi[i]), z[i]])
        w, h = resolutions[int(pick[i])]
        cams.append(look_at_camera(pos, [0.0, 0.0, 0.0], w, h, ty * w / h, ty))
    return cams


def multiview_batch(n: int = 1_000_000, n_cams: int = 64):
    scene = make_scene(n, seed=4, extent=5.0, sh_degree=3)
    return scene, sphere_cameras(n_cams, seed=4)


def pad_sh16(sh: np.ndarray) -> np.ndarray:
    """[n,K,3] -> [n,16,3] with zero padding (python_bindings/vol_marcher.cu:92)."""
    n, k, _ = sh.shape
    out = np.zeros((n, 16, 3), dtype=sh.dtype)
    out[:, :k] = sh
    return out

"""Scene and camera files in the layout INRIA-style Gaussian splatting trainers
write, so resident scenes can be loaded from real captures.

* PLY (binary little endian, one `vertex` element): x y z, nx ny nz,
  f_dc_0..2, f_rest_0..(3(K-1)-1) stored channel-major (all coefficients of
  red, then green, then blue), opacity (logit), scale_0..2 (log), rot_0..3
  (w, x, y, z, unnormalised).  `load_ply` returns those stored values and the
  activated ones the renderers take (sigmoid, exp, q/|q|); SH come back as
  [N, K, 3] (coefficient-major, the layout of python_bindings/vol_marcher.py's
  `shs`).
* cameras.json: a list of {id, img_name, width, height, position[3],
  rotation[3][3] (camera-to-world, columns are the camera axes), fx, fy}.
  `camera_from_json` builds the renderers' Camera (world_view_transform and
  full_proj_transform in the transposed tensor layout,
  python_bindings/vol_marcher.cu:32-40).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass
from typing import List

import numpy as np

from . import scenes


@dataclass
class PlyScene:
    means: np.ndarray          # [N,3]
    sh: np.ndarray             # [N,K,3]
    opacity_logits: np.ndarray # [N,1]
    log_scales: np.ndarray     # [N,3]
    rotations: np.ndarray      # [N,4] stored (unnormalised) w, x, y, z
    sh_degree: int

    @property
    def n(self) -> int:
        return self.means.shape[0]

    def activated(self):
        """(sh, opacity, means, scales, rotations) as the renderers take them."""
        q = self.rotations / np.maximum(np.linalg.norm(self.rotations, axis=1, keepdims=True), 1e-12)
        op = 1.0 / (1.0 + np.exp(-self.opacity_logits))
        return (self.sh.astype(np.float32), op.astype(np.float32), self.means.astype(np.float32),
                np.exp(self.log_scales).astype(np.float32), q.astype(np.float32))


def _fields(k: int) -> List[str]:
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{i}" for i in range(3 * (k - 1))]
    names += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    return names


def save_ply(path: str, scene: PlyScene) -> None:
    n, k, _ = scene.sh.shape
    names = _fields(k)
    data = np.zeros((n, len(names)), np.float32)
    data[:, 0:3] = scene.means
    data[:, 6:9] = scene.sh[:, 0, :]
    rest = scene.sh[:, 1:, :].transpose(0, 2, 1).reshape(n, -1)  # channel-major
    data[:, 9:9 + rest.shape[1]] = rest
    o = 9 + rest.shape[1]
    data[:, o] = scene.opacity_logits.reshape(-1)
    data[:, o + 1:o + 4] = scene.log_scales
    data[:, o + 4:o + 8] = scene.rotations
    header = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    header += "".join(f"property float {nm}\n" for nm in names) + "end_header\n"
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(data.astype("<f4").tobytes())


def load_ply(path: str) -> PlyScene:
    with open(path, "rb") as f:
        header = []
        while True:
            line = f.readline()
            if not line:
                raise ValueError("PLY header not terminated")
            line = line.decode("ascii").strip()
            header.append(line)
            if line == "end_header":
                break
        if "format binary_little_endian 1.0" not in header:
            raise ValueError("only binary little-endian PLY is supported")
        n = None
        props = []
        in_vertex = False
        for line in header:
            parts = line.split()
            if parts[:2] == ["element", "vertex"]:
                n = int(parts[2])
                in_vertex = True
            elif parts and parts[0] == "element":
                in_vertex = False
            elif parts and parts[0] == "property" and in_vertex:
                if parts[1] not in ("float", "float32"):
                    raise ValueError(f"unsupported property type {parts[1]}")
                props.append(parts[2])
        if n is None:
            raise ValueError("no vertex element")
        data = np.frombuffer(f.read(4 * n * len(props)), dtype="<f4").reshape(n, len(props))
    col = {nm: i for i, nm in enumerate(props)}
    get = lambda names: np.stack([data[:, col[nm]] for nm in names], axis=1).astype(np.float32)
    rest_names = sorted((nm for nm in props if nm.startswith("f_rest_")), key=lambda s: int(s.split("_")[-1]))
    k = 1 + len(rest_names) // 3
    sh = np.zeros((n, k, 3), np.float32)
    sh[:, 0, :] = get(["f_dc_0", "f_dc_1", "f_dc_2"])
    if k > 1:
        sh[:, 1:, :] = get(rest_names).reshape(n, 3, k - 1).transpose(0, 2, 1)
    degree = int(round(math.sqrt(k))) - 1
    return PlyScene(get(["x", "y", "z"]), sh, get(["opacity"]), get(["scale_0", "scale_1", "scale_2"]),
                    get(["rot_0", "rot_1", "rot_2", "rot_3"]), degree)


def ply_from_scene(scene: scenes.Scene) -> PlyScene:
    """The stored (pre-activation) form of a synthetic scene."""
    return PlyScene(scene.means.copy(), scene.sh.copy(), scene.opacity_logits.reshape(-1, 1).copy(), scene.log_scales.copy(),
                    scene.rotations.copy(), scene.sh_degree)


def camera_from_json(entry: dict, znear: float = 0.01, zfar: float = 100.0) -> scenes.Camera:
    """One cameras.json entry: rotation is camera-to-world (its columns are the
    camera's x, y, z axes in world space), position the camera centre."""
    R = np.array(entry["rotation"], np.float64)
    C = np.array(entry["position"], np.float64)
    w, h = int(entry["width"]), int(entry["height"])
    tan_x = w / (2.0 * float(entry["fx"]))
    tan_y = h / (2.0 * float(entry["fy"]))
    world_to_cam = np.eye(4)
    world_to_cam[:3, :3] = R.T
    world_to_cam[:3, 3] = -R.T @ C
    proj = scenes.projection_matrix(znear, zfar, tan_x, tan_y)
    return scenes.Camera(world_view_transform=world_to_cam.T.astype(np.float32),
                         full_proj_transform=(proj @ world_to_cam).T.astype(np.float32),
                         camera_center=C.astype(np.float32), tan_fovx=float(tan_x), tan_fovy=float(tan_y), width=w, height=h)


def load_cameras_json(path: str) -> List[scenes.Camera]:
    with open(path) as f:
        return [camera_from_json(e) for e in json.load(f)]

"""Writes inria_scene.ply and inria_cameras.json: a small scene and two cameras
in the file layout INRIA-style Gaussian splatting trainers write, built here
byte by byte (independently of scene_io.save_ply), so loading them checks the
reader against the layout itself.

PLY: binary little endian, one `vertex` element of float properties in the
trainer's order x y z nx ny nz f_dc_0..2 f_rest_0..44 opacity scale_0..2
rot_0..3, with f_rest channel-major (f_rest_{15 c + k - 1} is coefficient k of
colour channel c), opacity a logit, scales logarithms, rotations unnormalised
(w, x, y, z).  cameras.json: entries {id, img_name, width, height, position,
rotation (camera-to-world, rows of the 3x3 matrix), fx, fy}.

Values (seed 1234): 300 Gaussians, means uniform in [-1.2, 1.2]^3, normals
zero, f_dc uniform [-1, 1], f_rest normal(0, 0.15), opacity logits
normal(0, 1.5), log-scales normal(-2.6, 0.4), rotations normal(0, 1)^4.
Run: python tests/golden/make_inria_ply.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N = 300


def columns():
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{i}" for i in range(45)]
    names += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    return names


def values(rng):
    v = {}
    xyz = rng.uniform(-1.2, 1.2, (N, 3))
    v["x"], v["y"], v["z"] = xyz.T
    v["nx"] = v["ny"] = v["nz"] = np.zeros(N)
    dc = rng.uniform(-1.0, 1.0, (N, 3))
    for c in range(3):
        v[f"f_dc_{c}"] = dc[:, c]
    rest = rng.normal(0.0, 0.15, (N, 3, 15))   # [vertex, channel, coefficient 1..15]
    for c in range(3):
        for k in range(15):
            v[f"f_rest_{15 * c + k}"] = rest[:, c, k]
    v["opacity"] = rng.normal(0.0, 1.5, N)
    ls = rng.normal(-2.6, 0.4, (N, 3))
    for i in range(3):
        v[f"scale_{i}"] = ls[:, i]
    q = rng.normal(0.0, 1.0, (N, 4))
    for i in range(4):
        v[f"rot_{i}"] = q[:, i]
    return v


def main():
    rng = np.random.default_rng(1234)
    v = values(rng)
    names = columns()
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {N}"] + [f"property float {nm}" for nm in names] + ["end_header"]
    body = np.stack([np.asarray(v[nm], np.float32) for nm in names], axis=1).astype("<f4")
    with open(os.path.join(HERE, "inria_scene.ply"), "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(body.tobytes())
    # cameras: camera-to-world rotations whose columns are the camera's x (right),
    # y (down) and z (forward) axes, looking at the origin from distance 4
    cams = []
    for i, (pos, w, h, fx, fy) in enumerate([((0.5, -0.8, -4.0), 160, 120, 180.0, 180.0), ((3.2, 0.4, -2.4), 128, 96, 150.0, 140.0)]):
        c = np.array(pos, np.float64)
        z = -c / np.linalg.norm(c)
        up = np.array([0.0, 1.0, 0.0])
        y = -(up - (up @ z) * z)
        y /= np.linalg.norm(y)
        x = np.cross(y, z)
        R = np.stack([x, y, z], axis=1)
        cams.append({"id": i, "img_name": f"view_{i:03d}", "width": w, "height": h, "position": c.tolist(), "rotation": R.tolist(),
                     "fx": fx, "fy": fy})
    with open(os.path.join(HERE, "inria_cameras.json"), "w") as f:
        json.dump(cams, f, indent=1)


if __name__ == "__main__":
    main()

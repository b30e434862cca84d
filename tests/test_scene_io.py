"""Scene/camera files (CPU): PLY round trip, the channel-major f_rest order, the
activations of the stored values, and cameras.json against look_at cameras."""
import math

import numpy as np

from differentiable_gaussian_mixture_renderer_amd import scene_io, scenes


def test_ply_round_trip_and_activation(tmp_path):
    sc = scenes.make_scene(500, seed=3, extent=2.0, sh_degree=3)
    ply = scene_io.ply_from_scene(sc)
    path = str(tmp_path / "scene.ply")
    scene_io.save_ply(path, ply)
    back = scene_io.load_ply(path)
    assert back.n == 500 and back.sh_degree == 3
    for name in ("means", "sh", "opacity_logits", "log_scales", "rotations"):
        np.testing.assert_array_equal(getattr(back, name), getattr(ply, name).astype(np.float32))
    sh, op, mu, s, q = back.activated()
    np.testing.assert_allclose(op, sc.opacity, rtol=1e-6)
    np.testing.assert_allclose(s, sc.scales, rtol=1e-6)
    np.testing.assert_allclose(q, sc.rotations, rtol=1e-6, atol=1e-7)


def test_f_rest_is_channel_major(tmp_path):
    n, k = 2, 4
    sh = np.arange(n * k * 3, dtype=np.float32).reshape(n, k, 3)
    ply = scene_io.PlyScene(np.zeros((n, 3), np.float32), sh, np.zeros((n, 1), np.float32), np.zeros((n, 3), np.float32),
                            np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1)), 1)
    path = str(tmp_path / "order.ply")
    scene_io.save_ply(path, ply)
    raw = open(path, "rb").read()
    body = np.frombuffer(raw[raw.index(b"end_header\n") + len(b"end_header\n"):], dtype="<f4").reshape(n, -1)
    # f_rest_j of vertex i is sh[i, 1 + j % (k-1), j // (k-1)]
    for j in range(3 * (k - 1)):
        assert body[1, 9 + j] == sh[1, 1 + j % (k - 1), j // (k - 1)]
    assert scene_io.load_ply(path).sh_degree == 1


def test_cameras_json_matches_look_at():
    w, h = 320, 200
    ty = math.tan(math.radians(25.0))
    tx = ty * w / h
    ref = scenes.look_at_camera([1.0, -2.0, -5.0], [0.2, 0.1, 0.0], w, h, tx, ty)
    V = ref.world_view_transform.T.astype(np.float64)  # world-to-camera
    entry = {"id": 0, "img_name": "a", "width": w, "height": h, "position": ref.camera_center.tolist(),
             "rotation": V[:3, :3].T.tolist(), "fx": w / (2 * tx), "fy": h / (2 * ty)}
    cam = scene_io.camera_from_json(entry)
    np.testing.assert_allclose(cam.world_view_transform, ref.world_view_transform, atol=2e-6)
    np.testing.assert_allclose(cam.full_proj_transform, ref.full_proj_transform, atol=2e-6)
    assert abs(cam.tan_fovx - tx) < 1e-12 and abs(cam.tan_fovy - ty) < 1e-12


def _fixture_values():
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "make_inria_ply.py")
    spec = importlib.util.spec_from_file_location("make_inria_ply", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.values(np.random.default_rng(1234)), os.path.dirname(path)


def test_inria_fixture_loads_in_the_trainer_layout():
    """tests/golden/inria_scene.ply is written byte by byte in the trainer's
    property order (make_inria_ply.py, not save_ply): every loaded field
    equals the recipe's value, f_rest read channel-major."""
    import os
    v, d = _fixture_values()
    ply = scene_io.load_ply(os.path.join(d, "inria_scene.ply"))
    assert ply.n == 300 and ply.sh_degree == 3 and ply.sh.shape == (300, 16, 3)
    f = lambda a: np.asarray(a, np.float32)
    np.testing.assert_array_equal(ply.means, np.stack([f(v["x"]), f(v["y"]), f(v["z"])], 1))
    for c in range(3):
        np.testing.assert_array_equal(ply.sh[:, 0, c], f(v[f"f_dc_{c}"]))
        for k in range(1, 16):
            np.testing.assert_array_equal(ply.sh[:, k, c], f(v[f"f_rest_{15 * c + k - 1}"]))
    np.testing.assert_array_equal(ply.opacity_logits[:, 0], f(v["opacity"]))
    np.testing.assert_array_equal(ply.log_scales, np.stack([f(v[f"scale_{i}"]) for i in range(3)], 1))
    np.testing.assert_array_equal(ply.rotations, np.stack([f(v[f"rot_{i}"]) for i in range(4)], 1))
    sh, op, mu, s, q = ply.activated()
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, rtol=1e-6)
    np.testing.assert_allclose(op[:, 0], 1.0 / (1.0 + np.exp(-f(v["opacity"]).astype(np.float64))), rtol=1e-6)


def test_inria_cameras_fixture():
    """cameras.json entries look at the origin: the world-to-camera transform
    puts the origin on the +z axis at the camera's distance, and the
    projection maps it to the image centre (NDC 0, 0)."""
    import os
    _, d = _fixture_values()
    cams = scene_io.load_cameras_json(os.path.join(d, "inria_cameras.json"))
    assert len(cams) == 2
    for cam in cams:
        V = cam.world_view_transform.T.astype(np.float64)
        P = cam.full_proj_transform.T.astype(np.float64)
        o = V @ np.array([0.0, 0.0, 0.0, 1.0])
        np.testing.assert_allclose(o[:2], 0.0, atol=1e-5)
        assert abs(o[2] - np.linalg.norm(cam.camera_center)) < 1e-5
        clip = P @ np.array([0.0, 0.0, 0.0, 1.0])
        np.testing.assert_allclose(clip[:2] / clip[3], 0.0, atol=1e-6)
        assert cam.width in (160, 128)

"""Host logic of the batched form's views (no GPU)."""
import numpy as np

import yar_amd
from yar_amd import views as yv


def test_batch_views_shapes_and_poses(c1_path):
    scene = yar_amd.Scene(c1_path)
    cam = scene.desc.camera
    base = yar_amd.View()
    base.camera_to_world[:] = cam.camera_to_world[:]
    base.fov_y_degrees = cam.fov_y_degrees
    base.z_is_up = cam.z_is_up
    vs = yv.batch_views(base, 32)
    assert len(vs) == 32
    assert {v.film_width for v in vs} <= {128, 256, 512, 768, 1024}
    assert len({v.film_width for v in vs}) >= 4
    assert {v.x_pixel_sample_count * v.y_pixel_sample_count for v in vs} == {4, 16, 64}
    again = yv.batch_views(base, 32)
    assert all(list(a.camera_to_world) == list(b.camera_to_world) and a.film_width == b.film_width
               for a, b in zip(vs, again))
    assert len({v.rng_seed_offset for v in vs}) == 32
    c = np.array([0.5, 0.5, 0.5])
    m0 = np.array(base.camera_to_world[:]).reshape(3, 4)
    for v in vs:
        m = np.array(v.camera_to_world[:]).reshape(3, 4)
        R = m[:, :3]
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-5)           # still a rigid camera frame
        assert np.isclose(np.linalg.norm(m[:, 3] - c), np.linalg.norm(m0[:, 3] - c), rtol=1e-5)
        assert (v.region_x1, v.region_y1) == (v.film_width, v.film_height)


def test_orbit_identity_and_composition():
    v = yar_amd.View()
    v.camera_to_world[:] = [1, 0, 0, 0.5, 0, 1, 0, 0.5, 0, 0, 1, 3.2]
    same = yv.orbit_view(v, (0.5, 0.5, 0.5), 0.0, 0.0)
    assert list(same.camera_to_world) == list(v.camera_to_world)
    a = yv.orbit_view(yv.orbit_view(v, (0.5, 0.5, 0.5), 10.0, 0.0), (0.5, 0.5, 0.5), 15.0, 0.0)
    b = yv.orbit_view(v, (0.5, 0.5, 0.5), 25.0, 0.0)
    assert np.allclose(a.camera_to_world[:], b.camera_to_world[:], atol=1e-5)
    m = np.array(b.camera_to_world[:]).reshape(3, 4)
    # yaw about +y moves the camera on the horizontal circle through the start position
    assert np.isclose(m[1, 3], 0.5) and np.isclose(np.hypot(m[0, 3] - 0.5, m[2, 3] - 0.5), 2.7, rtol=1e-5)

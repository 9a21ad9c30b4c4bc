"""Camera views for the batched multi-view form (yar_gpu_set_views / yar_gpu_render_batch).

A view is the Scene_Overrides analog (kennyalive/YAR src/ref/reference_renderer.h:23): camera pose,
film size, render region, sample grid and RNG seed offset.  `orbit_view` turns a scene's camera
about a point (yaw about the world up axis, then pitch about the camera's right axis), which keeps
the camera conventions of the loaded scene intact; `batch_views` builds the 32-view batch of the
batched workload: seeded poses orbiting the box room of the config-1 scene, square films with sides
drawn uniformly from {128, 256, 512, 768, 1024} and 2x2 / 4x4 / 8x8 stratified sampling (4, 16, 64
spp) drawn uniformly.
"""
import math

import numpy as np

from ._abi import View

BATCH_SIDES = (128, 256, 512, 768, 1024)
BATCH_GRIDS = (2, 4, 8)     # x and y pixel sample counts -> 4, 16, 64 spp


def copy_view(v):
    out = View()
    for name, _ in View._fields_:
        val = getattr(v, name)
        if hasattr(val, "__len__"):
            getattr(out, name)[:] = val[:]
        else:
            setattr(out, name, val)
    return out


def _rot(axis, ang):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(ang), math.sin(ang)
    C = 1.0 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def orbit_view(v, center, yaw_deg, pitch_deg):
    """Copy of view `v` with its camera turned about `center` (degrees)."""
    m = np.array(v.camera_to_world[:], dtype=np.float64).reshape(3, 4)
    R, t = m[:, :3], m[:, 3]
    up = np.array([0.0, 0.0, 1.0]) if v.z_is_up else np.array([0.0, 1.0, 0.0])
    right = R[:, 0]
    rot = _rot(up, math.radians(yaw_deg)) @ _rot(right, math.radians(pitch_deg))
    c = np.asarray(center, dtype=np.float64)
    R2 = rot @ R
    t2 = c + rot @ (t - c)
    out = copy_view(v)
    out.camera_to_world[:] = [float(np.float32(x)) for x in np.concatenate([R2, t2[:, None]], axis=1).reshape(-1)]
    return out


def with_film(v, width, height, grid, seed_offset=None, region=None):
    out = copy_view(v)
    out.film_width, out.film_height = width, height
    if region is None:
        region = (0, 0, width, height)
    out.region_x0, out.region_y0, out.region_x1, out.region_y1 = region
    gx, gy = (grid, grid) if isinstance(grid, int) else grid
    out.x_pixel_sample_count, out.y_pixel_sample_count = gx, gy
    if seed_offset is not None:
        out.rng_seed_offset = seed_offset
    return out


def batch_views(base, count=32, center=(0.5, 0.5, 0.5), seed=1, sides=BATCH_SIDES, grids=BATCH_GRIDS):
    """The batched workload's views: `count` seeded poses (yaw uniform in [-25, 25] degrees, pitch
    uniform in [-8, 8] degrees around `center`), film side and sample grid drawn uniformly from
    `sides` and `grids`, seed offset = view index."""
    rng = np.random.Generator(np.random.PCG64(seed))
    views = []
    for i in range(count):
        yaw = float(rng.uniform(-25.0, 25.0))
        pitch = float(rng.uniform(-8.0, 8.0))
        side = int(sides[rng.integers(len(sides))])
        grid = int(grids[rng.integers(len(grids))])
        v = orbit_view(base, center, yaw, pitch)
        views.append(with_film(v, side, side, grid, seed_offset=i))
    return views


def view_samples(v):
    return (v.region_x1 - v.region_x0) * (v.region_y1 - v.region_y0) * v.x_pixel_sample_count * v.y_pixel_sample_count

"""Seeded ray batches for traversal parity (test helper)."""
import numpy as np

from yar_amd import ray_dtype


def random_rays(scene, n, seed=1, tmax=np.inf):
    """Rays over the scene: interior origins, outside origins aimed inside, camera rays and
    axis-aligned directions (the mix of src/ref/test_kdtree.cpp:46-83)."""
    d = scene.desc
    lo = np.full(3, np.inf, np.float32)
    hi = np.full(3, -np.inf, np.float32)
    for m in range(d.mesh_count):
        md = d.meshes[m]
        v = np.ctypeslib.as_array(md.vertices, shape=(md.vertex_count * 3,)).reshape(-1, 3)
        lo = np.minimum(lo, v.min(0))
        hi = np.maximum(hi, v.max(0))
    # huge ground planes: focus on the region where most geometry is
    span = hi - lo
    big = span > 100
    lo = np.where(big, -20, lo).astype(np.float32)
    hi = np.where(big, 20, hi).astype(np.float32)
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=ray_dtype())
    k = n // 4
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    # outside origins aimed at interior points
    c = (lo + hi) / 2
    rad = np.linalg.norm(hi - lo) * 0.75
    u = rng.normal(size=(k, 3)).astype(np.float32)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o[k:2 * k] = c + u * rad
    tgt = rng.uniform(lo, hi, size=(k, 3)).astype(np.float32)
    dd = tgt - o[k:2 * k]
    dirs[k:2 * k] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    # camera rays through random film points
    cam = np.array(d.camera.camera_to_world[:], np.float32).reshape(3, 4)
    w, h = d.config.film_width, d.config.film_height
    fx = rng.uniform(0, w, k).astype(np.float32)
    fy = rng.uniform(0, h, k).astype(np.float32)
    th = np.float32(np.tan(np.radians(d.camera.fov_y_degrees / 2)))
    cd = np.stack([(2 * fx / w - 1) * th * (w / h), -(2 * fy / h - 1) * th, -np.ones(k, np.float32)], 1)
    cd /= np.linalg.norm(cd, axis=1, keepdims=True)
    o[2 * k:3 * k] = cam[:, 3]
    dirs[2 * k:3 * k] = cd @ cam[:, :3].T
    # 1/32 of the rays axis-aligned
    na = n // 32
    ax = rng.integers(0, 3, na)
    sg = rng.choice([-1.0, 1.0], na).astype(np.float32)
    aligned = np.zeros((na, 3), np.float32)
    aligned[np.arange(na), ax] = sg
    dirs[n - na:] = aligned
    r["origin"] = o
    r["direction"] = dirs
    r["tmax"] = tmax
    return r


def rays_at_mesh(vertices, indices, n, seed=1, axis_aligned_fraction=1 / 32):
    """Rays from a box around a mesh towards random points on its triangles (grazing and
    axis-aligned directions included), as the reference's Ray_Generator does for mesh trees."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3)
    idx = np.asarray(indices, np.int32).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    lo, hi = v.min(0), v.max(0)
    pad = np.maximum((hi - lo) * 0.5, 1e-3)
    o = rng.uniform(lo - pad, hi + pad, size=(n, 3)).astype(np.float32)
    t = rng.integers(0, idx.shape[0], n)
    a, b = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    p = v[idx[t, 0]] + a[:, None] * (v[idx[t, 1]] - v[idx[t, 0]]) + b[:, None] * (v[idx[t, 2]] - v[idx[t, 0]])
    d = (p - o).astype(np.float32)
    na = int(n * axis_aligned_fraction)
    if na:
        ax = np.argmax(np.abs(d[:na]), axis=1)
        keep = np.zeros((na, 3), np.float32)
        keep[np.arange(na), ax] = np.sign(d[np.arange(na), ax])
        o[:na] = p[:na] - keep * np.linalg.norm(d[:na], axis=1, keepdims=True)
        d[:na] = keep
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, dtype=ray_dtype())
    r["origin"] = o
    r["direction"] = d
    r["tmax"] = np.inf
    return r

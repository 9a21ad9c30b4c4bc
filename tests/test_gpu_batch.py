"""Batched multi-view form: a view rendered inside a batch equals the same view rendered alone.

Per-sample work depends only on (view, pixel, sample), and film values are gathered per pixel in
render_tile's visitation order, so chunk composition, batch membership, chunk size and hipGraph
replay must not change a single bit of any view's image.
"""
import numpy as np
import pytest

import yar_amd
from yar_amd import views as yv

pytestmark = pytest.mark.gpu


def _c1_views(ctx):
    base = ctx.view(0)
    c = (0.5, 0.5, 0.5)
    return [
        yv.with_film(yv.orbit_view(base, c, -20.0, 5.0), 64, 64, 2, seed_offset=11),
        yv.with_film(yv.orbit_view(base, c, 10.0, -4.0), 96, 80, (2, 3), seed_offset=12),
        yv.with_film(yv.orbit_view(base, c, 0.0, 0.0), 160, 160, 1, seed_offset=13, region=(30, 20, 150, 100)),
        yv.with_film(yv.orbit_view(base, c, 25.0, 8.0), 128, 128, 4, seed_offset=14),
        yv.with_film(base, 200, 72, (3, 1), seed_offset=15),
    ]


def _alone(scene, views, **opts):
    out = []
    for v in views:
        ctx = yar_amd.Context(scene, **opts)
        ctx.set_views([v])
        img, _, var, _ = ctx.render(0, want_film=False)
        out.append((img, var))
        ctx.close()
    return out


@pytest.mark.parametrize("opts", [{}, {"max_paths_in_flight": 9000}, {"use_graph": 1}, {"bin_rays": 1},
                                  {"bin_rays": 3, "use_graph": 1, "max_paths_in_flight": 20000},
                                  {"lane_refill": 1}],
                         ids=["default", "small_chunks", "graph", "binned", "binned_graph", "refill"])
def test_batch_equals_single_view_renders(c1_path, opts):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene, **opts)
    views = _c1_views(ctx)
    ctx.set_views(views)
    imgs, var, st = ctx.render_batch(list(range(len(views))))
    ref = _alone(scene, views)
    for i, ((rimg, rvar), img, v) in enumerate(zip(ref, imgs, var)):
        diff = np.abs(img - rimg).max()
        print(f"view {i}: shape {img.shape}, max |batch - alone| {diff:.3g}, var {v:.6g} / {rvar:.6g}")
        assert img.shape == rimg.shape
        assert np.array_equal(img, rimg)
        assert v == rvar
    assert st.samples == sum(yv.view_samples(v) for v in views)


def test_batch_order_and_repeats(c1_path):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene)
    views = _c1_views(ctx)[:3]
    ctx.set_views(views)
    a, _, _ = ctx.render_batch([0, 1, 2])
    b, _, _ = ctx.render_batch([2, 0, 1, 0])
    assert np.array_equal(a[0], b[1]) and np.array_equal(a[0], b[3])
    assert np.array_equal(a[1], b[2]) and np.array_equal(a[2], b[0])


def test_render_tiles_subset_matches_full(c1_path):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene)
    v = yv.with_film(ctx.view(0), 200, 136, 2, seed_offset=3)
    ctx.set_views([v])
    n = ctx.tile_count(0)
    full_tiles, full_var, _ = ctx.render_tiles(list(range(n)), 0)
    pick = [n - 1, 1, 3]
    tiles, var, _ = ctx.render_tiles(pick, 0)
    for k, t in enumerate(pick):
        assert np.array_equal(tiles[k]["color_sum"], full_tiles[t]["color_sum"])
        assert np.array_equal(tiles[k]["weight_sum"], full_tiles[t]["weight_sum"])
        assert var[k] == full_var[t]


def test_graph_replay_and_launch_options_equal_eager_c2(c2_small_path):
    """hipGraph replay, chunk size, per-lane refill, stack storage and traversal block size (three- and
    two-wave blocks on the per-kernel occupancy grids) change no bit of the image."""
    scene = yar_amd.Scene(c2_small_path)
    region = (432, 432, 592, 592)
    res = []
    for opts in ({}, {"use_graph": 1}, {"use_graph": 1, "max_paths_in_flight": 20000}, {"lane_refill": 1},
                 {"lane_refill": 1, "stack_mode": 2, "block_size": 256}, {"block_size": 192}, {"block_size": 128}):
        ctx = yar_amd.Context(scene, **opts)
        v = ctx.view(0)
        v.x_pixel_sample_count = v.y_pixel_sample_count = 2
        v.region_x0, v.region_y0, v.region_x1, v.region_y1 = region
        ctx.set_views([v])
        img, _, var, _ = ctx.render(0, want_film=False)
        img2, _, var2, _ = ctx.render(0, want_film=False)     # second call replays the captured loop
        assert np.array_equal(img, img2) and var == var2
        res.append(img)
        ctx.close()
    for r in res[1:]:
        assert np.array_equal(res[0], r)


def test_sequential_sampling_matches_oracle(scene_dir, oracle):
    """max depth > 10: the sampler's tables run out and samples draw from the per-pixel fallback
    stream in sample order, so the engine renders one sample per pixel per pass."""
    path = yar_amd.generate_scene(1, str(scene_dir / "c1_depth12.pbrt"), seed=1)
    text = open(path).read()
    assert '"integer maxdepth" [ 5 ]' in text
    open(path, "w").write(text.replace('"integer maxdepth" [ 5 ]', '"integer maxdepth" [ 12 ]'))
    scene = yar_amd.Scene(path)
    assert scene.desc.config.max_light_bounces == 12
    ctx = yar_amd.Context(scene)
    v = ctx.view(0)
    v.x_pixel_sample_count = v.y_pixel_sample_count = 2
    v.region_x0, v.region_y0, v.region_x1, v.region_y1 = 96, 96, 160, 160
    ctx.set_views([v])
    img, film, var, st = ctx.render(0)
    oimg, ofilm, ovar, _, _ = oracle.OracleScene(scene).render(v)
    rel = np.abs(img - oimg) / np.maximum(np.abs(oimg), 1e-3)
    close = np.all(rel <= 1e-4, axis=-1).mean()
    mean_rel = abs(img.mean() - oimg.mean()) / max(oimg.mean(), 1e-12)
    print(f"depth 12: pixels within 1e-4 {close:.5f}, mean rel {mean_rel:.2e}")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()

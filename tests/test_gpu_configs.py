"""GPU parity at the benchmark configurations' real sizes (C ABI against the CPU oracle).

config 2: the full 2.18M-triangle scene (400 displaced icospheres, 2 km ground): 256K seeded rays and
a 64x64 crop at 64 spp.  config 3: the full 32-view batch (hipGraph bounce loop, the autotuner's
launch shape) equals the 32 single-view renders bit for bit.  config 4: the 500K-triangle room lit
through a slit: rays, a crop, and the guided render's mean against the plain one.  config 5: the
19.4M-triangle instanced scene at full size: rays and a crop at 256 spp, depth 16 (sequential sample
passes with the RNG fallback past ten bounces)."""
import os
import tempfile

import numpy as np
import pytest

import yar_amd
from yar_amd import views as yv
from tests import raygen

pytestmark = pytest.mark.gpu

_CACHE = os.path.join(tempfile.gettempdir(), "yar_scenes")


def _scene(config):
    os.makedirs(_CACHE, exist_ok=True)
    path = os.path.join(_CACHE, f"config{config}_s1_x1.0.pbrt")
    if not os.path.exists(path):
        yar_amd.generate_scene(config, path + ".tmp", seed=1)
        os.replace(path + ".tmp", path)
    return yar_amd.Scene(path)


@pytest.fixture(scope="module")
def c2():
    return _scene(2)


@pytest.fixture(scope="module")
def c4():
    return _scene(4)


def _hits_agree(g, o):
    same = (g["object"] == o["object"]) & (g["triangle"] == o["triangle"])
    both = (o["object"] >= 0) & same
    rel = np.abs(g["t"][both] - o["t"][both]) / np.maximum(np.abs(o["t"][both]), 1e-30)
    return same.mean(), (rel.max() if rel.size else 0.0)


def _check_rays(scene, osc, label, seed):
    rays = raygen.random_rays(scene, 1 << 18, seed=seed)
    ctx = yar_amd.Context(scene)
    g = ctx.trace_rays(rays)
    o = osc.trace(rays)
    same, max_rel = _hits_agree(g, o)
    rays["tmax"] = np.random.default_rng(seed).uniform(0.01, 20.0, rays.size).astype(np.float32)
    ga = ctx.trace_rays(rays, any_hit=True)
    oa = osc.trace(rays, any_hit=True)
    any_agree = (ga["t"] == oa["t"]).mean()
    print(f"{label}: hit ids {same:.6f}, max rel t {max_rel:.3g}, hit rate {(o['object'] >= 0).mean():.3f}, "
          f"any-hit {any_agree:.6f}", flush=True)
    ctx.close()
    assert same >= 0.9999 and max_rel <= 1e-5
    assert any_agree >= 0.9999


def _crop(view, side, nx=None):
    v = yv.copy_view(view)
    cx, cy = (v.region_x0 + v.region_x1) // 2, (v.region_y0 + v.region_y1) // 2
    v.region_x0, v.region_y0, v.region_x1, v.region_y1 = cx - side // 2, cy - side // 2, cx + side // 2, cy + side // 2
    if nx:
        v.x_pixel_sample_count = v.y_pixel_sample_count = nx
    return v


def _check_crop(scene, osc, view, label, close_min=0.99):
    ctx = yar_amd.Context(scene)
    ctx.set_views([view])
    img, film, var, st = ctx.render(0)
    ctx.close()
    oimg, ofilm, ovar, _, _ = osc.render(view)
    rel = np.abs(img - oimg) / np.maximum(np.abs(oimg), 1e-3)
    close = np.all(rel <= 1e-4, axis=-1).mean()
    exact = (np.all(film["color_sum"] == ofilm["color_sum"], axis=-1) & (film["weight_sum"] == ofilm["weight_sum"])).mean()
    mean_rel = abs(img.mean() - oimg.mean()) / max(oimg.mean(), 1e-12)
    var_rel = abs(var - ovar) / max(ovar, 1e-12)
    print(f"{label}: {st.samples} samples, bit-exact pixels {exact:.4f}, within 1e-4 rel {close:.4f}, "
          f"mean {img.mean():.6g} / {oimg.mean():.6g}, variance {var:.6g} / {ovar:.6g}", flush=True)
    assert np.isfinite(img).all()
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()        # footprints and sample counts exact
    assert close >= close_min
    assert mean_rel <= 2e-3
    assert var_rel <= 2e-2


def test_c2_full_scene_rays(c2, oracle):
    _check_rays(c2, oracle.OracleScene(c2), "c2 full (2.18M tris)", seed=31)


def test_c2_full_scene_crop_64spp(c2, oracle):
    view = _crop(yar_amd.Context(c2).view(0), 64)
    _check_crop(c2, oracle.OracleScene(c2), view, "c2 full, 64x64 crop, 64 spp")


def test_c3_full_batch_equals_single_views(c1_path):
    scene = yar_amd.Scene(c1_path)
    probe = yar_amd.Context(scene)
    views = yv.batch_views(probe.view(0), 32)
    probe.close()
    opts = dict(use_graph=1, block_size=128, stack_mode=1, traversal_stack_depth=8, lane_refill=1,
                max_paths_in_flight=1 << 23)
    ctx = yar_amd.Context(scene, **opts)
    ctx.set_views(views)
    imgs, var, st = ctx.render_batch(list(range(len(views))))
    assert st.samples == sum(yv.view_samples(v) for v in views)
    single = yar_amd.Context(scene)
    for i, v in enumerate(views):
        single.set_views([v])
        img, _, rvar, _ = single.render(0, want_film=False)
        assert np.array_equal(img, imgs[i]), i
        assert rvar == var[i]
    print(f"c3: 32 views, {st.samples} samples, batch == single-view renders", flush=True)


def test_c4_full_scene_rays_and_crop(c4, oracle):
    osc = oracle.OracleScene(c4)
    _check_rays(c4, osc, "c4 full (500K tris)", seed=41)
    view = _crop(yar_amd.Context(c4).view(0), 48)
    _check_crop(c4, osc, view, "c4 full, 48x48 crop, 64 spp")


def test_c4_guided_mean_matches_plain(c4):
    ctx = yar_amd.Context(c4)
    view = _crop(ctx.view(0), 192, nx=4)
    ctx.set_views([view])
    plain, pvar, _ = ctx.render_batch([0])
    guide = yar_amd.Guide(seed=3)
    guided, gvar, gst = ctx.render_guided(guide, [0])
    n = plain[0].shape[0] * plain[0].shape[1]
    lum = lambda im: (0.2126729 * im[..., 0] + 0.7151522 * im[..., 1] + 0.0721750 * im[..., 2]).mean()
    sigma = np.sqrt(pvar[0] / n + gvar[0] / n)
    print(f"c4 guided: plain {lum(plain[0]):.6g}, guided {lum(guided[0]):.6g}, sigma {sigma:.3g}, "
          f"passes {gst.passes}, train nats {gst.train_nats:.3f}", flush=True)
    assert abs(lum(plain[0]) - lum(guided[0])) <= 5 * sigma


def test_c5_full_scene_rays_and_crop(oracle):
    c5 = _scene(5)
    assert c5.triangle_count() > 19_000_000
    osc = oracle.OracleScene(c5)
    _check_rays(c5, osc, "c5 full (19.4M tris)", seed=51)
    view = _crop(yar_amd.Context(c5).view(0), 16)
    # depth-16 paths over low-roughness conductors and glass carry a device sin/cos ulp difference through
    # many more discrete events than the shallower configs: fewer pixels stay within 1e-4 (measured 96.9%),
    # while the mean and the variance still agree to 1e-4 and 3e-3
    _check_crop(c5, osc, view, "c5 full, 16x16 crop, 256 spp, depth 16", close_min=0.95)

"""Guided form on the GPU: the learned proposal trained online inside yar_gpu_render_guided keeps the
estimator unbiased (same mean as the plain tracer within the noise bound) and lowers the error at
equal sample count where indirect light dominates."""
import numpy as np
import pytest

import yar_amd
from yar_amd import views as yv
from tests import scenes

pytestmark = pytest.mark.gpu


def _lum(img):
    return 0.2126729 * img[..., 0] + 0.7151522 * img[..., 1] + 0.0721750 * img[..., 2]


def test_guided_mean_matches_plain_c1(c1_path):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene)
    base = ctx.view(0)
    region = (64, 64, 192, 192)
    ref_v = yv.with_film(base, 256, 256, 16, seed_offset=100, region=region)
    g_v = yv.with_film(base, 256, 256, 4, seed_offset=7, region=region)
    ctx.set_views([ref_v, g_v])
    ref, _, ref_var, _ = ctx.render(0, want_film=False)
    guide = yar_amd.Guide(seed=2, first_pass_spp=4)
    imgs, var, st = ctx.render_guided(guide, [1])
    img = imgs[0]
    n = img.shape[0] * img.shape[1]
    diff = _lum(img).mean() - _lum(ref).mean()
    sigma = np.sqrt(var[0] / n + ref_var / n)
    print(f"guided mean {_lum(img).mean():.6f} plain {_lum(ref).mean():.6f} diff {diff:.3g} sigma {sigma:.3g} "
          f"passes {st.passes} records {st.train_records} nats {st.train_nats:.3f}")
    assert np.isfinite(img).all()
    assert st.passes == 2 and st.train_records > 0          # 16 spp in passes of 4 (BSDF only) and 12 (guided)
    assert abs(diff) < 5 * sigma


def test_guided_lowers_error_where_indirect_light_dominates():
    sc = scenes.indirect_room_scene(res=96, spp_side=8)
    ctx = yar_amd.Context(sc.desc)
    base = ctx.view(0)
    ctx.set_views([yv.with_film(base, 96, 96, 64, seed_offset=1000), yv.with_film(base, 96, 96, 8, seed_offset=1)])
    ref, _, _, _ = ctx.render(0, want_film=False)
    plain, _, _, _ = ctx.render(1, want_film=False)
    # four passes (8, 8, 16, 32 spp) so the mixture is retrained as the samples accumulate
    guide = yar_amd.Guide(seed=1, pass_count=4)
    imgs, _, st = ctx.render_guided(guide, [1])
    mse_p = float(((plain - ref) ** 2).mean())
    mse_g = float(((imgs[0] - ref) ** 2).mean())
    print(f"indirect room 64 spp: MSE plain {mse_p:.4g}, guided {mse_g:.4g} (ratio {mse_g / mse_p:.3f}), "
          f"mean plain {_lum(plain).mean():.5f} guided {_lum(imgs[0]).mean():.5f} ref {_lum(ref).mean():.5f}, "
          f"nats {st.train_nats:.3f}")
    assert st.train_nats > 0.5          # the trained mixture beats the uniform density on the tracer's samples
    assert mse_g < 0.92 * mse_p


def test_guided_batch_of_views(c1_path):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene, max_paths_in_flight=20000)
    base = ctx.view(0)
    c = (0.5, 0.5, 0.5)
    views = [yv.with_film(yv.orbit_view(base, c, -15, 3), 96, 64, 4, seed_offset=1),
             yv.with_film(yv.orbit_view(base, c, 10, -2), 64, 64, 2, seed_offset=2)]
    ctx.set_views(views)
    guide = yar_amd.Guide(seed=3)
    imgs, var, st = ctx.render_guided(guide, [0, 1])
    plain, pvar, _ = ctx.render_batch([0, 1])
    for i in range(2):
        n = imgs[i].shape[0] * imgs[i].shape[1]
        diff = _lum(imgs[i]).mean() - _lum(plain[i]).mean()
        sigma = np.sqrt(var[i] / n + pvar[i] / n)
        print(f"view {i}: guided {_lum(imgs[i]).mean():.5f} plain {_lum(plain[i]).mean():.5f} sigma {sigma:.3g}")
        assert imgs[i].shape == plain[i].shape and np.isfinite(imgs[i]).all()
        assert abs(diff) < 5 * sigma
    assert st.samples == sum(yv.view_samples(v) for v in views)

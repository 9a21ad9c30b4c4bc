"""Checkpointed rendering on the GPU (render_scene with a checkpoint directory,
src/ref/reference_renderer.cpp:446-584): an uninterrupted checkpointed render equals the plain render
bit for bit, its tile files hold exactly the Film_Tiles render_tiles returns, and a session resumed
after losing some tile files renders only those tiles and ends with the same image."""
import os

import numpy as np
import pytest

import yar_amd

pytestmark = pytest.mark.gpu


def test_checkpoint_resume_equals_uninterrupted_render(c1_path, tmp_path):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene)
    v = ctx.view(0)
    v.x_pixel_sample_count = v.y_pixel_sample_count = 2
    ctx.set_views([v])
    ref_img, ref_film, ref_var, _ = ctx.render(0)
    n = ctx.tile_count(0)
    d = str(tmp_path / "ck")
    img, film, var, t, st = ctx.render_checkpointed(d, "c1.pbrt", tiles_per_batch=5)
    assert np.array_equal(img, ref_img) and film.tobytes() == ref_film.tobytes() and var == ref_var
    assert st.samples == (v.region_x1 - v.region_x0) * (v.region_y1 - v.region_y0) * 4 and t > 0
    names = sorted(os.listdir(d))
    assert names == ["checkpoint"] + [f"tile_{i:04d}" for i in range(n)]
    tiles, tvar, _ = ctx.render_tiles(list(range(n)))
    ck = yar_amd.Checkpoint(d, "c1.pbrt", n, 4)
    for i in range(n):
        b, px, tv = ck.tile(i)
        assert b == ctx.tile_bounds(i)[1] and px.tobytes() == tiles[i].tobytes() and tv == tvar[i]
    ck.close()
    # an interrupted session: some tiles never reached the directory
    lost = [0, 3, 7, n - 1]
    for i in lost:
        os.remove(os.path.join(d, f"tile_{i:04d}"))
    img2, film2, var2, t2, st2 = ctx.render_checkpointed(d, "c1.pbrt", tiles_per_batch=3)
    spp_px = sum((ctx.tile_bounds(i)[0][2] - ctx.tile_bounds(i)[0][0]) * (ctx.tile_bounds(i)[0][3] - ctx.tile_bounds(i)[0][1])
                 for i in lost)
    assert st2.samples == 4 * spp_px                                # only the lost tiles were rendered
    assert np.array_equal(img2, ref_img) and film2.tobytes() == ref_film.tobytes() and var2 == ref_var
    assert t2 > 0
    # a finished checkpoint renders nothing
    img3, _, _, _, st3 = ctx.render_checkpointed(d, "c1.pbrt")
    assert st3.samples == 0 and np.array_equal(img3, ref_img)

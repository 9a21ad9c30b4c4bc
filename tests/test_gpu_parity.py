"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs (C ABI calls)."""
import numpy as np
import pytest

import yar_amd
from tests import raygen

pytestmark = pytest.mark.gpu


def _hits_agree(g, o):
    same = (g["object"] == o["object"]) & (g["triangle"] == o["triangle"])
    both = (o["object"] >= 0) & same
    rel = np.abs(g["t"][both] - o["t"][both]) / np.maximum(np.abs(o["t"][both]), 1e-30)
    return same.mean(), (rel.max() if rel.size else 0.0), (g["t"][both] == o["t"][both]).mean() if both.any() else 1.0


@pytest.mark.parametrize("which,opts", [("c1", {}), ("c2_small", {}),
                                        ("c2_small", {"traversal_stack_depth": 2}),
                                        ("c2_small", {"traversal_stack_depth": 32, "stack_mode": 2}),
                                        ("c2_small", {"kdtree_builder": 1}),
                                        ("c2_small", {"traversal_stack_depth": 2, "lane_refill": 1})],
                         ids=["c1", "c2_small", "c2_small-restart", "c2_small-private", "c2_small-binned-trees",
                              "c2_small-restart-refill"])
def test_trace_rays_closest_matches_oracle(which, opts, c1_path, c2_small_path, oracle):
    """Short stacks restart from the root (kd-restart) when they overflow; the hit records must not
    depend on the stack storage or depth, on per-lane refill, nor on
    the tree (binned SAH fast path against the oracle's reference-builder trees)."""
    path = c1_path if which == "c1" else c2_small_path
    scene = yar_amd.Scene(path)
    rays = raygen.random_rays(scene, 1 << 18, seed=7)
    ctx = yar_amd.Context(scene, **opts)
    g = ctx.trace_rays(rays)
    o = oracle.OracleScene(scene).trace(rays)
    same, max_rel, exact = _hits_agree(g, o)
    print(f"{which}: hit-id agreement {same:.6f}, max rel t err {max_rel:.3g}, bit-exact t {exact:.6f}, "
          f"hit rate {(o['object'] >= 0).mean():.3f}")
    assert same >= 0.9999            # hit (object, triangle) identical on >= 99.99% of the batch
    assert max_rel <= 1e-5           # t within 1e-5 relative where the hit agrees


def test_trace_rays_any_matches_oracle(c1_path, oracle):
    scene = yar_amd.Scene(c1_path)
    rays = raygen.random_rays(scene, 1 << 18, seed=9)
    rng = np.random.default_rng(3)
    rays["tmax"] = rng.uniform(0.01, 3.0, rays.size).astype(np.float32)
    ctx = yar_amd.Context(scene)
    g = ctx.trace_rays(rays, any_hit=True)
    o = oracle.OracleScene(scene).trace(rays, any_hit=True)
    agree = (g["t"] == o["t"]).mean()
    print(f"any-hit agreement {agree:.6f}, occluded {(o['t'] > 0).mean():.3f}")
    assert agree >= 0.9999


def _render_pair(path, oracle, nx=None, region=None, seed_offset=0):
    scene = yar_amd.Scene(path)
    ctx = yar_amd.Context(scene)
    v = ctx.view(0)
    if nx:
        v.x_pixel_sample_count = v.y_pixel_sample_count = nx
    if region:
        v.region_x0, v.region_y0, v.region_x1, v.region_y1 = region
    v.rng_seed_offset = seed_offset
    ctx.set_views([v])
    img, film, var, st = ctx.render(0)
    oimg, ofilm, ovar, _, _ = oracle.OracleScene(scene).render(v)
    return img, film, var, st, oimg, ofilm, ovar


def _compare_images(img, oimg, film, ofilm, label):
    exact_px = np.all(film["color_sum"] == ofilm["color_sum"], axis=-1) & (film["weight_sum"] == ofilm["weight_sum"])
    rel = np.abs(img - oimg) / np.maximum(np.abs(oimg), 1e-3)
    close_px = np.all(rel <= 1e-4, axis=-1)
    mean_rel = abs(img.mean() - oimg.mean()) / max(oimg.mean(), 1e-12)
    print(f"{label}: bit-exact pixels {exact_px.mean():.5f}, pixels within 1e-4 rel {close_px.mean():.5f}, "
          f"image mean gpu {img.mean():.6f} cpu {oimg.mean():.6f} (rel {mean_rel:.2e})")
    return exact_px.mean(), close_px.mean(), mean_rel


def test_render_c1_matches_oracle(c1_path, oracle):
    img, film, var, st, oimg, ofilm, ovar = _render_pair(c1_path, oracle)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 16spp")
    # Paths diverge only where device sin/cos differ from glibc by an ulp and the perturbed
    # direction changes a discrete event; such pixels are rare and the image mean is unaffected.
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert np.isfinite(img).all()
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_c2_small_matches_oracle(c2_small_path, oracle):
    # 160x160 crop through the sphere field: diffuse, GGX conductor, smooth glass, point light
    img, film, var, st, oimg, ofilm, ovar = _render_pair(c2_small_path, oracle, nx=4, region=(432, 432, 592, 592))
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c2-small crop 16spp")
    assert close >= 0.98
    assert mean_rel <= 5e-3


def test_render_plastic_and_mirror_matches_oracle(c1_path, oracle, tmp_path):
    # the config-1 room with a mirror wall, a plastic wall and plastic triangles (pbrt3 plastic,
    # perfect reflector): Pbrt3_Plastic_BRDF (bsdf_pbrt.cpp:264), delta reflection (delta_scattering.cpp:27)
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    lines[mats[0]] = 'Material "plastic" "rgb Kd" [ 0.3 0.4 0.5 ] "rgb Ks" [ 0.5 0.5 0.5 ] "float roughness" [ 0.08 ]'
    lines[mats[2]] = 'Material "mirror" "rgb Kr" [ 0.85 0.8 0.75 ]'
    lines[mats[-1]] = 'Material "plastic" "rgb Kd" [ 0.6 0.3 0.2 ] "float roughness" [ 0.3 ]'
    p = tmp_path / "c1_plastic.pbrt"
    p.write_text("\n".join(lines) + "\n")
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 plastic+mirror 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


@pytest.mark.parametrize("nx", [12, 20], ids=["144spp-u8", "400spp-u16"])
def test_render_high_spp_matches_oracle(c1_path, oracle, nx):
    # sample counts above 128 take the one-wave-per-run shuffle kernel (u8 entries up to 256 spp,
    # u16 beyond): pixel_sampling.cpp:54-85 with every sampler dimension used by depth-5 paths
    img, film, var, st, oimg, ofilm, ovar = _render_pair(c1_path, oracle, nx=nx, region=(96, 96, 112, 112))
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, f"c1 16x16 crop {nx * nx} spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_coated_diffuse_matches_oracle(c1_path, oracle, tmp_path):
    # the config-1 room with an anisotropic substrate wall and substrate triangles:
    # Ashikhmin_Shirley_Phong_BRDF (bsdf.cpp:427-495)
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    lines[mats[1]] = ('Material "substrate" "rgb Kd" [ 0.5 0.3 0.2 ] "rgb Ks" [ 0.08 0.08 0.08 ] '
                      '"float uroughness" [ 0.05 ] "float vroughness" [ 0.25 ]')
    lines[mats[-1]] = 'Material "substrate" "rgb Kd" [ 0.2 0.4 0.6 ]'
    p = tmp_path / "c1_substrate.pbrt"
    p.write_text("\n".join(lines) + "\n")
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 substrate 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_rough_glass_matches_oracle(c1_path, oracle, tmp_path):
    # the config-1 room with rough glass triangles and a rough glass wall (Rough_Glass_BSDF,
    # bsdf.cpp:268-420: GGX reflection and transmission, light sampled through the surface)
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    lines[mats[3]] = 'Material "glass" "float uroughness" [ 0.3 ] "float vroughness" [ 0.3 ]'
    lines[mats[-1]] = ('Material "glass" "rgb Kr" [ 0.9 0.9 0.9 ] "rgb Kt" [ 0.8 0.9 1.0 ] "float index" [ 1.4 ] '
                       '"float uroughness" [ 0.15 ] "float vroughness" [ 0.15 ]')
    p = tmp_path / "c1_rough_glass.pbrt"
    p.write_text("\n".join(lines) + "\n")
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 rough glass 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_spot_directional_mesh_lights_match_oracle(c1_path, oracle, tmp_path):
    # the config-1 room lit by every light type at once: its rectangle light, a spot light aimed at
    # the floor, a distant light through the open front and a triangle-fan emitter under the ceiling
    # (import_pbrt_non_area_light, pbrt_scene.cpp:903-956; Triangle_Mesh_Light, light_sampling.cpp)
    from tests import scenes
    v, idx = scenes.square_fan(0.12, 0.0, facing_down=True)
    pts = " ".join(f"{x + 0.25:.6g} {y + 0.92:.6g} {z + 0.6:.6g}" for x, y, z in v)
    extra = ('LightSource "spot" "rgb I" [ 2 1.5 1 ] "point from" [ 0.7 0.9 0.5 ] "point to" [ 0.6 0 0.4 ] '
             '"float coneangle" [ 25 ] "float conedeltaangle" [ 6 ]\n'
             'LightSource "distant" "rgb L" [ 0.6 0.7 0.8 ] "point from" [ 0.3 0.8 4 ] "point to" [ 0.5 0.4 0 ]\n'
             'AttributeBegin\nAreaLightSource "diffuse" "rgb L" [ 6 5 4 ]\n'
             f'Shape "trianglemesh" "integer indices" [ {" ".join(map(str, idx))} ] "point P" [ {pts} ]\n'
             'AttributeEnd\n')
    text = open(c1_path).read().replace("WorldBegin\n", "WorldBegin\n" + extra, 1)
    p = tmp_path / "c1_lights.pbrt"
    p.write_text(text)
    d = yar_amd.Scene(str(p)).desc
    assert (d.spot_light_count, d.directional_light_count, d.mesh_light_count) == (1, 1, 1)
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 spot+distant+mesh lights 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


SPHERES = '''AttributeBegin
Translate 0.3 0.75 0.4
AreaLightSource "diffuse" "rgb L" [ 8 7 6 ]
Shape "sphere" "float radius" [ 0.06 ]
AttributeEnd
AttributeBegin
Translate 0.7 0.22 0.5
Rotate 35 0.3 1 0.2
Scale 1 1.4 0.8
Material "substrate" "rgb Kd" [ 0.5 0.3 0.2 ] "rgb Ks" [ 0.1 0.1 0.1 ] "float uroughness" [ 0.05 ] "float vroughness" [ 0.3 ]
Shape "sphere" "float radius" [ 0.15 ]
AttributeEnd
AttributeBegin
Translate 0.35 0.15 0.7
Material "metal" "rgb eta" [ 0.2 0.9 1.1 ] "rgb k" [ 3.9 2.4 2.1 ] "float roughness" [ 0.1 ]
Shape "sphere" "float radius" [ 0.12 ]
AttributeEnd
'''


def test_render_spheres_and_sphere_light_match_oracle(c1_path, oracle, tmp_path):
    # the config-1 room with a sphere light and two tessellated spheres under object transforms
    # (rotation, non-uniform scale) with anisotropic coated-diffuse and GGX metal materials:
    # import_pbrt_sphere (pbrt_scene.cpp:745-758), direct_lighting_from_sphere_light
    # (direct_lighting.cpp:196-260), object-space traversal through world_to_object (kdtree.cpp:219-327)
    text = open(c1_path).read().replace("WorldBegin\n", "WorldBegin\n" + SPHERES, 1)
    p = tmp_path / "c1_spheres.pbrt"
    p.write_text(text)
    scene = yar_amd.Scene(str(p))
    assert scene.desc.sphere_light_count == 1
    rays = raygen.random_rays(scene, 1 << 18, seed=5)
    # half the batch aimed from the room's interior at points around the three spheres
    rng = np.random.default_rng(11)
    h = rays.size // 2
    centres = np.array([[0.3, 0.75, 0.4], [0.7, 0.22, 0.5], [0.35, 0.15, 0.7]], np.float32)
    org = rng.uniform(0.02, 0.98, size=(h, 3)).astype(np.float32)
    tgt = centres[rng.integers(0, 3, h)] + rng.normal(scale=0.12, size=(h, 3)).astype(np.float32)
    dd = tgt - org
    rays["origin"][:h] = org
    rays["direction"][:h] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    rays["tmax"][:h] = np.inf
    g = yar_amd.Context(scene).trace_rays(rays)
    o = oracle.OracleScene(scene).trace(rays)
    same, max_rel, exact = _hits_agree(g, o)
    print(f"spheres trace: hit-id agreement {same:.6f}, max rel t err {max_rel:.3g}, bit-exact t {exact:.6f}")
    assert same >= 0.9999
    assert max_rel <= 1e-5
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 spheres + sphere light 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_uber_matches_oracle(c1_path, oracle, tmp_path):
    # the config-1 room with pbrt3 uber walls and triangles: Kd + Ks (BSDF layers only), Kd + Kr
    # (Fresnel mirror chosen per component, direct light in both branches) and Kd + Ks + opacity
    # (pass-through): Pbrt3_Uber_BRDF (bsdf_pbrt.cpp:14-91), get_pbrt_uber_info (delta_scattering.cpp:102-152)
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    lines[mats[0]] = 'Material "uber" "rgb Kd" [ 0.5 0.4 0.3 ] "rgb Ks" [ 0.3 0.3 0.3 ] "float roughness" [ 0.05 ]'
    lines[mats[1]] = 'Material "uber" "rgb Kd" [ 0.2 0.3 0.4 ] "rgb Ks" [ 0 0 0 ] "rgb Kr" [ 0.8 0.8 0.8 ] "float index" [ 1.7 ]'
    lines[mats[-1]] = ('Material "uber" "rgb Kd" [ 0.6 0.3 0.2 ] "rgb Ks" [ 0.2 0.2 0.2 ] "rgb opacity" [ 0.5 0.6 0.7 ] '
                       '"float uroughness" [ 0.02 ] "float vroughness" [ 0.3 ]')
    p = tmp_path / "c1_uber.pbrt"
    p.write_text("\n".join(lines) + "\n")
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 uber 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_environment_light_matches_oracle(c1_path, oracle, tmp_path):
    # the config-1 room under a turned constant sky besides its rectangle light, with a mirror wall so
    # delta rays also leave the scene: Environment_Light_Sampler (light_sampling.cpp:11-73),
    # direct_lighting_from_environment_light (direct_lighting.cpp:324-384), misses at the camera and
    # after delta bounces (path_tracing.cpp:27-33, 86-96)
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    lines[mats[2]] = 'Material "mirror" "rgb Kr" [ 0.85 0.8 0.75 ]'
    text = "\n".join(lines) + "\n"
    text = text.replace("WorldBegin\n", 'WorldBegin\nAttributeBegin\nRotate 30 1 0.5 0\n'
                        'LightSource "infinite" "rgb L" [ 0.3 0.4 0.5 ]\nAttributeEnd\n', 1)
    p = tmp_path / "c1_env.pbrt"
    p.write_text(text)
    assert yar_amd.Scene(str(p)).desc.environment_light_count == 1
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 environment light 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_mix_matches_oracle(c1_path, oracle, tmp_path):
    # the config-1 room with mix materials (matte + plastic, substrate + GGX metal, matte + rough
    # glass): Mix_BSDF (bsdf.cpp:497-570), get_mix_info's u_scat rescaling (delta_scattering.cpp:154-189);
    # the renderer switches to its mix-aware shade kernel for this scene
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    named = ('MakeNamedMaterial "ma" "string type" [ "matte" ] "rgb Kd" [ 0.7 0.3 0.2 ]\n'
             'MakeNamedMaterial "pl" "string type" [ "plastic" ] "rgb Kd" [ 0.1 0.2 0.5 ] "float roughness" [ 0.05 ]\n'
             'MakeNamedMaterial "su" "string type" [ "substrate" ] "rgb Kd" [ 0.4 0.5 0.2 ] "float uroughness" [ 0.05 ] "float vroughness" [ 0.2 ]\n'
             'MakeNamedMaterial "me" "string type" [ "metal" ] "rgb eta" [ 0.2 0.9 1.1 ] "rgb k" [ 3.9 2.4 2.1 ] "float roughness" [ 0.1 ]\n'
             'MakeNamedMaterial "rg" "string type" [ "glass" ] "float uroughness" [ 0.2 ] "float vroughness" [ 0.2 ]\n')
    lines[mats[0]] = 'Material "mix" "string namedmaterial1" [ "ma" ] "string namedmaterial2" [ "pl" ] "rgb amount" [ 0.3 0.5 0.7 ]'
    lines[mats[1]] = 'Material "mix" "string namedmaterial1" [ "su" ] "string namedmaterial2" [ "me" ]'
    lines[mats[-1]] = 'Material "mix" "string namedmaterial1" [ "ma" ] "string namedmaterial2" [ "rg" ] "rgb amount" [ 0.6 0.6 0.6 ]'
    text = "\n".join(lines) + "\n"
    p = tmp_path / "c1_mix.pbrt"
    p.write_text(text.replace("WorldBegin\n", "WorldBegin\n" + named, 1))
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 mix 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()


def test_render_mix_with_delta_layers_matches_oracle(c1_path, oracle, tmp_path):
    # mix materials with one delta sub-material: mirror + matte on a wall, plastic + smooth glass on
    # the triangles (the delta layer picked with its mix weight, the other BSDF scaled: get_mix_info,
    # delta_scattering.cpp:154-189; create_mix_bsdf / Scaled_BSDF, bsdf.cpp:572-629)
    lines = open(c1_path).read().splitlines()
    mats = [i for i, l in enumerate(lines) if l.startswith('Material "matte"')]
    named = ('MakeNamedMaterial "mi" "string type" [ "mirror" ] "rgb Kr" [ 0.9 0.85 0.8 ]\n'
             'MakeNamedMaterial "ma" "string type" [ "matte" ] "rgb Kd" [ 0.5 0.3 0.2 ]\n'
             'MakeNamedMaterial "gl" "string type" [ "glass" ] "float index" [ 1.5 ]\n'
             'MakeNamedMaterial "pl" "string type" [ "plastic" ] "rgb Kd" [ 0.2 0.4 0.6 ] "float roughness" [ 0.1 ]\n')
    lines[mats[2]] = 'Material "mix" "string namedmaterial1" [ "mi" ] "string namedmaterial2" [ "ma" ] "rgb amount" [ 0.6 0.6 0.6 ]'
    lines[mats[-1]] = 'Material "mix" "string namedmaterial1" [ "pl" ] "string namedmaterial2" [ "gl" ] "rgb amount" [ 0.3 0.4 0.5 ]'
    text = "\n".join(lines) + "\n"
    p = tmp_path / "c1_mix_delta.pbrt"
    p.write_text(text.replace("WorldBegin\n", "WorldBegin\n" + named, 1))
    img, film, var, st, oimg, ofilm, ovar = _render_pair(str(p), oracle, nx=4)
    exact, close, mean_rel = _compare_images(img, oimg, film, ofilm, "c1 mix with delta layers 16spp")
    assert close >= 0.99
    assert mean_rel <= 1e-3
    assert abs(var - ovar) / max(ovar, 1e-12) <= 1e-2
    assert (film["weight_sum"] == ofilm["weight_sum"]).all()

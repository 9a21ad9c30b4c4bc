"""pbrt3-subset loader semantics (load_pbrt_scene, src/lib/pbrt_scene.cpp) - CPU only."""
import numpy as np
import pytest

import yar_amd

HEAD = '''LookAt 0 0 5  0 0 0  0 1 0
Camera "perspective" "float fov" [ 40 ]
Film "image" "integer xresolution" [ 200 ] "integer yresolution" [ 100 ]
Sampler "stratified" "integer xsamples" [ 2 ] "integer ysamples" [ 3 ]
Integrator "path" "integer maxdepth" [ 7 ]
WorldBegin
'''


def _load(tmp_path, body, head=HEAD, name="s.pbrt"):
    p = tmp_path / name
    p.write_text(head + body + "WorldEnd\n")
    return yar_amd.Scene(str(p))


def _arr(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


LIGHT = '''AttributeBegin
AreaLightSource "diffuse" "rgb L" [ 4 5 6 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ -1 2 -1  1 2 -1  1 2 1  -1 2 1 ]
AttributeEnd
'''


def test_camera_film_sampler_integrator(tmp_path):
    s = _load(tmp_path, LIGHT)
    d = s.desc
    c2w = np.array(d.camera.camera_to_world[:]).reshape(3, 4)
    # LookAt frame (Parser.inl:695) -> camera basis (-right, up, -forward) for y-up (pbrt_scene.cpp:1020-1030)
    assert np.allclose(c2w[:, :3], np.eye(3), atol=1e-6)
    assert np.allclose(c2w[:, 3], [0, 0, 5], atol=1e-6)
    assert d.camera.z_is_up == 0
    assert d.camera.fov_y_degrees == pytest.approx(40.0)
    c = d.config
    assert (c.film_width, c.film_height) == (200, 100)
    assert (c.region_x0, c.region_y0, c.region_x1, c.region_y1) == (0, 0, 200, 100)
    assert (c.x_pixel_sample_count, c.y_pixel_sample_count) == (2, 3)
    assert c.max_light_bounces == 7 and c.russian_roulette_threshold == 1.0
    assert c.russian_roulette_bounce_count_threshold == 3
    assert c.pixel_filter_radius == 0.5 and c.pbrt3_scene == 1
    assert np.isinf(c.max_rgb_component_value_of_film_sample) and c.film_radiance_scale == 1.0


def test_rectangle_light_detection(tmp_path):
    s = _load(tmp_path, LIGHT)
    d = s.desc
    assert d.rect_light_count == 1 and d.object_count == 1
    L = d.rect_lights[0]
    l2w = np.array(L.light_to_world[:]).reshape(3, 4)
    assert list(L.size) == pytest.approx([2.0, 2.0])
    assert np.allclose(l2w[:, 3], [0, 2, 0])                 # mid point
    x, y, z = l2w[:, 0], l2w[:, 1], l2w[:, 2]
    assert np.allclose(x, [1, 0, 0]) and np.allclose(y, [0, 0, 1])   # first right angle (p0->p1, p1->p2)
    assert np.allclose(z, np.cross(x, y))
    assert list(L.emitted_radiance) == [4, 5, 6]
    o = d.objects[0]
    assert o.area_light == 0 and o.material == -1


def test_reverse_orientation_flips_light_frame(tmp_path):
    s = _load(tmp_path, LIGHT.replace("AreaLightSource", "ReverseOrientation\nAreaLightSource"))
    l2w = np.array(s.desc.rect_lights[0].light_to_world[:]).reshape(3, 4)
    assert np.allclose(l2w[:, 0], [-1, 0, 0]) and np.allclose(l2w[:, 2], [0, 1, 0])
    assert s.desc.meshes[0].reverse_geometric_normal_orientation == 1


def test_default_uvs_and_degenerate_removal(tmp_path):
    body = LIGHT + '''Material "matte" "rgb Kd" [ 0.1 0.2 0.3 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ 0 0 0  1 0 0  1 1 0  0 1 0 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 2 3 4 ] "point P" [ 0 0 0  1 0 0  0 1 0  0 1 0  5 5 5 ]
'''
    d = _load(tmp_path, body).desc
    assert d.mesh_count == 4
    quad, tri, other = d.meshes[1], d.meshes[2], d.meshes[3]
    assert np.array_equal(_arr(quad.uvs, 8), [0, 0, 1, 0, 1, 1, 1, 0])      # pbrt_scene.cpp:704-712
    assert np.array_equal(_arr(tri.uvs, 6), [0, 0, 1, 0, 1, 1])              # pbrt_scene.cpp:714-720
    assert not other.uvs
    assert other.triangle_count == 1                                          # (2,3,4) has coincident vertices
    assert list(d.materials[d.objects[1].material].reflectance) == pytest.approx([0.1, 0.2, 0.3])


def test_materials_metal_glass(tmp_path):
    body = LIGHT + '''Material "metal" "rgb eta" [ 0.2 0.9 1.1 ] "rgb k" [ 3.9 2.4 2.1 ] "float uroughness" [ 0.3 ] "float vroughness" [ 0.1 ] "bool remaproughness" "false"
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]
Material "glass"
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 1  1 0 1  0 1 1 ]
'''
    d = _load(tmp_path, body).desc
    m = d.materials[d.objects[1].material]
    assert m.type == yar_amd._abi.YAR_MATERIAL_METAL and m.roughness_is_alpha == 1
    assert (m.u_roughness, m.v_roughness) == pytest.approx((0.3, 0.1)) and m.eta_i == 1.0
    g = d.materials[d.objects[2].material]
    assert g.type == yar_amd._abi.YAR_MATERIAL_GLASS
    assert g.index_of_refraction == pytest.approx(1.5) and list(g.reflectance) == [1, 1, 1]


def test_unsupported_inputs_are_parse_errors(tmp_path):
    for body in ['Material "translucent"\nShape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]\n',
                 'Material "glass" "float uroughness" [ 0.2 ]\nShape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]\n',
                 'Material "metal"\nShape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]\n']:
        with pytest.raises(yar_amd.YarError) as e:
            _load(tmp_path, LIGHT + body)
        assert e.value.code == 3


def test_tall_image_fov_and_crop_window(tmp_path):
    head = HEAD.replace('"integer xresolution" [ 200 ] "integer yresolution" [ 100 ]',
                        '"integer xresolution" [ 100 ] "integer yresolution" [ 200 ] "float cropwindow" [ 0.1 0.5 0.25 0.75 ]')
    d = _load(tmp_path, LIGHT, head=head).desc
    # pbrt fov is the narrower (horizontal) extent for tall images (pbrt_scene.cpp:1034-1042)
    want = np.degrees(2 * np.arctan(np.tan(np.radians(20.0)) * 2.0))
    assert d.camera.fov_y_degrees == pytest.approx(want, rel=1e-5)
    c = d.config
    # crop window ceil(res * fraction), x mirrored for the right-handed camera (pbrt_scene.cpp:1104-1112)
    assert (c.region_x0, c.region_x1) == (100 - 50, 100 - 10)
    assert (c.region_y0, c.region_y1) == (50, 150)


def test_sampler_fallback_and_missing_integrator(tmp_path):
    head = HEAD.replace('Sampler "stratified" "integer xsamples" [ 2 ] "integer ysamples" [ 3 ]',
                        'Sampler "halton" "integer pixelsamples" [ 10 ]').replace('Integrator "path" "integer maxdepth" [ 7 ]\n', '')
    c = _load(tmp_path, LIGHT, head=head).desc.config
    assert (c.x_pixel_sample_count, c.y_pixel_sample_count) == (4, 4)        # k = ceil(sqrt(10))
    assert c.max_light_bounces == 32                                          # raytracer_config.h:21


def test_point_light(tmp_path):
    d = _load(tmp_path, LIGHT + 'LightSource "point" "rgb I" [ 1 2 3 ] "rgb scale" [ 2 2 2 ] "point from" [ 1 2 3 ]\n').desc
    assert d.point_light_count == 1
    assert list(d.point_lights[0].position) == [1, 2, 3]
    assert list(d.point_lights[0].intensity) == [2, 4, 6]


def test_shape_transform_is_baked_into_world_space(tmp_path):
    """pbrt-parser bakes the CTM into P and N (impl/semantic/Geometry.cpp:248-252), so a transformed
    area light is detected as a rectangle in world space and every object transform is the identity."""
    body = '''AttributeBegin
Translate 0 3 0
Rotate 90 0 0 1
AreaLightSource "diffuse" "rgb L" [ 1 1 1 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ -1 0 -0.5  1 0 -0.5  1 0 0.5  -1 0 0.5 ]
AttributeEnd
AttributeBegin
Translate 10 0 0
Scale 2 1 1
Material "matte"
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ] "normal N" [ 1 1 0  1 1 0  1 1 0 ]
AttributeEnd
'''
    s = _load(tmp_path, body)
    d = s.desc
    assert d.rect_light_count == 1 and d.object_count == 2
    # Rotate 90 about z maps x -> y; the quad spans y in [2, 4] and z in [-0.5, 0.5] at x = 0
    P = _arr(d.meshes[0].vertices, 12).reshape(4, 3)
    assert np.allclose(P, [[0, 2, -0.5], [0, 4, -0.5], [0, 4, 0.5], [0, 2, 0.5]], atol=1e-6)
    L = d.rect_lights[0]
    l2w = np.array(L.light_to_world[:]).reshape(3, 4)
    assert list(L.size) == pytest.approx([2.0, 1.0])
    assert np.allclose(l2w[:, 3], [0, 3, 0], atol=1e-6)          # world-space mid point
    assert np.allclose(l2w[:, 0], [0, 1, 0], atol=1e-6)          # first edge p0->p1, now along +y
    assert np.allclose(l2w[:, 1], [0, 0, 1], atol=1e-6)
    # the emitter's triangles lie inside the world-space rectangle the light samples
    c, ex, ey = l2w[:, 3], l2w[:, 0], l2w[:, 1]
    q = P - c
    assert np.all(np.abs(q @ ex) <= 0.5 * L.size[0] + 1e-6) and np.all(np.abs(q @ ey) <= 0.5 * L.size[1] + 1e-6)
    for o in (d.objects[0], d.objects[1]):
        assert np.allclose(np.array(o.object_to_world[:]).reshape(3, 4), np.eye(3, 4))
        assert np.allclose(np.array(o.world_to_object[:]).reshape(3, 4), np.eye(3, 4))
    # Translate 10 0 0 * Scale 2 1 1: points scale then move; normals use the inverse transpose
    P2 = _arr(d.meshes[1].vertices, 9).reshape(3, 3)
    assert np.allclose(P2, [[10, 0, 0], [12, 0, 0], [10, 1, 0]])
    N2 = _arr(d.meshes[1].normals, 9).reshape(3, 3)
    assert np.allclose(N2, [[0.5, 1, 0]] * 3)


def test_plastic_and_mirror_materials(tmp_path):
    # pbrt-parser defaults (Kd = Ks = 0.25, roughness 0.1, mirror Kr = 0.9) and r0 = Y(sRGB -> XYZ of Ks)
    # (pbrt_scene.cpp:443-447, 503-525)
    body = LIGHT + '''Material "plastic" "rgb Kd" [ 0.1 0.2 0.3 ] "rgb Ks" [ 0.4 0.5 0.6 ] "float roughness" [ 0.05 ]
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]
Material "plastic"
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 1 1 0 1 0 1 1]
Material "mirror" "rgb Kr" [ 0.8 0.7 0.6 ]
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 2 1 0 2 0 1 2]
Material "mirror"
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 3 1 0 3 0 1 3]
'''
    d = _load(tmp_path, body).desc
    m = [d.materials[i] for i in range(d.material_count)]
    assert [x.type for x in m] == [5, 5, 4, 4]
    assert list(m[0].reflectance) == pytest.approx([0.1, 0.2, 0.3])
    assert m[0].specular_r0 == pytest.approx(0.2126729 * 0.4 + 0.7151522 * 0.5 + 0.0721750 * 0.6)
    assert m[0].u_roughness == pytest.approx(0.05) and m[0].roughness_is_alpha == 0
    assert list(m[1].reflectance) == pytest.approx([0.25] * 3) and m[1].u_roughness == pytest.approx(0.1)
    assert m[1].specular_r0 == pytest.approx(0.25, rel=1e-6)
    assert list(m[2].reflectance) == pytest.approx([0.8, 0.7, 0.6]) and list(m[3].reflectance) == pytest.approx([0.9] * 3)


def test_substrate_material(tmp_path):
    # pbrt "substrate" -> Coated_Diffuse_Material (pbrt_scene.cpp:531-561); pbrt-parser defaults
    # Kd = Ks = 0.5, uroughness = vroughness = 0.1, remaproughness true
    body = LIGHT + '''Material "substrate" "rgb Kd" [ 0.1 0.2 0.3 ] "rgb Ks" [ 0.04 0.05 0.06 ] "float uroughness" [ 0.02 ] "float vroughness" [ 0.3 ] "bool remaproughness" "false"
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]
Material "substrate"
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 1 1 0 1 0 1 1]
'''
    d = _load(tmp_path, body).desc
    m = [d.materials[i] for i in range(d.material_count)]
    assert [x.type for x in m] == [6, 6]
    assert list(m[0].reflectance) == pytest.approx([0.1, 0.2, 0.3])
    assert list(m[0].specular) == pytest.approx([0.04, 0.05, 0.06])
    assert (m[0].u_roughness, m[0].v_roughness, m[0].roughness_is_alpha) == (pytest.approx(0.02), pytest.approx(0.3), 1)
    assert list(m[1].reflectance) == pytest.approx([0.5] * 3) and list(m[1].specular) == pytest.approx([0.5] * 3)
    assert (m[1].u_roughness, m[1].v_roughness, m[1].roughness_is_alpha) == (pytest.approx(0.1), pytest.approx(0.1), 0)
    bad = LIGHT + '''Texture "t" "spectrum" "imagemap" "string filename" "x.png"
Material "substrate" "texture Kd" "t"
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]
'''
    with pytest.raises(yar_amd.YarError):
        _load(tmp_path, bad)


def test_rough_glass_material(tmp_path):
    # glass with u = v roughness > 0 -> Rough_Glass_BSDF (pbrt_scene.cpp:449-462); u != v is rejected
    body = LIGHT + '''Material "glass" "float uroughness" [ 0.2 ] "float vroughness" [ 0.2 ] "float index" [ 1.33 ]
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]
'''
    d = _load(tmp_path, body).desc
    m = d.materials[0]
    assert m.type == 3 and m.u_roughness == pytest.approx(0.2) and m.v_roughness == pytest.approx(0.2)
    assert m.index_of_refraction == pytest.approx(1.33) and m.roughness_is_alpha == 0
    bad = LIGHT + '''Material "glass" "float uroughness" [ 0.2 ] "float vroughness" [ 0.1 ]
Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]
'''
    with pytest.raises(yar_amd.YarError):
        _load(tmp_path, bad)


def test_spot_and_distant_lights(tmp_path):
    """import_pbrt_non_area_light (pbrt_scene.cpp:903-956): spot from/to go through the CTM, cone
    angles are degrees; a distant light shines along from - to with L * scale as irradiance."""
    body = LIGHT + '''AttributeBegin
Translate 0 1 0
LightSource "spot" "rgb I" [ 3 2 1 ] "point from" [ 0 2 0 ] "point to" [ 0 2 -4 ] "float coneangle" [ 20 ] "float conedeltaangle" [ 4 ]
AttributeEnd
LightSource "spot" "rgb I" [ 1 1 1 ]
LightSource "distant" "rgb L" [ 2 2 2 ] "rgb scale" [ 0.5 1 2 ] "point from" [ 0 0 0 ] "point to" [ 0 -3 0 ]
LightSource "distant"
'''
    d = _load(tmp_path, body).desc
    assert d.spot_light_count == 2 and d.directional_light_count == 2 and d.point_light_count == 0
    s0, s1 = d.spot_lights[0], d.spot_lights[1]
    assert list(s0.position) == [0, 3, 0]
    assert np.allclose(list(s0.direction), [0, 0, -1])
    assert s0.cone_angle == pytest.approx(np.radians(20), rel=1e-6)
    assert s0.penumbra_angle == pytest.approx(np.radians(4), rel=1e-6)
    assert list(s0.intensity) == [3, 2, 1]
    assert s1.cone_angle == pytest.approx(np.radians(30), rel=1e-6)          # pbrt3 defaults
    assert s1.penumbra_angle == pytest.approx(np.radians(5), rel=1e-6)
    l0, l1 = d.directional_lights[0], d.directional_lights[1]
    assert np.allclose(list(l0.direction), [0, 1, 0])                          # from - to, normalised
    assert list(l0.irradiance) == [1, 2, 4]
    assert np.allclose(list(l1.direction), [0, 0, -1])                         # default to = (0, 0, 1)
    assert list(l1.irradiance) == [1, 1, 1]


def test_point_light_from_goes_through_ctm(tmp_path):
    body = LIGHT + 'AttributeBegin\nTranslate 1 0 0\nScale 2 2 2\nLightSource "point" "point from" [ 1 1 1 ]\nAttributeEnd\n'
    d = _load(tmp_path, body).desc
    assert list(d.point_lights[0].position) == [3, 2, 2]
    assert list(d.point_lights[0].intensity) == [1, 1, 1]


def test_triangle_mesh_area_light(tmp_path):
    """import_pbrt_shape (pbrt_scene.cpp:786-794): an emitter that is not a rectangle becomes a
    triangle-mesh light; its object's light index counts after every rectangle light."""
    fan = '''AttributeBegin
AreaLightSource "diffuse" "rgb L" [ 7 8 9 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 0 3 4 ] "point P" [ 0 3 0  1 3 0  1 3 1  0 3 2  -1 3 1 ]
AttributeEnd
'''
    d = _load(tmp_path, fan + LIGHT + 'Material "matte"\nShape "trianglemesh" "integer indices" [ 0 1 2 ] '
              '"point P" [ 0 0 0  1 0 0  0 1 0 ]\n').desc
    assert d.rect_light_count == 1 and d.mesh_light_count == 1
    assert d.mesh_lights[0].object == 0
    assert list(d.mesh_lights[0].emitted_radiance) == [7, 8, 9]
    objs = [d.objects[i] for i in range(d.object_count)]
    assert [o.area_light for o in objs] == [1, 0, -1]                          # mesh light after 1 rect
    assert objs[0].material == -1 and objs[2].material >= 0


SPHERES = '''AttributeBegin
Translate 1 2 3
Scale 2 2 2
AreaLightSource "diffuse" "rgb L" [ 3 4 5 ] "integer nsamples" [ 4 ]
Shape "sphere" "float radius" [ 0.25 ]
AttributeEnd
AttributeBegin
Translate -1 0 0
Rotate 30 0 1 0
Material "matte"
Shape "sphere" "float radius" [ 0.25 ]
AttributeEnd
Material "matte"
Shape "sphere"
'''


def test_sphere_shapes_are_tessellated_geodesic_spheres(tmp_path):
    """import_pbrt_sphere (pbrt_scene.cpp:745-758): one level-6 geodesic tessellation per radius,
    shared by every sphere of that radius, in object space under the shape's transform."""
    d = _load(tmp_path, LIGHT + SPHERES).desc
    assert d.mesh_count == 3 and d.object_count == 4                         # rect, two r=0.25, one r=1
    objs = [d.objects[i] for i in range(d.object_count)]
    assert objs[1].mesh == objs[2].mesh != objs[3].mesh
    m = d.meshes[objs[1].mesh]
    assert m.triangle_count == 20 * 4 ** 6 and m.vertex_count == 10 * 4 ** 6 + 2   # every shared vertex merged
    v = _arr(m.vertices, 3 * m.vertex_count).reshape(-1, 3)
    n = _arr(m.normals, 3 * m.vertex_count).reshape(-1, 3)
    uv = _arr(m.uvs, 2 * m.vertex_count).reshape(-1, 2)
    idx = _arr(m.indices, 3 * m.triangle_count).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 0.25, rtol=1e-6)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, rtol=1e-6)
    assert (np.sum(n * v, axis=1) > 0.2499).all()                            # outward, near radial
    # faces wind outwards: the geometric normal points away from the centre
    a, b, c = v[idx[:, 0]], v[idx[:, 1]], v[idx[:, 2]]
    assert (np.sum(np.cross(b - a, c - a) * (a + b + c), axis=1) > 0).all()
    # the first level keeps the icosahedron's vertex 0 = (g, 1, 0)/sqrt(1 + g^2) first
    g = (1 + np.sqrt(5)) / 2
    assert np.allclose(v[0], np.array([g, 1, 0]) / np.sqrt(1 + g * g) * 0.25, atol=1e-7)
    # uv: v = acos(-z)/pi is 0 at the -z pole, u = atan2(y, x)/2pi in [0, 1)
    assert np.allclose(uv[:, 1], np.arccos(np.clip(-v[:, 2] / 0.25, -1, 1)) / np.pi, atol=2e-6)
    assert (uv[:, 0] >= 0).all() and (uv[:, 0] < 1).all()
    o2w = np.array(list(objs[1].object_to_world)).reshape(3, 4)
    assert np.allclose(o2w, [[2, 0, 0, 1], [0, 2, 0, 2], [0, 0, 2, 3]])
    w2o = np.array(list(objs[1].world_to_object)).reshape(3, 4)
    assert np.allclose(w2o, [[0.5, 0, 0, -0.5], [0, 0.5, 0, -1], [0, 0, 0.5, -1.5]])
    assert d.meshes[objs[3].mesh].vertex_count == m.vertex_count
    assert np.allclose(np.linalg.norm(_arr(d.meshes[objs[3].mesh].vertices, 30).reshape(-1, 3), axis=1), 1.0)


def test_sphere_light(tmp_path):
    """Diffuse_Sphere_Light (pbrt_scene.cpp:857-866): centre = the transform's translation, radius
    unscaled; area-light indices run rectangles, spheres, meshes."""
    fan = '''AttributeBegin
AreaLightSource "diffuse" "rgb L" [ 1 1 1 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 3 0  1 3 0  1 3 1 ]
AttributeEnd
'''
    d = _load(tmp_path, fan + LIGHT + SPHERES).desc
    assert (d.rect_light_count, d.sphere_light_count, d.mesh_light_count) == (1, 1, 1)
    sl = d.sphere_lights[0]
    assert sl.object == 2 and list(sl.position) == [1, 2, 3] and sl.radius == 0.25
    assert list(sl.emitted_radiance) == [3, 4, 5] and sl.sample_count == 4
    objs = [d.objects[i] for i in range(d.object_count)]
    assert [o.area_light for o in objs] == [2, 0, 1, -1, -1]
    assert objs[2].material == -1 and objs[3].material >= 0


def test_uber_material(tmp_path):
    """import_pbrt_material's uber branch (pbrt_scene.cpp:571-636) with pbrt-parser's defaults
    (Kd = Ks = 0.25, Kr = Kt = 0, opacity 1, roughness 0.1, index 1.5); v roughness follows u."""
    body = LIGHT + '''Material "uber"
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]
Material "uber" "rgb Kd" [ 0.1 0.2 0.3 ] "rgb Ks" [ 0 0 0 ] "rgb Kr" [ 0.5 0.5 0.5 ] "rgb opacity" [ 0.4 0.5 0.6 ] "float index" [ 1.33 ] "float uroughness" [ 0.02 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 1  1 0 1  0 1 1 ]
'''
    d = _load(tmp_path, body).desc
    m0, m1 = d.materials[0], d.materials[1]
    assert m0.type == m1.type == yar_amd._abi.YAR_MATERIAL_PBRT3_UBER
    assert np.allclose(list(m0.reflectance), 0.25) and np.allclose(list(m0.specular), 0.25)
    assert list(m0.delta_reflectance) == [0, 0, 0] and list(m0.opacity) == [1, 1, 1]
    assert m0.u_roughness == m0.v_roughness == pytest.approx(0.1) and m0.index_of_refraction == 1.5
    assert m0.roughness_is_alpha == 0
    assert np.allclose(list(m1.reflectance), [0.1, 0.2, 0.3]) and list(m1.specular) == [0, 0, 0]
    assert list(m1.delta_reflectance) == [0.5, 0.5, 0.5] and np.allclose(list(m1.opacity), [0.4, 0.5, 0.6])
    assert m1.u_roughness == m1.v_roughness == pytest.approx(0.02) and m1.index_of_refraction == pytest.approx(1.33)


@pytest.mark.parametrize("params", ['"rgb Kt" [ 0.5 0.5 0.5 ]', '"texture Kd" "checks"', '"float alpha" [ 0.5 ]',
                                    '"rgb Kd" [ 0 0 0 ] "rgb Ks" [ 0 0 0 ]', '"bool remaproughness" "false"'])
def test_uber_unsupported_forms_are_parse_errors(tmp_path, params):
    body = LIGHT + f'Material "uber" {params}\nShape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]\n'
    with pytest.raises(yar_amd.YarError):
        _load(tmp_path, body)


def test_infinite_light_without_map(tmp_path):
    """import_pbrt_non_area_light's infinite branch (pbrt_scene.cpp:958-980): radiance L * scale, the
    light frame is the current transform; environment maps are not supported."""
    body = LIGHT + 'AttributeBegin\nRotate 90 1 0 0\nLightSource "infinite" "rgb L" [ 0.5 1 2 ] "rgb scale" [ 2 2 2 ] "integer nsamples" [ 4 ]\nAttributeEnd\n'
    d = _load(tmp_path, body).desc
    assert d.environment_light_count == 1
    e = d.environment_light[0]
    assert list(e.radiance) == [1, 2, 4] and e.sample_count == 4
    l2w = np.array(list(e.light_to_world)).reshape(3, 4)
    w2l = np.array(list(e.world_to_light)).reshape(3, 4)
    assert np.allclose(l2w[:, :3], [[1, 0, 0], [0, 0, -1], [0, 1, 0]], atol=1e-6)
    assert np.allclose(l2w[:, :3] @ w2l[:, :3], np.eye(3), atol=1e-6)
    d = _load(tmp_path, 'LightSource "infinite"\n', name="only_env.pbrt").desc
    assert d.environment_light_count == 1 and list(d.environment_light[0].radiance) == [1, 1, 1]
    for bad in ('LightSource "infinite" "string mapname" "sky.exr"\n',
                'LightSource "infinite"\nLightSource "infinite"\n'):
        with pytest.raises(yar_amd.YarError):
            _load(tmp_path, bad, name="bad.pbrt")


def test_mix_material(tmp_path):
    """import_pbrt_material's mix branch (pbrt_scene.cpp:563-569) over pbrt-parser's named
    sub-materials (Materials.cpp:376-402), amount default 0.5."""
    body = LIGHT + '''MakeNamedMaterial "a" "string type" [ "matte" ] "rgb Kd" [ 0.1 0.2 0.3 ]
MakeNamedMaterial "b" "string type" [ "plastic" ]
Material "mix" "string namedmaterial1" [ "a" ] "string namedmaterial2" [ "b" ] "rgb amount" [ 0.2 0.4 0.6 ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]
Material "mix" "string namedmaterial1" [ "b" ] "string namedmaterial2" [ "a" ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 1  1 0 1  0 1 1 ]
'''
    d = _load(tmp_path, body).desc
    mats = [d.materials[i] for i in range(d.material_count)]
    mixes = [m for m in mats if m.type == yar_amd._abi.YAR_MATERIAL_MIX]
    assert len(mixes) == 2
    a, b = mixes[0].mix_materials
    assert mats[a].type == yar_amd._abi.YAR_MATERIAL_DIFFUSE and mats[b].type == yar_amd._abi.YAR_MATERIAL_PLASTIC
    assert np.allclose(list(mixes[0].mix_amount), [0.2, 0.4, 0.6])
    assert list(mixes[1].mix_materials) == [b, a] and list(mixes[1].mix_amount) == [0.5, 0.5, 0.5]


@pytest.mark.parametrize("subs", [('"matte"', '"uber"'), ('"mix"', '"matte"'), ('"mirror"', '"glass"')],
                         ids=["uber", "nested-mix", "two-delta-layers"])
def test_mix_unsupported_sub_materials_are_parse_errors(tmp_path, subs):
    body = LIGHT + f'''MakeNamedMaterial "a" "string type" [ {subs[0]} ]
MakeNamedMaterial "b" "string type" [ {subs[1]} ]
Material "mix" "string namedmaterial1" [ "a" ] "string namedmaterial2" [ "b" ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]
'''
    with pytest.raises(yar_amd.YarError):
        _load(tmp_path, body)


def test_mix_with_one_delta_layer_loads(tmp_path):
    body = LIGHT + '''MakeNamedMaterial "a" "string type" [ "mirror" ]
MakeNamedMaterial "b" "string type" [ "matte" ]
Material "mix" "string namedmaterial1" [ "a" ] "string namedmaterial2" [ "b" ]
Shape "trianglemesh" "integer indices" [ 0 1 2 ] "point P" [ 0 0 0  1 0 0  0 1 0 ]
'''
    d = _load(tmp_path, body).desc
    mix = [d.materials[i] for i in range(d.material_count) if d.materials[i].type == yar_amd._abi.YAR_MATERIAL_MIX][0]
    assert d.materials[mix.mix_materials[0]].type == yar_amd._abi.YAR_MATERIAL_MIRROR

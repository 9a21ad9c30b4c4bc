"""Hand-built yar_scene_desc instances for tests (analytic scenes, fixture meshes)."""
import ctypes as C

import numpy as np

from yar_amd import _abi


def identity12():
    return [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]


class DescScene:
    """Owns numpy buffers and the ctypes SceneDesc pointing at them."""

    def __init__(self, meshes, materials, objects, rects=(), points=(), camera=None, config=None, spots=(), dirs=(),
                 mesh_lights=()):
        self._keep = []
        md = (_abi.MeshDesc * len(meshes))()
        for i, m in enumerate(meshes):
            v = np.ascontiguousarray(m["vertices"], np.float32).ravel()
            idx = np.ascontiguousarray(m["indices"], np.int32).ravel()
            self._keep += [v, idx]
            md[i].vertices = v.ctypes.data_as(C.POINTER(C.c_float))
            md[i].indices = idx.ctypes.data_as(C.POINTER(C.c_int32))
            md[i].vertex_count = v.size // 3
            md[i].triangle_count = idx.size // 3
            if m.get("normals") is not None:
                n = np.ascontiguousarray(m["normals"], np.float32).ravel()
                self._keep.append(n)
                md[i].normals = n.ctypes.data_as(C.POINTER(C.c_float))
            if m.get("uvs") is not None:
                uv = np.ascontiguousarray(m["uvs"], np.float32).ravel()
                self._keep.append(uv)
                md[i].uvs = uv.ctypes.data_as(C.POINTER(C.c_float))
            md[i].reverse_geometric_normal_orientation = int(m.get("reverse", 0))
        mt = (_abi.MaterialDesc * max(1, len(materials)))()
        for i, m in enumerate(materials):
            mt[i].type = m.get("type", _abi.YAR_MATERIAL_DIFFUSE)
            mt[i].reflectance[:] = m.get("reflectance", [0.5, 0.5, 0.5])
            mt[i].transmittance[:] = m.get("transmittance", [1, 1, 1])
            mt[i].eta[:] = m.get("eta", [0.2, 0.9, 1.1])
            mt[i].k[:] = m.get("k", [3.9, 2.4, 2.1])
            mt[i].eta_i = m.get("eta_i", 1.0)
            mt[i].u_roughness = m.get("u_roughness", 0.0)
            mt[i].v_roughness = m.get("v_roughness", m.get("u_roughness", 0.0))
            mt[i].index_of_refraction = m.get("ior", 1.5)
            mt[i].roughness_is_alpha = m.get("roughness_is_alpha", 1)
            mt[i].specular[:] = m.get("specular", [0, 0, 0])
            mt[i].delta_reflectance[:] = m.get("delta_reflectance", [0, 0, 0])
            mt[i].opacity[:] = m.get("opacity", [1, 1, 1])
            mt[i].mix_materials[:] = m.get("mix_materials", [0, 0])
            mt[i].mix_amount[:] = m.get("mix_amount", [0.5, 0.5, 0.5])
        ob = (_abi.ObjectDesc * len(objects))()
        for i, o in enumerate(objects):
            ob[i].mesh = o["mesh"]
            ob[i].material = o.get("material", -1)
            ob[i].area_light = o.get("light", -1)
            ob[i].object_to_world[:] = o.get("o2w", identity12())
            ob[i].world_to_object[:] = o.get("w2o", identity12())
        rl = (_abi.RectLightDesc * max(1, len(rects)))()
        for i, r in enumerate(rects):
            rl[i].light_to_world[:] = r["l2w"]
            rl[i].emitted_radiance[:] = r["L"]
            rl[i].size[:] = r["size"]
            rl[i].sample_count = 1
        pl = (_abi.PointLightDesc * max(1, len(points)))()
        for i, p in enumerate(points):
            pl[i].position[:] = p["position"]
            pl[i].intensity[:] = p["intensity"]
        sp = (_abi.SpotLightDesc * max(1, len(spots)))()
        for i, l in enumerate(spots):
            sp[i].position[:] = l["position"]
            sp[i].direction[:] = l["direction"]
            sp[i].cone_angle = l["cone_angle"]
            sp[i].penumbra_angle = l["penumbra_angle"]
            sp[i].intensity[:] = l["intensity"]
        dl = (_abi.DirectionalLightDesc * max(1, len(dirs)))()
        for i, l in enumerate(dirs):
            dl[i].direction[:] = l["direction"]
            dl[i].irradiance[:] = l["irradiance"]
        ml = (_abi.MeshLightDesc * max(1, len(mesh_lights)))()
        for i, l in enumerate(mesh_lights):
            ml[i].object = l["object"]
            ml[i].emitted_radiance[:] = l["L"]
        self._keep += [md, mt, ob, rl, pl, sp, dl, ml]
        d = _abi.SceneDesc()
        d.meshes, d.mesh_count = md, len(meshes)
        d.materials, d.material_count = mt, len(materials)
        d.objects, d.object_count = ob, len(objects)
        d.rect_lights, d.rect_light_count = rl, len(rects)
        d.point_lights, d.point_light_count = pl, len(points)
        d.spot_lights, d.spot_light_count = sp, len(spots)
        d.directional_lights, d.directional_light_count = dl, len(dirs)
        d.mesh_lights, d.mesh_light_count = ml, len(mesh_lights)
        cam = camera or {}
        d.camera.camera_to_world[:] = cam.get("c2w", [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 5])
        d.camera.fov_y_degrees = cam.get("fov", 45.0)
        d.camera.z_is_up = cam.get("z_is_up", 0)
        cfg = dict(film_width=64, film_height=64, region=(0, 0, 64, 64), nx=1, ny=1, maxdepth=5,
                   rr_bounce=3, rr_thr=1.0, scale=1.0, max_component=float("inf"), seed_offset=0, pbrt3=1)
        cfg.update(config or {})
        c = d.config
        c.film_width, c.film_height = cfg["film_width"], cfg["film_height"]
        c.region_x0, c.region_y0, c.region_x1, c.region_y1 = cfg["region"]
        c.x_pixel_sample_count, c.y_pixel_sample_count = cfg["nx"], cfg["ny"]
        c.max_light_bounces = cfg["maxdepth"]
        c.russian_roulette_bounce_count_threshold = cfg["rr_bounce"]
        c.russian_roulette_threshold = cfg["rr_thr"]
        c.film_radiance_scale = cfg["scale"]
        c.max_rgb_component_value_of_film_sample = cfg["max_component"]
        c.pixel_filter_radius = 0.5
        c.rng_seed_offset = cfg["seed_offset"]
        c.pbrt3_scene = cfg["pbrt3"]
        self.d = d

    def desc_ptr(self):
        return C.pointer(self.d)

    @property
    def desc(self):
        return self.d


def parallel_light_scene(albedo=0.5, Le=1.0, h=1.0, half=1.0, maxdepth=1, spp_side=8, res=16, fov=2.0):
    """Diffuse plane y=0 under a 2*half x 2*half rectangular light at height h facing down; the
    camera looks straight down at the origin from y = h/2 with a narrow field of view.  With
    maxdepth 1 the pixel radiance is albedo * Le * F, F = point-to-parallel-rectangle form factor."""
    plane = {"vertices": [[-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]], "indices": [0, 1, 2, 0, 2, 3]}
    lq = {"vertices": [[-half, h, -half], [half, h, -half], [half, h, half], [-half, h, half]],
          "indices": [0, 1, 2, 0, 2, 3]}
    # light frame: x, y in the plane, z = normal pointing down (-y)
    l2w = [1, 0, 0, 0,
           0, 0, -1, h,
           0, 1, 0, 0]
    # camera at (0, h/2, 0) looking down -y: columns right=(1,0,0), up=(0,0,-1), back=(0,1,0)
    c2w = [1, 0, 0, 0,
           0, 0, 1, h / 2,
           0, -1, 0, 0]
    return DescScene(
        meshes=[plane, lq],
        materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [albedo] * 3}],
        objects=[{"mesh": 0, "material": 0}, {"mesh": 1, "light": 0}],
        rects=[{"l2w": l2w, "L": [Le] * 3, "size": [2 * half, 2 * half]}],
        camera={"c2w": c2w, "fov": fov},
        config={"film_width": res, "film_height": res, "region": (0, 0, res, res), "nx": spp_side, "ny": spp_side,
                "maxdepth": maxdepth})


def parallel_rect_form_factor(half, h):
    """Form factor from a differential area to a parallel 2*half x 2*half rectangle centred at
    height h above it: four corner rectangles, F_corner = 1/(2 pi) [A/sqrt(1+A^2) atan(B/sqrt(1+A^2))
    + B/sqrt(1+B^2) atan(A/sqrt(1+B^2))] with A = B = half/h."""
    A = B = half / h
    f = (A / np.sqrt(1 + A * A) * np.arctan(B / np.sqrt(1 + A * A)) +
         B / np.sqrt(1 + B * B) * np.arctan(A / np.sqrt(1 + B * B))) / (2 * np.pi)
    return 4 * f


def regression_mesh_scene(index):
    import json
    import os
    data = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kdtree_regression_meshes.json")))
    m = data["meshes"][index]
    lq = {"vertices": [[-0.1, 40, -0.1], [0.1, 40, -0.1], [0.1, 40, 0.1], [-0.1, 40, 0.1]], "indices": [0, 1, 2, 0, 2, 3]}
    return DescScene(meshes=[{"vertices": m["vertices"], "indices": m["indices"]}, lq],
                     materials=[{"reflectance": [0.5, 0.5, 0.5]}],
                     objects=[{"mesh": 0, "material": 0}, {"mesh": 1, "light": 0}],
                     rects=[{"l2w": [1, 0, 0, 0, 0, 0, -1, 40, 0, 1, 0, 0], "L": [1, 1, 1], "size": [0.2, 0.2]}])


def _quad(a, b, c, d):
    return {"vertices": [a, b, c, d], "indices": [0, 1, 2, 0, 2, 3]}


def indirect_room_scene(res=64, spp_side=4, maxdepth=5, albedo=0.7, Le=50.0):
    """Closed 4 x 3 x 4 diffuse room lit only indirectly: a 0.4 x 0.4 emitter at height 2.6 faces
    the ceiling, so every point below it receives light by way of the bright ceiling patch (next-event
    estimation sees the emitter only from above its plane).  The camera looks at the far wall."""
    W, Hh, D = 4.0, 3.0, 4.0
    walls = [
        _quad([0, 0, 0], [W, 0, 0], [W, 0, D], [0, 0, D]),          # floor
        _quad([0, Hh, 0], [0, Hh, D], [W, Hh, D], [W, Hh, 0]),      # ceiling
        _quad([0, 0, 0], [0, Hh, 0], [W, Hh, 0], [W, 0, 0]),        # z = 0
        _quad([0, 0, D], [W, 0, D], [W, Hh, D], [0, Hh, D]),        # z = D
        _quad([0, 0, 0], [0, 0, D], [0, Hh, D], [0, Hh, 0]),        # x = 0
        _quad([W, 0, 0], [W, Hh, 0], [W, Hh, D], [W, 0, D]),        # x = W
    ]
    h, s = 2.6, 0.2
    light = _quad([2 - s, h, 2 - s], [2 - s, h, 2 + s], [2 + s, h, 2 + s], [2 + s, h, 2 - s])
    # light frame: local x -> world x, local y -> world -z, normal (local z) -> world +y
    l2w = [1, 0, 0, 2,
           0, 0, 1, h,
           0, -1, 0, 2]
    # camera at (2, 1.4, 3.9) looking down -z: columns right (1,0,0), up (0,1,0), back (0,0,1)
    c2w = [1, 0, 0, 2,
           0, 1, 0, 1.4,
           0, 0, 1, 3.9]
    meshes = walls + [light]
    objects = [{"mesh": i, "material": 0} for i in range(len(walls))] + [{"mesh": len(walls), "light": 0}]
    return DescScene(
        meshes=meshes,
        materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [albedo] * 3}],
        objects=objects,
        rects=[{"l2w": l2w, "L": [Le] * 3, "size": [2 * s, 2 * s]}],
        camera={"c2w": c2w, "fov": 70.0},
        config={"film_width": res, "film_height": res, "region": (0, 0, res, res), "nx": spp_side, "ny": spp_side,
                "maxdepth": maxdepth})


def point_light_scene(albedo=0.5, intensity=2.0, h=1.0, spp_side=4, res=16, fov=2.0, maxdepth=1):
    """Diffuse plane y=0 under an isotropic point light at (0, h, 0); the camera looks straight down
    at the origin from y = h/2 with a narrow field of view.  With maxdepth 1 a point at distance r
    from the axis has radiance albedo/pi * I * h / (h^2 + r^2)^(3/2) (direct_lighting.cpp:26-47)."""
    plane = {"vertices": [[-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]], "indices": [0, 1, 2, 0, 2, 3]}
    c2w = [1, 0, 0, 0,
           0, 0, 1, h / 2,
           0, -1, 0, 0]
    return DescScene(
        meshes=[plane],
        materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [albedo] * 3}],
        objects=[{"mesh": 0, "material": 0}],
        points=[{"position": [0.0, h, 0.0], "intensity": [intensity] * 3}],
        camera={"c2w": c2w, "fov": fov},
        config={"film_width": res, "film_height": res, "region": (0, 0, res, res), "nx": spp_side, "ny": spp_side,
                "maxdepth": maxdepth})


def lit_plane_scene(albedo=0.5, spots=(), dirs=(), emitters=(), spp_side=4, res=16, fov=2.0, h=1.0, maxdepth=1):
    """Diffuse plane y=0 seen straight down from (0, h/2, 0) with a narrow field of view, lit by spot and
    directional lights and by emitter meshes given as (vertices, indices, L) (triangle-mesh lights:
    object index = 1 + their position in the list, area light = rect count (0) + that position)."""
    plane = {"vertices": [[-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]], "indices": [0, 1, 2, 0, 2, 3]}
    meshes = [plane] + [{"vertices": v, "indices": ix} for v, ix, _ in emitters]
    objects = [{"mesh": 0, "material": 0}] + [{"mesh": 1 + i, "light": i} for i in range(len(emitters))]
    c2w = [1, 0, 0, 0,
           0, 0, 1, h / 2,
           0, -1, 0, 0]
    return DescScene(
        meshes=meshes,
        materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [albedo] * 3}],
        objects=objects,
        spots=list(spots), dirs=list(dirs),
        mesh_lights=[{"object": 1 + i, "L": L} for i, (_, _, L) in enumerate(emitters)],
        camera={"c2w": c2w, "fov": fov},
        config={"film_width": res, "film_height": res, "region": (0, 0, res, res), "nx": spp_side, "ny": spp_side,
                "maxdepth": maxdepth})


def square_fan(half, y, facing_down=True):
    """A 2*half square at height y as four triangles around its centre (not a rectangle light: a
    triangle-mesh emitter); faces down by default."""
    c = [0.0, y, 0.0]
    corners = [[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]]
    v = [c] + corners
    idx = []
    for i in range(4):
        a, b = 1 + i, 1 + (i + 1) % 4
        idx += [0, a, b] if facing_down else [0, b, a]
    return v, idx

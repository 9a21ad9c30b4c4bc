"""Pins of oracle functions that no reference fixture covers, by what the mathematics fixes:
Snell's law, total internal reflection and Fresnel frequencies for smooth glass
(delta_scattering.cpp:53-100), the eta^2 radiance scaling through an equilibrium (furnace) argument,
adjust_shading_normal (shading_context.cpp:17-63) against its invariants and a closed-form planar case,
the pinhole camera (camera.cpp:16-32) against its geometry, the box-filter footprint (film.cpp:27-50)
against the inclusive-support definition, and the SAH split choice (kdtree_builder.cpp:322-423)
against brute force over every candidate plane.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import yar_amd
from tests import oracle_lib as O

F3 = C.c_float * 3


def _L():
    L = O.lib()
    P = C.c_void_p
    L.oref_glass_scatter.argtypes = [P, P, C.c_int32, C.c_float, P, P, C.c_float, P, P]
    L.oref_glass_scatter.restype = C.c_int32
    L.oref_adjust_normal.argtypes = [P, P, P, P]
    L.oref_adjust_normal.restype = C.c_int32
    L.oref_camera_ray.argtypes = [C.POINTER(yar_amd.View), C.c_float, C.c_float, P, P]
    L.oref_film_footprint.argtypes = [C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float, P]
    return L


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def glass(n, wo, flipped, ior, u, kr=(1, 1, 1), kt=(1, 1, 1)):
    n, wo, kr, kt = _f32(n), _f32(wo), _f32(kr), _f32(kt)
    d = np.zeros(3, np.float32)
    att = np.zeros(3, np.float32)
    r = _L().oref_glass_scatter(n.ctypes.data, wo.ctypes.data, flipped, ior, kr.ctypes.data, kt.ctypes.data, u,
                                d.ctypes.data, att.ctypes.data)
    return bool(r), d.astype(np.float64), att.astype(np.float64)


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _textbook_fresnel(cos_i, n1, n2):
    """Unpolarised Fresnel reflectance for light going from index n1 into n2 (float64)."""
    sin_t = n1 / n2 * np.sqrt(np.maximum(0.0, 1 - cos_i ** 2))
    if sin_t >= 1:
        return 1.0
    cos_t = np.sqrt(1 - sin_t ** 2)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("ior", [1.5, 1.33])
def test_glass_refraction_obeys_snell(ior):
    rng = np.random.default_rng(5)
    ns, wos = _unit(rng, 400), _unit(rng, 400)
    checked = 0
    for n, wo in zip(ns, wos):
        if wo @ n < 0:
            wo = -wo
        for flipped in (0, 1):
            n1, n2 = (1.0, ior) if not flipped else (ior, 1.0)        # wo's side first
            sin_i = np.linalg.norm(np.cross(wo, n))
            if n1 / n2 * sin_i >= 0.999:                               # total internal reflection
                continue
            refl, d, att = glass(n, wo, flipped, ior, 0.9999999, kt=(0.5, 0.7, 0.9))
            assert not refl
            assert abs(np.linalg.norm(d) - 1) < 2e-6                   # unit direction
            assert d @ n < 0                                           # transmitted to the far side
            assert abs(d @ np.cross(wo, n)) < 2e-6                     # plane of incidence
            sin_t = np.linalg.norm(np.cross(d, n))
            assert abs(n1 * sin_i - n2 * sin_t) < 5e-6                 # Snell
            # radiance crossing into index n2 (towards the camera: from n2 into n1) scales by (n1/n2)^2
            assert np.allclose(att, np.array([0.5, 0.7, 0.9]) * (n1 / n2) ** 2, rtol=1e-6)
            checked += 1
    assert checked > 500


def test_glass_total_internal_reflection():
    rng = np.random.default_rng(6)
    ior = 1.5
    crit = np.arcsin(1 / ior)
    for _ in range(300):
        n = _unit(rng, 1)[0]
        theta = rng.uniform(crit + 1e-3, np.pi / 2 - 1e-3)
        t = np.cross(n, _unit(rng, 1)[0])
        t /= np.linalg.norm(t)
        wo = np.cos(theta) * n + np.sin(theta) * t
        for u in (0.0, 0.5, 0.9999999):
            refl, d, att = glass(n, wo, 1, ior, u, kr=(0.3, 0.6, 0.9))
            assert refl
            mirror = 2 * (wo @ n) * n - wo
            assert np.allclose(d, mirror, atol=2e-6)
            assert np.allclose(att, [0.3, 0.6, 0.9])


@pytest.mark.parametrize("flipped", [0, 1])
def test_glass_reflection_frequency_is_fresnel(flipped):
    ior = 1.5
    n = np.array([0.0, 0.0, 1.0])
    rng = np.random.default_rng(7 + flipped)
    for theta in (0.0, 0.4, 0.7, 1.0, 1.3, 1.5):
        wo = np.array([np.sin(theta), 0.0, np.cos(theta)])
        n1, n2 = (1.0, ior) if not flipped else (ior, 1.0)
        F = _textbook_fresnel(np.cos(theta), n1, n2)
        us = rng.uniform(size=4000).astype(np.float32)
        frac = np.mean([glass(n, wo, flipped, ior, float(u))[0] for u in us])
        sigma = np.sqrt(max(F * (1 - F), 1e-4) / us.size)
        assert abs(frac - F) < 5 * sigma + 1e-3, (theta, frac, F)


def _icosphere(level):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v), np.array(f)


def _furnace_scene(tmp_path, eye, look, name):
    """A closed glass icosphere (radius 1, outward winding, Kr = Kt = 1) inside a cube of six
    rectangular emitters of radiance 1 (half size 3)."""
    v, f = _icosphere(3)
    P = " ".join(f"{x:.8f}" for x in v.ravel())
    I = " ".join(str(i) for i in f.ravel())
    s = 3.0
    quads = []
    for axis in range(3):
        for sign in (-1, 1):
            c = np.zeros(3)
            c[axis] = sign * s
            a, b = np.eye(3)[(axis + 1) % 3] * s, np.eye(3)[(axis + 2) % 3] * s
            pts = [c - a - b, c + a - b, c + a + b, c - a + b]
            if sign > 0:
                pts = pts[::-1]                                     # emit towards the centre
            quads.append(pts)
    body = f'LookAt {eye[0]} {eye[1]} {eye[2]}  {look[0]} {look[1]} {look[2]}  0 1 0\n'
    body += 'Camera "perspective" "float fov" [ 50 ]\nFilm "image" "integer xresolution" [ 24 ] "integer yresolution" [ 24 ]\n'
    body += 'Sampler "stratified" "integer xsamples" [ 2 ] "integer ysamples" [ 2 ]\n'
    body += 'Integrator "path" "integer maxdepth" [ 16 ]\nWorldBegin\n'
    body += f'AttributeBegin\nMaterial "glass" "float index" [ 1.5 ]\nShape "trianglemesh" "integer indices" [ {I} ] "point P" [ {P} ]\nAttributeEnd\n'
    for q in quads:
        pts = " ".join(f"{x:.6f}" for p in q for x in p)
        body += ('AttributeBegin\nAreaLightSource "diffuse" "rgb L" [ 1 1 1 ]\n'
                 f'Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ {pts} ]\nAttributeEnd\n')
    body += "WorldEnd\n"
    p = tmp_path / name
    p.write_text(body)
    return yar_amd.Scene(str(p))


def test_glass_furnace_radiance_scales_with_eta_squared(tmp_path):
    # Equilibrium radiance inside a lossless dielectric of index eta surrounded by a uniform emitter L
    # is eta^2 L (radiance / eta^2 is invariant), and back outside it is L again.
    inside = _furnace_scene(tmp_path, (0, 0, 0), (0.3, 0.1, -1), "in.pbrt")
    ctx = O.OracleScene(inside, threads=4)
    v = yar_amd.View()
    d = inside.desc
    for name, _ in yar_amd.View._fields_:
        pass
    v.camera_to_world[:] = d.camera.camera_to_world[:]
    v.fov_y_degrees = d.camera.fov_y_degrees
    v.z_is_up = d.camera.z_is_up
    v.film_width = v.film_height = 24
    v.region_x1 = v.region_y1 = 24
    v.x_pixel_sample_count = v.y_pixel_sample_count = 2
    img, _, _, _, _ = ctx.render(v)
    # every path refracts out exactly once (reflections keep throughput 1): each pixel is 2.25 L
    assert np.allclose(img, 2.25, rtol=1e-5), (img.min(), img.max())

    outside = _furnace_scene(tmp_path, (0, 0, 2.5), (0, 0, 0), "out.pbrt")
    ctx = O.OracleScene(outside, threads=4)
    v.camera_to_world[:] = outside.desc.camera.camera_to_world[:]
    v.x_pixel_sample_count = v.y_pixel_sample_count = 8
    img, _, _, ps, _ = ctx.render(v, per_sample=True)
    # enter (x 1/2.25) and leave (x 2.25): radiance L; Russian roulette keeps the mean
    assert np.median(ps) == pytest.approx(1.0, rel=1e-5)
    assert abs(ps.mean() - 1.0) < 0.02, ps.mean()


def adjust(wo, ng, n):
    wo, ng, n = _f32(wo), _f32(ng), _f32(n)
    out = np.zeros(3, np.float32)
    ch = _L().oref_adjust_normal(wo.ctypes.data, ng.ctypes.data, n.ctypes.data, out.ctypes.data)
    return bool(ch), out.astype(np.float64)


def test_adjust_shading_normal_identity_and_grazing():
    rng = np.random.default_rng(8)
    ng = np.array([0.0, 0.0, 1.0])
    unchanged = 0
    for _ in range(500):
        wo = _unit(rng, 1)[0]
        wo[2] = abs(wo[2]) + 0.05
        wo /= np.linalg.norm(wo)
        n = ng + 0.2 * rng.normal(size=3)
        n[2] = abs(n[2])
        n /= np.linalg.norm(n)
        r = 2 * (wo @ n) * n - wo
        if r @ ng >= 1e-4:
            ch, out = adjust(wo, ng, n)
            assert not ch and np.array_equal(out.astype(np.float32), n.astype(np.float32))
            unchanged += 1
    assert unchanged > 100
    # below the grazing threshold (wo . ng < 0.0017453) the geometric normal is used
    wo = np.array([np.cos(0.001), 0.0, np.sin(0.001)])
    n = np.array([-0.6, 0.0, 0.8])
    ch, out = adjust(wo, ng, n)
    assert ch and np.array_equal(out, ng)


def test_adjust_shading_normal_closed_form_plane():
    # planar case: ng = z, wo at angle theta from z towards +x, n tilted by alpha towards -x so that the
    # mirror of wo about n (at angle -(theta + 2 alpha)) falls below the horizon.  The adjusted normal is
    # the half vector of wo and the horizon direction -x lifted by 1e-4 ng.
    ng = np.array([0.0, 0.0, 1.0])
    for theta in (0.2, 0.5, 0.9, 1.3):
        for alpha in (0.5, 0.7, 1.0):
            if theta + 2 * alpha <= np.pi / 2 + 1e-3:
                continue
            wo = np.array([np.sin(theta), 0.0, np.cos(theta)])
            n = np.array([-np.sin(alpha), 0.0, np.cos(alpha)])
            ch, out = adjust(wo, ng, n)
            h = wo + np.array([-1.0, 0.0, 1e-4])
            assert ch and np.allclose(out, h / np.linalg.norm(h), atol=2e-6), (theta, alpha, out)


def test_adjust_shading_normal_keeps_reflection_above_surface():
    rng = np.random.default_rng(9)
    changed = 0
    for _ in range(2000):
        ng = _unit(rng, 1)[0]
        wo = _unit(rng, 1)[0]
        if wo @ ng < 0.01:
            wo = wo - 2 * (wo @ ng) * ng + 0.02 * ng
            wo /= np.linalg.norm(wo)
        n = ng + 0.8 * rng.normal(size=3)
        if n @ ng < 0.05:
            n = n - 2 * (n @ ng) * ng + 0.05 * ng
        n /= np.linalg.norm(n)
        ch, out = adjust(wo, ng, n)
        assert abs(np.linalg.norm(out) - 1) < 1e-5
        r = 2 * (wo @ out) * out - wo
        assert r @ ng > -1e-5
        changed += ch
    assert changed > 200


def _view(c2w, fov, w, h, zup=0):
    v = yar_amd.View()
    v.camera_to_world[:] = list(np.asarray(c2w, np.float32).ravel())
    v.fov_y_degrees = fov
    v.z_is_up = zup
    v.film_width, v.film_height = w, h
    v.region_x1, v.region_y1 = w, h
    v.x_pixel_sample_count = v.y_pixel_sample_count = 1
    return v


def camray(v, fx, fy):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    _L().oref_camera_ray(C.byref(v), fx, fy, o.ctypes.data, d.ctypes.data)
    return o.astype(np.float64), d.astype(np.float64)


@pytest.mark.parametrize("zup", [0, 1])
def test_pinhole_camera_geometry(zup):
    rng = np.random.default_rng(10)
    a = rng.normal(size=(3, 3))
    q, _ = np.linalg.qr(a)
    q *= np.sign(np.linalg.det(q))                                  # rotation
    pos = np.array([1.5, -2.0, 3.0])
    c2w = np.concatenate([q, pos[:, None]], 1)
    fov, w, h = 60.0, 200, 100
    v = _view(c2w, fov, w, h, zup)
    fwd = q @ (np.array([0, 1, 0]) if zup else np.array([0, 0, -1]))
    up = q @ (np.array([0, 0, 1]) if zup else np.array([0, 1, 0]))
    right = np.cross(fwd, up)
    o, d = camray(v, w / 2, h / 2)
    assert np.allclose(o, pos, atol=1e-6) and np.allclose(d, fwd, atol=1e-6)
    t = np.tan(np.radians(fov / 2))
    # top edge centre: the vertical half field of view; right edge centre: atan(aspect tan(fov/2))
    _, d = camray(v, w / 2, 0)
    assert np.degrees(np.arccos(d @ fwd)) == pytest.approx(fov / 2, abs=1e-4) and d @ up > 0
    _, d = camray(v, w, h / 2)
    assert np.arccos(d @ fwd) == pytest.approx(np.arctan(t * w / h), abs=2e-6) and d @ right > 0
    # any film point: direction proportional to fwd + (2 fx/w - 1) t aspect right - (2 fy/h - 1) t up
    for fx, fy in rng.uniform([0, 0], [w, h], size=(50, 2)):
        _, d = camray(v, fx, fy)
        e = fwd + (2 * fx / w - 1) * t * (w / h) * right - (2 * fy / h - 1) * t * up
        assert np.allclose(d, e / np.linalg.norm(e), atol=2e-6)


def footprint(x, y, jx, jy, r=0.5):
    out = np.zeros(4, np.int32)
    _L().oref_film_footprint(x, y, jx, jy, r, out.ctypes.data)
    return out


def test_film_footprint_is_inclusive_box_support():
    rng = np.random.default_rng(11)
    one_minus = np.float32(1) - np.float32(2 ** -24)
    cases = [(10, 20, 0.0, 0.0), (10, 20, 0.5, 0.5), (0, 0, 0.0, 0.25), (7, 3, float(one_minus), 0.5),
             ((1 << 20) + 3, 5, float(one_minus), 0.0)]
    cases += [(int(x), int(y), float(jx), float(jy)) for x, y, jx, jy in
              zip(rng.integers(0, 4096, 300), rng.integers(0, 4096, 300),
                  rng.uniform(0, 1, 300).astype(np.float32), rng.uniform(0, 1, 300).astype(np.float32))]
    for x, y, jx, jy in cases:
        fx = float(np.float32(x) + np.float32(jx))                    # film_pos = pixel + jitter in float32
        fy = float(np.float32(y) + np.float32(jy))
        px = [p for p in range(x - 2, x + 3) if abs(p + 0.5 - fx) <= 0.5]
        py = [p for p in range(y - 2, y + 3) if abs(p + 0.5 - fy) <= 0.5]
        got = footprint(x, y, jx, jy)
        assert (got[0], got[2], got[1], got[3]) == (px[0], px[-1] + 1, py[0], py[-1] + 1), (x, y, jx, jy, got)
    # a zero jitter puts the sample on the corner of four pixel squares: it reaches the left/upper pixels
    assert list(footprint(10, 20, 0.0, 0.0)) == [9, 19, 11, 21]
    # a jitter that rounds the float sum up to the next integer reaches the pixel to the right only
    assert list(footprint((1 << 20) + 3, 5, float(one_minus), 0.5))[:3:2] == [(1 << 20) + 3, (1 << 20) + 5]


def _brute_root_split(lo, hi):
    """Best SAH plane over all candidate positions (float64 costs) and its margin to the runner-up."""
    n = lo.shape[0]
    blo, bhi = lo.min(0), hi.max(0)
    diag = bhi - blo
    total = 2 * (diag[0] * diag[1] + diag[0] * diag[2] + diag[1] * diag[2])
    cands = []
    for axis in range(3):
        o0, o1 = [a for a in range(3) if a != axis]
        for t in np.unique(np.concatenate([lo[:, axis], hi[:, axis]])):
            if not (blo[axis] < t < bhi[axis]):
                continue
            nb = int((lo[:, axis] < t).sum())                       # reach below the plane
            na = int((hi[:, axis] > t).sum())                       # reach above the plane
            ab = 2 * diag[o0] * diag[o1] + 2 * (diag[o0] + diag[o1]) * (t - blo[axis])
            aa = 2 * diag[o0] * diag[o1] + 2 * (diag[o0] + diag[o1]) * (bhi[axis] - t)
            bonus = 0.3 if (nb == 0 or na == 0) else 0.0
            cands.append(((1 - bonus) * (ab / total * nb + aa / total * na), axis, float(t)))
    cands.sort(key=lambda c: c[0])
    return cands, n


def test_sah_root_split_matches_brute_force():
    rng = np.random.default_rng(12)
    checked = leaves = 0
    for trial in range(300):
        n = int(rng.integers(3, 14))
        lo = rng.uniform(0, 10, (n, 3)).astype(np.float32)
        hi = (lo + rng.uniform(0, 3, (n, 3))).astype(np.float32)
        if trial % 3 == 0:
            k = rng.integers(0, 3)
            hi[::2, k] = lo[::2, k]                                   # some planar boxes
        words, _, _ = O.kdtree_boxes(np.concatenate([lo, hi], 1))
        prod = yar_amd.build_box_kdtree(np.concatenate([lo, hi], 1)).words
        assert np.array_equal(words, prod)
        cands, n = _brute_root_split(lo.astype(np.float64), hi.astype(np.float64))
        w0, w1 = int(words[0]), int(words[1])
        best = cands[0] if cands else None
        if best is None or best[0] >= n * (1 - 1e-6):
            if best is None or best[0] > n * (1 + 1e-6):
                assert (w0 & 3) == 3                                   # no plane beats the leaf cost N
                leaves += 1
            continue
        if len(cands) > 1 and cands[1][0] - best[0] < 1e-5 * best[0]:
            continue                                                  # near tie: float order decides
        assert (w0 & 3) == best[1], (trial, best, w0 & 3)
        assert np.array([w1], np.uint32).view(np.float32)[0] == np.float32(best[2])
        checked += 1
    assert checked > 150 and leaves >= 0


def plastic(pbrt3, kd, r0, alpha, wo, wi=None, u=None, flipped=0):
    L = O.lib()
    P = C.c_void_p
    L.oref_plastic_eval.argtypes = [C.c_int32, P, C.c_float, C.c_float, C.c_int32, P, P, P, P]
    L.oref_plastic_sample.argtypes = [C.c_int32, P, C.c_float, C.c_float, C.c_int32, P, P, P, P, P]
    L.oref_plastic_sample.restype = C.c_int32
    kd, wo = _f32(kd), _f32(wo)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    if wi is not None:
        wi = _f32(wi)
        L.oref_plastic_eval(pbrt3, kd.ctypes.data, r0, alpha, flipped, wo.ctypes.data, wi.ctypes.data, f.ctypes.data,
                            C.byref(pdf))
        return f.astype(np.float64), pdf.value
    u = _f32(u)
    out = np.zeros(3, np.float32)
    ok = L.oref_plastic_sample(pbrt3, kd.ctypes.data, r0, alpha, flipped, wo.ctypes.data, u.ctypes.data,
                               out.ctypes.data, f.ctypes.data, C.byref(pdf))
    return bool(ok), out.astype(np.float64), f.astype(np.float64), pdf.value


def _hemi(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


@pytest.mark.parametrize("pbrt3", [0, 1])
def test_plastic_closed_forms(pbrt3):
    # r0 = 0: the diffuse lobe alone, Kd / pi.  Normal incidence: both Fresnel forms give
    # ((1.5 - 1) / (1.5 + 1))^2 = 0.04, G = 1 and D(n) = 1 / (pi alpha^2), so the specular term is
    # r0 * 0.04 / (4 pi alpha^2).
    kd, n = [0.3, 0.5, 0.7], [0, 0, 1]
    f, _ = plastic(pbrt3, kd, 0.0, 0.2, _hemi(0.7, 0.3), _hemi(0.4, 2.0))
    assert np.allclose(f, np.array(kd) / np.pi, rtol=1e-6)
    for alpha in (0.1, 0.3):
        f, pdf = plastic(pbrt3, kd, 0.8, alpha, n, n)
        assert np.allclose(f, np.array(kd) / np.pi + 0.8 * 0.04 / (4 * np.pi * alpha ** 2), rtol=2e-5)
        # pdf = (cos / pi + G1 D (wo.wh) / (wo.n) / (4 wo.wh)) / 2 with wh = n
        assert pdf == pytest.approx(0.5 * (1 / np.pi + 1 / (np.pi * alpha ** 2) / 4), rel=2e-5)


@pytest.mark.parametrize("pbrt3,alpha", [(1, 0.15), (0, 0.15), (1, 0.4)])
def test_plastic_sampling_matches_quadrature_and_reciprocity(pbrt3, alpha):
    rng = np.random.default_rng(13)
    kd, r0 = [0.4, 0.3, 0.2], 0.6
    for theta_o in (0.3, 1.0):
        wo = _hemi(theta_o, 0.4)
        # reciprocity f(wo, wi) = f(wi, wo)
        for _ in range(20):
            wi = _hemi(rng.uniform(0, 1.5), rng.uniform(0, 2 * np.pi))
            f1, _ = plastic(pbrt3, kd, r0, alpha, wo, wi)
            f2, _ = plastic(pbrt3, kd, r0, alpha, wi, wo)
            assert np.allclose(f1, f2, rtol=1e-5)
        # quadrature of the albedo int f cos and of the pdf over the hemisphere (midpoint rule in
        # (cos theta, phi)) against the importance-sampled estimate and the fraction of valid samples
        nt, nphi = 400, 200
        ct = (np.arange(nt) + 0.5) / nt
        ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
        alb, mass = 0.0, 0.0
        for c in ct:
            st = np.sqrt(1 - c * c)
            for p in ph[::4]:
                wi = np.array([st * np.cos(p), st * np.sin(p), c])
                f, pdf = plastic(pbrt3, kd, r0, alpha, wo, wi)
                alb += f[1] * c
                mass += pdf
        dA = (1.0 / nt) * (2 * np.pi / (nphi / 4))
        alb *= dA
        mass *= dA
        n = 6000
        est, valid = 0.0, 0
        for u in rng.uniform(size=(n, 3)):
            ok, wi, f, pdf = plastic(pbrt3, kd, r0, alpha, wo, u=u)
            if ok:
                est += f[1] * wi[2] / pdf
                valid += 1
        est /= n
        assert est == pytest.approx(alb, rel=0.03), (theta_o, est, alb)
        assert mass == pytest.approx(valid / n, abs=0.02), (theta_o, mass, valid / n)


def test_mirror_reflects_point_lit_plane_closed_form():
    # Camera at (0, 1, 0) looking along +x at a 45-degree mirror centred at (1, 1, 0) (normal
    # (-1, -1, 0)/sqrt 2): every camera ray turns straight down onto the diffuse plane y = 0 near
    # (1, 0, 0), lit by a point light at (4, 2, 0) that the mirror does not shadow.  With maxdepth 2
    # (the delta bounce counts, delta_scattering.cpp:27-33, path_tracing.cpp:60-84) the radiance is
    # Kr * albedo/pi * I * cos / d^2 with cos = 2 / sqrt 13, d^2 = 13.
    from tests.scenes import DescScene
    from tests.test_oracle_fixtures import _view_from_desc
    from yar_amd import _abi
    s = 0.3 / np.sqrt(2)
    c = np.array([1.0, 1.0, 0.0])
    e1, e2 = np.array([0, 0, 0.3]), np.array([s, -s, 0])
    quad = {"vertices": [c - e1 - e2, c + e1 - e2, c + e1 + e2, c - e1 + e2], "indices": [0, 1, 2, 0, 2, 3]}
    plane = {"vertices": [[-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]], "indices": [0, 1, 2, 0, 2, 3]}
    kr, albedo, intensity = [0.9, 0.6, 0.3], 0.5, 2.0
    sc = DescScene(
        meshes=[plane, quad],
        materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [albedo] * 3},
                   {"type": _abi.YAR_MATERIAL_MIRROR, "reflectance": kr}],
        objects=[{"mesh": 0, "material": 0}, {"mesh": 1, "material": 1}],
        points=[{"position": [4.0, 2.0, 0.0], "intensity": [intensity] * 3}],
        camera={"c2w": [0, 0, -1, 0, 0, 1, 0, 1, 1, 0, 0, 0], "fov": 1.0},
        config={"film_width": 16, "film_height": 16, "region": (0, 0, 16, 16), "nx": 2, "ny": 2, "maxdepth": 2})
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    want = np.array(kr) * albedo / np.pi * intensity * 2.0 / 13.0 ** 1.5
    # +-0.5 degree of view moves the floor point by +-0.017: +-1.2% across the image, linear, so the
    # image mean keeps only the quadratic residue
    assert np.allclose(img, want, rtol=2e-2), (img.min(axis=(0, 1)), img.max(axis=(0, 1)), want)
    assert np.allclose(img.mean(axis=(0, 1)), want, rtol=1e-3)
    # the mirror tints by Kr, channel by channel
    assert np.allclose(img[..., 0] / img[..., 1], kr[0] / kr[1], rtol=1e-5)
    # maxdepth 1 ends the path at the delta bounce: nothing reaches the camera
    sc.desc.config.max_light_bounces = 1
    img1, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    assert (img1 == 0).all()


def coated(kd, r0, ax, ay, wo, wi=None, u=None):
    L = O.lib()
    P = C.c_void_p
    L.oref_coated_eval.argtypes = [P, P, C.c_float, C.c_float, P, P, P, P]
    L.oref_coated_sample.argtypes = [P, P, C.c_float, C.c_float, P, P, P, P, P]
    L.oref_coated_sample.restype = C.c_int32
    kd, r0, wo = _f32(kd), _f32(r0), _f32(wo)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    if wi is not None:
        wi = _f32(wi)
        L.oref_coated_eval(kd.ctypes.data, r0.ctypes.data, ax, ay, wo.ctypes.data, wi.ctypes.data, f.ctypes.data,
                           C.byref(pdf))
        return f.astype(np.float64), pdf.value
    u = _f32(u)
    out = np.zeros(3, np.float32)
    ok = L.oref_coated_sample(kd.ctypes.data, r0.ctypes.data, ax, ay, wo.ctypes.data, u.ctypes.data, out.ctypes.data,
                              f.ctypes.data, C.byref(pdf))
    return bool(ok), out.astype(np.float64), f.astype(np.float64), pdf.value


def test_coated_diffuse_closed_forms():
    # Ashikhmin-Shirley coated diffuse (bsdf.cpp:440-495).  Normal incidence: F = r0, anisotropic
    # GGX D(n) = 1/(pi ax ay), so f = Rd (1 - r0) 28/(23 pi) (31/32)^2 + r0 / (4 pi ax ay).
    kd, r0 = np.array([0.3, 0.5, 0.7]), np.array([0.05, 0.1, 0.2])
    for ax, ay in ((0.1, 0.1), (0.08, 0.3)):
        f, pdf = coated(kd, r0, ax, ay, [0, 0, 1], [0, 0, 1])
        want = kd * (1 - r0) * 28 / (23 * np.pi) * (31 / 32) ** 2 + r0 / (4 * np.pi * ax * ay)
        assert np.allclose(f, want, rtol=2e-5), (f, want)
        # pdf = (cos/pi + VNDF reflection pdf) / 2, at normal incidence G1 = 1: D(n) / 4
        assert pdf == pytest.approx(0.5 * (1 / np.pi + 1 / (4 * np.pi * ax * ay)), rel=2e-5)
    # r0 = 1: no diffuse term, F = 1, and the specular term is the textbook anisotropic GGX
    # D(h) = 1 / (pi ax ay (hx^2/ax^2 + hy^2/ay^2 + hz^2)^2) over 4 (wi.h) max(wo.z, wi.z)
    ax, ay = 0.15, 0.4
    rng = np.random.default_rng(5)
    for _ in range(20):
        wo = _hemi(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi))
        wi = _hemi(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi))
        h = (wo + wi) / np.linalg.norm(wo + wi)
        D = 1 / (np.pi * ax * ay * (h[0] ** 2 / ax ** 2 + h[1] ** 2 / ay ** 2 + h[2] ** 2) ** 2)
        want = D / (4 * wi.dot(h) * max(wo[2], wi[2]))
        f, _ = coated(kd, [1, 1, 1], ax, ay, wo, wi)
        assert np.allclose(f, want, rtol=1e-4), (f, want)


@pytest.mark.parametrize("ax,ay", [(0.15, 0.15), (0.1, 0.35)])
def test_coated_diffuse_sampling_matches_quadrature_and_reciprocity(ax, ay):
    rng = np.random.default_rng(17)
    kd, r0 = [0.4, 0.3, 0.2], [0.3, 0.3, 0.3]
    for theta_o in (0.3, 1.0):
        wo = _hemi(theta_o, 0.7)
        for _ in range(20):
            wi = _hemi(rng.uniform(0, 1.5), rng.uniform(0, 2 * np.pi))
            f1, _ = coated(kd, r0, ax, ay, wo, wi)
            f2, _ = coated(kd, r0, ax, ay, wi, wo)
            assert np.allclose(f1, f2, rtol=1e-5)
        nt, nphi = 400, 200
        ct = (np.arange(nt) + 0.5) / nt
        ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
        alb, mass = 0.0, 0.0
        for c in ct:
            st = np.sqrt(1 - c * c)
            for p in ph[::2]:
                wi = np.array([st * np.cos(p), st * np.sin(p), c])
                f, pdf = coated(kd, r0, ax, ay, wo, wi)
                alb += f[1] * c
                mass += pdf
        dA = (1.0 / nt) * (2 * np.pi / (nphi / 2))
        alb *= dA
        mass *= dA
        n = 6000
        est, valid = 0.0, 0
        for u in rng.uniform(size=(n, 3)):
            ok, wi, f, pdf = coated(kd, r0, ax, ay, wo, u=u)
            if ok:
                est += f[1] * wi[2] / pdf
                valid += 1
        est /= n
        assert est == pytest.approx(alb, rel=0.03), (theta_o, est, alb)
        assert mass == pytest.approx(valid / n, abs=0.02), (theta_o, mass, valid / n)


def rough_glass(ior, alpha, wo, wi=None, u=None, flipped=0, kr=(1, 1, 1), kt=(1, 1, 1)):
    L = O.lib()
    P = C.c_void_p
    L.oref_rough_glass_eval.argtypes = [C.c_float, C.c_float, P, P, C.c_int32, P, P, P, P]
    L.oref_rough_glass_sample.argtypes = [C.c_float, C.c_float, P, P, C.c_int32, P, P, P, P, P]
    L.oref_rough_glass_sample.restype = C.c_int32
    kr, kt, wo = _f32(kr), _f32(kt), _f32(wo)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    if wi is not None:
        wi = _f32(wi)
        L.oref_rough_glass_eval(ior, alpha, kr.ctypes.data, kt.ctypes.data, flipped, wo.ctypes.data, wi.ctypes.data,
                                f.ctypes.data, C.byref(pdf))
        return f.astype(np.float64), pdf.value
    u = _f32(u)
    out = np.zeros(3, np.float32)
    ok = L.oref_rough_glass_sample(ior, alpha, kr.ctypes.data, kt.ctypes.data, flipped, wo.ctypes.data, u.ctypes.data,
                                   out.ctypes.data, f.ctypes.data, C.byref(pdf))
    return bool(ok), out.astype(np.float64), f.astype(np.float64), pdf.value


@pytest.mark.parametrize("flipped", [0, 1], ids=["entering", "leaving"])
def test_rough_glass_smooth_limit_is_fresnel_with_radiance_scaling(flipped):
    # As alpha -> 0 the sampled albedo of Rough_Glass_BSDF (Kr = Kt = 1) tends to the smooth
    # interface's F + (1 - F) (eta_o / eta_i)^2 (radiance scaling of the transmitted part), with F
    # the unpolarised Fresnel reflectance, and to 1 past the critical angle.
    ior = 1.5
    eta_o, eta_i = (1.0, ior) if not flipped else (ior, 1.0)
    rng = np.random.default_rng(23)
    for theta in (0.2, 0.6, 1.0, 1.3):
        wo = _hemi(theta, 0.3)
        F = _textbook_fresnel(np.cos(theta), eta_o, eta_i)
        want = F + (1 - F) * (eta_o / eta_i) ** 2
        n = 4000
        est = 0.0
        for u in rng.uniform(size=(n, 3)):
            ok, wi, f, pdf = rough_glass(ior, 0.003, wo, u=u, flipped=flipped)
            if ok:
                est += f[1] * abs(wi[2]) / pdf
        est /= n
        assert est == pytest.approx(want, rel=0.03), (theta, est, want, F)


@pytest.mark.parametrize("flipped,alpha", [(0, 0.2), (1, 0.2), (0, 0.45)])
def test_rough_glass_generalised_reciprocity_and_sampling_bounds(flipped, alpha):
    ior = 1.5
    eta_o, eta_i = (1.0, ior) if not flipped else (ior, 1.0)
    rng = np.random.default_rng(29)
    prev_gap = None
    for theta_o in (0.35, 0.9):
        wo = _hemi(theta_o, 1.1)
        # transmission: f(wo -> wi) seen from the other side is scaled by (eta_i / eta_o)^2
        # (Walter et al. 2007, eq. 21 in the radiance convention); reflection is reciprocal
        for _ in range(20):
            wt = _hemi(rng.uniform(0.1, 1.3), rng.uniform(0, 2 * np.pi)) * np.array([1, 1, -1])
            f1, _ = rough_glass(ior, alpha, wo, wt, flipped=flipped)
            f2, _ = rough_glass(ior, alpha, wt * np.array([1, 1, -1]), wo * np.array([1, 1, -1]), flipped=1 - flipped)
            if f1[1] > 0:
                assert f2[1] / f1[1] == pytest.approx((eta_i / eta_o) ** 2, rel=1e-3)   # float32 Fresnel and half-vector
            wr = _hemi(rng.uniform(0.1, 1.3), rng.uniform(0, 2 * np.pi))
            g1, _ = rough_glass(ior, alpha, wo, wr, flipped=flipped)
            g2, _ = rough_glass(ior, alpha, wr, wo, flipped=flipped)
            assert np.allclose(g1, g2, rtol=1e-5)
        # Rough_Glass_BSDF::sample rejects a microfacet normal whose mirror direction falls below the
        # surface before it chooses reflection or refraction (bsdf.cpp:341-344), so its samples cover
        # less than the support of pdf(): the sampled albedo and the valid-sample fraction are bounded
        # by the quadratures of f|cos| and of pdf over the sphere (equality in the smooth limit, see
        # the test above)
        nt, nphi = 300, 100
        ct = (np.arange(2 * nt) + 0.5) / nt - 1
        ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
        alb, mass = 0.0, 0.0
        for c in ct:
            st = np.sqrt(1 - c * c)
            for p in ph:
                wi = np.array([st * np.cos(p), st * np.sin(p), c])
                f, pdf = rough_glass(ior, alpha, wo, wi, flipped=flipped)
                alb += f[1] * abs(c)
                mass += pdf
        dA = (1.0 / nt) * (2 * np.pi / nphi)
        alb *= dA
        mass *= dA
        n = 6000
        est, valid = 0.0, 0
        for u in rng.uniform(size=(n, 3)):
            ok, wi, f, pdf = rough_glass(ior, alpha, wo, u=u, flipped=flipped)
            if ok:
                est += f[1] * abs(wi[2]) / pdf
                valid += 1
        est /= n
        assert mass <= 1.02
        assert est <= alb * 1.03 and valid / n <= mass + 0.02, (theta_o, est, alb, valid / n, mass)
        assert est > 0.5 * alb


def _render_lit_plane(**kw):
    from tests import scenes
    from tests.test_oracle_fixtures import _view_from_desc
    sc = scenes.lit_plane_scene(**kw)
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    return img


def test_spot_light_cone_and_penumbra_closed_forms():
    # direct_lighting.cpp:49-89: a spot at (0, 1, 0) lights the plane point below it with
    # albedo/pi * I * cos / d^2 inside the cone, k^4 of that in the penumbra with
    # k = (cos w - cos cone) / (cos(cone - delta) - cos cone), nothing outside.
    rad = np.pi / 180
    def spot(tilt_deg, cone=30.0, delta=10.0):
        t = tilt_deg * rad
        return {"position": [0, 1, 0], "direction": [np.sin(t), -np.cos(t), 0], "cone_angle": cone * rad,
                "penumbra_angle": delta * rad, "intensity": [2.0, 2.0, 2.0]}
    full = 0.5 / np.pi * 2.0
    img = _render_lit_plane(spots=[spot(0.0)], fov=0.5)
    assert np.allclose(img, full, rtol=1e-3), (img.min(), img.max(), full)
    img = _render_lit_plane(spots=[spot(45.0)], fov=0.5)
    assert (img == 0).all()
    img = _render_lit_plane(spots=[spot(25.0)], fov=0.2)
    k = (np.cos(25 * rad) - np.cos(30 * rad)) / (np.cos(20 * rad) - np.cos(30 * rad))
    # the view covers +-0.1 degree around the point below the light; k^4 varies by ~3% per 0.1 degree
    assert img.mean() == pytest.approx(full * k ** 4, rel=1e-2), (img.mean(), full * k ** 4)


def test_triangle_mesh_light_seen_directly_is_one_sided():
    # get_emitted_radiance (direct_lighting.cpp:399-408): a camera ray that hits a triangle-mesh
    # light sees Le on its front and black on its back (the shading normal had to be flipped)
    from tests import scenes
    for facing_down, want in ((False, 2.5), (True, 0.0)):
        v, idx = scenes.square_fan(1.0, 0.25, facing_down=facing_down)
        img = _render_lit_plane(emitters=[(v, idx, [2.5, 2.5, 2.5])])
        assert (img == np.float32(want)).all(), (facing_down, img.min(), img.max())


def test_directional_light_closed_form():
    # direct_lighting.cpp:91-107: L = albedo/pi * E * cos(theta) for a light from direction theta
    for theta in (0.0, 0.6, 1.2):
        d = [np.sin(theta), np.cos(theta), 0.0]
        img = _render_lit_plane(dirs=[{"direction": d, "irradiance": [3.0, 2.0, 1.0]}], fov=1.0)
        want = 0.5 / np.pi * np.array([3.0, 2.0, 1.0]) * np.cos(theta)
        assert np.allclose(img, want, rtol=1e-5), (theta, img.mean(axis=(0, 1)), want)


def test_triangle_mesh_light_form_factor_and_one_sidedness():
    # a 2x2 square emitter at height 1 built from four triangles around its centre is a triangle-mesh
    # light (not a rectangle); with MIS of light and BSDF sampling (direct_lighting.cpp:262-322) the
    # plane below it receives albedo * Le * F, F the point-to-parallel-square form factor
    from tests import scenes
    v, idx = scenes.square_fan(1.0, 1.0, facing_down=True)
    img = _render_lit_plane(emitters=[(v, idx, [1.0, 1.0, 1.0])], spp_side=16)
    want = 0.5 * scenes.parallel_rect_form_factor(1.0, 1.0)
    assert img.mean() == pytest.approx(want, rel=0.01), (img.mean(), want)
    # facing up, the plane sees the emitter's back: light samples there have pdf 0 and contribute
    # nothing, while BSDF samples that hit the back keep MIS weight 1 (direct_lighting.cpp:313-320,
    # light_sampling.cpp:175-177), so the BSDF half alone carries the same albedo * Le * F
    v, idx = scenes.square_fan(1.0, 1.0, facing_down=False)
    img = _render_lit_plane(emitters=[(v, idx, [1.0, 1.0, 1.0])], spp_side=16)
    assert img.mean() == pytest.approx(want, rel=0.02), (img.mean(), want)


def _sphere_light_scene(tmp_path, centre, radius, spp_side=16, look_up=False, name="sl.pbrt"):
    """A diffuse plane y = 0 (albedo 0.5) seen straight down from (0, 0.5, 0) through a 0.5-degree
    view (or a view straight up at the light), lit only by a pbrt sphere light of radiance 1."""
    look = "0 2 0  1 0 0" if look_up else "0 0 0  0 0 1"
    body = (f'LookAt 0 0.5 0  {look}\nCamera "perspective" "float fov" [ 0.5 ]\n'
            'Film "image" "integer xresolution" [ 16 ] "integer yresolution" [ 16 ]\n'
            f'Sampler "stratified" "integer xsamples" [ {spp_side} ] "integer ysamples" [ {spp_side} ]\n'
            'Integrator "path" "integer maxdepth" [ 1 ]\nWorldBegin\n'
            'AttributeBegin\nMaterial "matte" "rgb Kd" [ 0.5 0.5 0.5 ]\n'
            'Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ -50 0 -50  -50 0 50  50 0 50  50 0 -50 ]\n'
            'AttributeEnd\n'
            f'AttributeBegin\nTranslate {centre[0]} {centre[1]} {centre[2]}\nAreaLightSource "diffuse" "rgb L" [ 1 1 1 ]\n'
            f'Shape "sphere" "float radius" [ {radius} ]\nAttributeEnd\nWorldEnd\n')
    p = tmp_path / name
    p.write_text(body)
    return yar_amd.Scene(str(p))


def test_sphere_light_irradiance_closed_form(tmp_path):
    # direct_lighting.cpp:196-260, light_sampling.cpp:76-121: a uniformly emitting sphere wholly above
    # a surface gives irradiance pi L (R/d)^2 cos(theta) (theta between the normal and the centre), so
    # the diffuse plane shows albedo * L * R^2 cos(theta) / d^2.  Light samples land on the analytic
    # sphere, BSDF samples count where they hit its level-6 tessellation (inscribed, ~1e-4 smaller
    # in solid angle), and the two MIS halves must sum to that value.
    from tests.test_oracle_fixtures import _view_from_desc
    for centre, cos_t, d2 in (((0, 2, 0), 1.0, 4.0), ((1, 2, 0), 2 / np.sqrt(5), 5.0), ((0.5, 1, 0.3), 1 / np.sqrt(1.34), 1.34)):
        sc = _sphere_light_scene(tmp_path, centre, 0.5)
        img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
        want = 0.5 * 1.0 * 0.25 * cos_t / d2
        assert img.mean() == pytest.approx(want, rel=1e-3), (centre, img.mean(), want)


def test_sphere_light_seen_directly(tmp_path):
    # get_emitted_radiance (direct_lighting.cpp:396-397): the sphere emits its L towards every viewer
    from tests.test_oracle_fixtures import _view_from_desc
    sc = _sphere_light_scene(tmp_path, (0, 2, 0), 0.5, spp_side=2, look_up=True)
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    assert (img == 1.0).all(), (img.min(), img.max())


def uber(kd, ks, opacity, ior, ax, ay, wo, wi=None, u=None, flipped=0):
    L = O.lib()
    P = C.c_void_p
    F = C.c_float
    L.oref_uber_eval.argtypes = [P, P, P, F, F, F, C.c_int32, P, P, P, P]
    L.oref_uber_sample.argtypes = [P, P, P, F, F, F, C.c_int32, P, P, P, P, P]
    L.oref_uber_sample.restype = C.c_int32
    kd, ks, op, wo = _f32(kd), _f32(ks), _f32(opacity), _f32(wo)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    if wi is not None:
        wi = _f32(wi)
        L.oref_uber_eval(kd.ctypes.data, ks.ctypes.data, op.ctypes.data, ior, ax, ay, flipped, wo.ctypes.data,
                         wi.ctypes.data, f.ctypes.data, C.byref(pdf))
        return f.astype(np.float64), pdf.value
    u = _f32(u)
    out = np.zeros(3, np.float32)
    ok = L.oref_uber_sample(kd.ctypes.data, ks.ctypes.data, op.ctypes.data, ior, ax, ay, flipped, wo.ctypes.data,
                            u.ctypes.data, out.ctypes.data, f.ctypes.data, C.byref(pdf))
    return bool(ok), out.astype(np.float64), f.astype(np.float64), pdf.value


def _smith_lambda(v, ax, ay):
    c2 = v[2] ** 2
    s2 = 1 - c2
    a2 = (v[0] ** 2 * ax * ax + v[1] ** 2 * ay * ay) / s2 if s2 > 0 else 0.0
    return 0.5 * (-1 + np.sqrt(1 + a2 * s2 / c2))


def test_uber_closed_forms():
    # Pbrt3_Uber_BRDF (bsdf_pbrt.cpp:14-91): Ks = 0 leaves Kd * opacity / pi; Kd = 0 leaves the
    # textbook anisotropic GGX reflection F(wi.h) G D / (4 cos_o cos_i) of a dielectric of index
    # eta (1/eta when the surface is left), scaled by Ks * opacity; the pdf averages cosine and VNDF.
    rng = np.random.default_rng(9)
    kd, op = np.array([0.3, 0.5, 0.7]), np.array([0.9, 0.6, 0.4])
    for _ in range(5):
        wo = _hemi(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi))
        wi = _hemi(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi))
        f, _ = uber(kd, [0, 0, 0], op, 1.5, 0.2, 0.2, wo, wi)
        assert np.allclose(f, kd * op / np.pi, rtol=1e-6)
    ax, ay = 0.12, 0.35
    for flipped in (0, 1):
        eta = 1.5 if not flipped else 1 / 1.5
        for _ in range(20):
            wo = _hemi(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi))
            wi = _hemi(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi))
            h = (wo + wi) / np.linalg.norm(wo + wi)
            D = 1 / (np.pi * ax * ay * (h[0] ** 2 / ax ** 2 + h[1] ** 2 / ay ** 2 + h[2] ** 2) ** 2)
            G = 1 / (1 + _smith_lambda(wo, ax, ay) + _smith_lambda(wi, ax, ay))
            want = _textbook_fresnel(wi.dot(h), 1.0, eta) * G * D / (4 * wo[2] * wi[2])
            f, _ = uber([0, 0, 0], [1, 1, 1], [1, 1, 1], 1.5, ax, ay, wo, wi, flipped=flipped)
            assert np.allclose(f, want, rtol=2e-4), (flipped, f, want)
            # opacity scales both lobes channel by channel; f is reciprocal
            f2, _ = uber(kd, [1, 1, 1], op, 1.5, ax, ay, wo, wi, flipped=flipped)
            assert np.allclose(f2, op * (kd / np.pi + want), rtol=2e-4)
            f3, _ = uber(kd, [1, 1, 1], op, 1.5, ax, ay, wi, wo, flipped=flipped)
            assert np.allclose(f2, f3, rtol=1e-5)
    # normal incidence: G1 = 1, D(n) = 1/(pi ax ay), so pdf = (1/pi + 1/(4 pi ax ay)) / 2
    _, pdf = uber(kd, [1, 1, 1], op, 1.5, ax, ay, [0, 0, 1], [0, 0, 1])
    assert pdf == pytest.approx(0.5 * (1 / np.pi + 1 / (4 * np.pi * ax * ay)), rel=2e-5)


def test_uber_sampling_matches_quadrature():
    rng = np.random.default_rng(21)
    kd, ks, op = [0.4, 0.3, 0.2], [0.5, 0.5, 0.5], [0.8, 0.8, 0.8]
    ax, ay = 0.1, 0.3
    for theta_o in (0.3, 1.0):
        wo = _hemi(theta_o, 0.7)
        nt, nphi = 400, 200
        ct = (np.arange(nt) + 0.5) / nt
        ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
        alb, mass = 0.0, 0.0
        for c in ct:
            st = np.sqrt(1 - c * c)
            for p in ph[::2]:
                wi = np.array([st * np.cos(p), st * np.sin(p), c])
                f, pdf = uber(kd, ks, op, 1.5, ax, ay, wo, wi)
                alb += f[1] * c
                mass += pdf
        dA = (1.0 / nt) * (2 * np.pi / (nphi / 2))
        alb *= dA
        mass *= dA
        n = 6000
        est, valid = 0.0, 0
        for u in rng.uniform(size=(n, 3)):
            ok, wi, f, pdf = uber(kd, ks, op, 1.5, ax, ay, wo, u=u)
            if ok:
                est += f[1] * wi[2] / pdf
                valid += 1
        est /= n
        assert est == pytest.approx(alb, rel=0.03), (theta_o, est, alb)
        assert mass == pytest.approx(valid / n, abs=0.02), (theta_o, mass, valid / n)


def test_uber_delta_reflection_is_fresnel_weighted_mirror():
    # an uber surface with only Kr (one component, no BSDF layer) is a mirror whose reflection is
    # weighted by the dielectric Fresnel term at the incidence angle (get_pbrt_uber_info,
    # delta_scattering.cpp:115-131): the 45-degree mirror scene then shows
    # Kr * F * albedo/pi * I * cos / d^2 with F = F(cos 45, 1.5) where the quad faces the camera, and
    # F = 1 where it faces away (the ray leaves the dielectric: 45 degrees is past the critical 41.8)
    from tests.scenes import DescScene
    from tests.test_oracle_fixtures import _view_from_desc
    from yar_amd import _abi
    s = 0.3 / np.sqrt(2)
    c = np.array([1.0, 1.0, 0.0])
    e1, e2 = np.array([0, 0, 0.3]), np.array([s, -s, 0])
    plane = {"vertices": [[-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]], "indices": [0, 1, 2, 0, 2, 3]}
    kr, albedo, intensity = [0.9, 0.6, 0.3], 0.5, 2.0
    for facing, F in (([0, 2, 1, 0, 3, 2], _textbook_fresnel(np.cos(np.pi / 4), 1.0, 1.5)), ([0, 1, 2, 0, 2, 3], 1.0)):
        quad = {"vertices": [c - e1 - e2, c + e1 - e2, c + e1 + e2, c - e1 + e2], "indices": facing}
        sc = DescScene(
            meshes=[plane, quad],
            materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [albedo] * 3},
                       {"type": _abi.YAR_MATERIAL_PBRT3_UBER, "reflectance": [0, 0, 0], "specular": [0, 0, 0],
                        "delta_reflectance": kr, "ior": 1.5, "u_roughness": 0.2}],
            objects=[{"mesh": 0, "material": 0}, {"mesh": 1, "material": 1}],
            points=[{"position": [4.0, 2.0, 0.0], "intensity": [intensity] * 3}],
            camera={"c2w": [0, 0, -1, 0, 0, 1, 0, 1, 1, 0, 0, 0], "fov": 1.0},
            config={"film_width": 16, "film_height": 16, "region": (0, 0, 16, 16), "nx": 2, "ny": 2, "maxdepth": 2})
        img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
        want = np.array(kr) * F * albedo / np.pi * intensity * 2.0 / 13.0 ** 1.5
        assert np.allclose(img.mean(axis=(0, 1)), want, rtol=2e-3), (F, img.mean(axis=(0, 1)), want)


def test_uber_opacity_passes_light_through():
    # Kd plus opacity o: two components, each picked with probability 1/2.  The pass-through
    # (delta_scattering.cpp:135-140) carries 2 (1 - o) of what lies behind, the BSDF layer is divided
    # by its selection probability 1/2, and direct lighting is estimated in both branches
    # (path_tracing.cpp:53-79), so a point-lit uber plane in front of an emitter of radiance Le shows
    # Kd o / pi * I / d^2 + (1 - o) Le.
    from tests.scenes import DescScene
    from tests.test_oracle_fixtures import _view_from_desc
    from yar_amd import _abi
    kd, o, intensity, le = 0.5, 0.3, 2.0, 1.0
    big = lambda y, up: {"vertices": [[-50, y, -50], [-50, y, 50], [50, y, 50], [50, y, -50]],
                         "indices": [0, 2, 1, 0, 3, 2] if up else [0, 1, 2, 0, 2, 3]}
    rect_l2w = [1, 0, 0, 0, 0, 0, 1, -1, 0, -1, 0, 0]                      # columns x, y, z = +y normal
    sc = DescScene(
        meshes=[big(0.0, True), big(-1.0, True)],
        materials=[{"type": _abi.YAR_MATERIAL_PBRT3_UBER, "reflectance": [kd] * 3, "specular": [0, 0, 0],
                    "opacity": [o] * 3, "ior": 1.5, "u_roughness": 0.2}],
        objects=[{"mesh": 0, "material": 0}, {"mesh": 1, "light": 0}],
        rects=[{"l2w": rect_l2w, "L": [le] * 3, "size": [100.0, 100.0]}],
        points=[{"position": [0.0, 1.0, 0.0], "intensity": [intensity] * 3}],
        camera={"c2w": [1, 0, 0, 0, 0, 0, 1, 0.5, 0, -1, 0, 0], "fov": 0.5},
        config={"film_width": 16, "film_height": 16, "region": (0, 0, 16, 16), "nx": 8, "ny": 8, "maxdepth": 1})
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    want = kd * o / np.pi * intensity + (1 - o) * le
    assert img.mean() == pytest.approx(want, rel=2e-3), (img.mean(), want)


def _env_scene(tmp_path, body, look, rot="", spp=16, maxdepth=1, name="env.pbrt"):
    text = (f'LookAt {look}\nCamera "perspective" "float fov" [ 0.5 ]\n'
            'Film "image" "integer xresolution" [ 16 ] "integer yresolution" [ 16 ]\n'
            f'Sampler "stratified" "integer xsamples" [ {spp} ] "integer ysamples" [ {spp} ]\n'
            f'Integrator "path" "integer maxdepth" [ {maxdepth} ]\nWorldBegin\n'
            f'AttributeBegin\n{rot}LightSource "infinite" "rgb L" [ 0.5 1 1.5 ] "rgb scale" [ 2 2 2 ]\nAttributeEnd\n'
            + body + 'WorldEnd\n')
    p = tmp_path / name
    p.write_text(text)
    return yar_amd.Scene(str(p))


@pytest.mark.parametrize("rot", ["", "Rotate 70 1 0.3 0\n"], ids=["aligned", "rotated"])
def test_environment_light_furnace(tmp_path, rot):
    # direct_lighting.cpp:324-384, light_sampling.cpp:11-73: a diffuse plane under a constant sky of
    # radiance L receives irradiance pi L, so it shows albedo * L; the MIS of the (phi, theta)-uniform
    # light samples (pdf 1 / (2 pi^2 sin theta) in the light frame) and of the BSDF samples must sum
    # to that however the light frame is turned.  A camera ray that leaves the scene sees L.
    from tests.test_oracle_fixtures import _view_from_desc
    L = np.array([1.0, 2.0, 3.0])
    plane = ('AttributeBegin\nMaterial "matte" "rgb Kd" [ 0.5 0.5 0.5 ]\nShape "trianglemesh" "integer indices" '
             '[ 0 1 2 0 2 3 ] "point P" [ -1e4 0 -1e4  -1e4 0 1e4  1e4 0 1e4  1e4 0 -1e4 ]\nAttributeEnd\n')
    sc = _env_scene(tmp_path, plane, "0 0.5 0  0 0 0  0 0 1", rot)
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    assert np.allclose(img.mean(axis=(0, 1)), 0.5 * L, rtol=2e-3), (img.mean(axis=(0, 1)), 0.5 * L)
    sky = _env_scene(tmp_path, plane, "0 0.5 0  0 2 0  0 0 1", rot, spp=2, name="sky.pbrt")
    img, _, _, _, _ = O.OracleScene(sky).render(_view_from_desc(sky.desc))
    assert np.allclose(img, L, rtol=1e-6), (img.min(), img.max())


def test_environment_light_convex_object_and_delta_miss(tmp_path):
    # A convex diffuse object sees only the sky from every point, so at any path depth it shows
    # albedo * L (the continuation rays leave the scene and add nothing after the first vertex);
    # a mirror seen by the camera reflects the sky: Kr * L through the delta-miss branch
    # (path_tracing.cpp:86-96).
    from tests.test_oracle_fixtures import _view_from_desc
    L = np.array([1.0, 2.0, 3.0])
    ball = 'AttributeBegin\nMaterial "matte" "rgb Kd" [ 0.6 0.6 0.6 ]\nShape "sphere" "float radius" [ 1 ]\nAttributeEnd\n'
    sc = _env_scene(tmp_path, ball, "0 0 5  0 0 0  0 1 0", maxdepth=6)
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    assert np.allclose(img.mean(axis=(0, 1)), 0.6 * L, rtol=3e-3), (img.mean(axis=(0, 1)), 0.6 * L)
    mirror = ('AttributeBegin\nMaterial "mirror" "rgb Kr" [ 0.8 0.5 0.25 ]\nShape "trianglemesh" "integer indices" '
              '[ 0 1 2 0 2 3 ] "point P" [ -1e4 0 -1e4  -1e4 0 1e4  1e4 0 1e4  1e4 0 -1e4 ]\nAttributeEnd\n')
    sc = _env_scene(tmp_path, mirror, "0 0.5 0  0.3 0 0.2  0 1 0", spp=2, maxdepth=2, name="m.pbrt")
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    assert np.allclose(img, np.array([0.8, 0.5, 0.25]) * L, rtol=1e-5), (img.min(axis=(0, 1)), img.max(axis=(0, 1)))


def _mix_plane(tmp_path, material_lines, light, look="0 0.5 0  0 0 0  0 0 1", spp=16, name="mix.pbrt"):
    text = (f'LookAt {look}\nCamera "perspective" "float fov" [ 0.5 ]\n'
            'Film "image" "integer xresolution" [ 8 ] "integer yresolution" [ 8 ]\n'
            f'Sampler "stratified" "integer xsamples" [ {spp} ] "integer ysamples" [ {spp} ]\n'
            'Integrator "path" "integer maxdepth" [ 1 ]\nWorldBegin\n' + light +
            'MakeNamedMaterial "one" "string type" [ "plastic" ] "rgb Kd" [ 0.2 0.5 0.3 ] "rgb Ks" [ 0.4 0.4 0.4 ] "float roughness" [ 0.05 ]\n'
            'MakeNamedMaterial "two" "string type" [ "substrate" ] "rgb Kd" [ 0.6 0.2 0.1 ] "float uroughness" [ 0.1 ] "float vroughness" [ 0.3 ]\n'
            'MakeNamedMaterial "red" "string type" [ "matte" ] "rgb Kd" [ 0.8 0.1 0.3 ]\n'
            'MakeNamedMaterial "blue" "string type" [ "matte" ] "rgb Kd" [ 0.1 0.3 0.9 ]\n'
            + material_lines +
            'Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ -1e4 0 -1e4  -1e4 0 1e4  1e4 0 1e4  1e4 0 -1e4 ]\n'
            'WorldEnd\n')
    p = tmp_path / name
    p.write_text(text)
    return yar_amd.Scene(str(p))


def test_mix_of_diffuse_point_lit_closed_form(tmp_path):
    # Mix_BSDF (bsdf.cpp:497-521): f = amount f1 + (1 - amount) f2 per channel, so a point-lit mix of
    # two matte materials shows (a Kd1 + (1 - a) Kd2) / pi * I / d^2
    from tests.test_oracle_fixtures import _view_from_desc
    a = np.array([0.3, 0.6, 0.9])
    sc = _mix_plane(tmp_path, 'Material "mix" "string namedmaterial1" [ "red" ] "string namedmaterial2" [ "blue" ] '
                    '"rgb amount" [ 0.3 0.6 0.9 ]\n', 'LightSource "point" "rgb I" [ 2 2 2 ] "point from" [ 0 1 0 ]\n', spp=2)
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    want = (a * np.array([0.8, 0.1, 0.3]) + (1 - a) * np.array([0.1, 0.3, 0.9])) / np.pi * 2.0
    assert np.allclose(img.mean(axis=(0, 1)), want, rtol=1e-4), (img.mean(axis=(0, 1)), want)


def test_mix_is_linear_in_its_sub_materials(tmp_path):
    # Under a constant sky, the plane's radiance is the sky-weighted albedo of its BSDF; for a mix it
    # must equal a R1 + (1 - a) R2 with R1, R2 the renders of each sub-material alone.  This pins the
    # sample choice (probability 1 - luminance(1 - amount) for the second BSDF), the mixed pdf and the
    # mixed f together: any of them wrong biases the MIS estimate.
    from tests.test_oracle_fixtures import _view_from_desc
    sky = 'LightSource "infinite" "rgb L" [ 1 1 1 ]\n'
    look = "0 0.5 0.4  0 0 0  0 1 0"
    a = np.array([0.25, 0.5, 0.7])
    means = []
    for i, mat in enumerate(['NamedMaterial "one"\n', 'NamedMaterial "two"\n',
                             'Material "mix" "string namedmaterial1" [ "one" ] "string namedmaterial2" [ "two" ] '
                             '"rgb amount" [ 0.25 0.5 0.7 ]\n']):
        sc = _mix_plane(tmp_path, mat, sky, look=look, spp=32, name=f"m{i}.pbrt")
        img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
        means.append(img.mean(axis=(0, 1)).astype(np.float64))
    want = a * means[0] + (1 - a) * means[1]
    assert np.allclose(means[2], want, rtol=5e-3), (means, want)


def test_mix_with_a_delta_layer(tmp_path):
    # get_mix_info / create_mix_bsdf (delta_scattering.cpp:154-189, bsdf.cpp:599-629): with a mirror as
    # the first sub-material and matte as the second, the mirror is picked with probability
    # m = luminance(amount) and its reflection is not reweighted, and the matte BSDF is scaled by
    # 1 - luminance(amount) (the luminance, not per channel) and estimated in both branches.
    # (1) Seen straight down and point-lit: (1 - lum(a)) Kd / pi * I per channel.
    from tests.test_oracle_fixtures import _view_from_desc
    a = np.array([0.2, 0.5, 0.9])
    lum = 0.2126729 * a[0] + 0.7151522 * a[1] + 0.0721750 * a[2]
    text = ('LookAt 0 0.5 0  0 0 0  0 0 1\nCamera "perspective" "float fov" [ 0.5 ]\n'
            'Film "image" "integer xresolution" [ 8 ] "integer yresolution" [ 8 ]\n'
            'Sampler "stratified" "integer xsamples" [ 16 ] "integer ysamples" [ 16 ]\n'
            'Integrator "path" "integer maxdepth" [ 1 ]\nWorldBegin\n'
            'LightSource "point" "rgb I" [ 2 2 2 ] "point from" [ 0 1 0 ]\n'
            'MakeNamedMaterial "mi" "string type" [ "mirror" ] "rgb Kr" [ 0.9 0.9 0.9 ]\n'
            'MakeNamedMaterial "ma" "string type" [ "matte" ] "rgb Kd" [ 0.8 0.4 0.2 ]\n'
            'Material "mix" "string namedmaterial1" [ "mi" ] "string namedmaterial2" [ "ma" ] "rgb amount" [ 0.2 0.5 0.9 ]\n'
            'Shape "trianglemesh" "integer indices" [ 0 1 2 0 2 3 ] "point P" [ -1e4 0 -1e4  -1e4 0 1e4  1e4 0 1e4  1e4 0 -1e4 ]\n'
            'WorldEnd\n')
    p = tmp_path / "mixd.pbrt"
    p.write_text(text)
    sc = yar_amd.Scene(str(p))
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    want = (1 - lum) * np.array([0.8, 0.4, 0.2]) / np.pi * 2.0
    assert np.allclose(img.mean(axis=(0, 1)), want, rtol=1e-3), (img.mean(axis=(0, 1)), want)
    # (2) The same mix as the 45-degree mirror of the mirror closed form, with a black matte layer:
    # only the picked mirror carries light, Kr * lum(a) * albedo/pi * I * cos / d^2
    from tests.scenes import DescScene
    from yar_amd import _abi
    s = 0.3 / np.sqrt(2)
    c = np.array([1.0, 1.0, 0.0])
    e1, e2 = np.array([0, 0, 0.3]), np.array([s, -s, 0])
    quad = {"vertices": [c - e1 - e2, c + e1 - e2, c + e1 + e2, c - e1 + e2], "indices": [0, 1, 2, 0, 2, 3]}
    plane = {"vertices": [[-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]], "indices": [0, 1, 2, 0, 2, 3]}
    sc = DescScene(
        meshes=[plane, quad],
        materials=[{"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [0.5] * 3},
                   {"type": _abi.YAR_MATERIAL_MIRROR, "reflectance": [0.9, 0.6, 0.3]},
                   {"type": _abi.YAR_MATERIAL_DIFFUSE, "reflectance": [0.0] * 3},
                   {"type": _abi.YAR_MATERIAL_MIX, "mix_materials": [1, 2], "mix_amount": list(a)}],
        objects=[{"mesh": 0, "material": 0}, {"mesh": 1, "material": 3}],
        points=[{"position": [4.0, 2.0, 0.0], "intensity": [2.0] * 3}],
        camera={"c2w": [0, 0, -1, 0, 0, 1, 0, 1, 1, 0, 0, 0], "fov": 1.0},
        config={"film_width": 16, "film_height": 16, "region": (0, 0, 16, 16), "nx": 16, "ny": 16, "maxdepth": 2})
    img, _, _, _, _ = O.OracleScene(sc).render(_view_from_desc(sc.desc))
    want = np.array([0.9, 0.6, 0.3]) * lum * 0.5 / np.pi * 2.0 * 2.0 / 13.0 ** 1.5
    assert np.allclose(img.mean(axis=(0, 1)), want, rtol=3e-3), (img.mean(axis=(0, 1)), want)

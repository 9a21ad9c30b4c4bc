// shading.hpp - shading frame, BSDFs and MIS helpers of the wavefront tracer (device inline code).
//
// Restates src/ref/shading_context.cpp:17-240 (local geometry, shading-normal adaptation, ray-origin
// offsets), src/ref/bsdf.cpp:15-200 (Diffuse_BRDF, Metal_BRDF), src/ref/scattering.cpp (Fresnel,
// anisotropic GGX) and src/ref/sampling.cpp:43-53,311-359 (cosine and VNDF sampling) with the
// reference's operation order.
#pragma once

#include "device.hpp"

namespace yg {

// Whether a BSDF can scatter light arriving from direction wi: reflection above the shading
// normal, transmission below it (direct_lighting.cpp:133-135).
YG_INLINE bool scattering_possible(const DevMaterial& m, float n_dot_wi) {
    return n_dot_wi > 0.f || (n_dot_wi < 0.f && (m.type == kRoughGlassType || (m.type == kMixType && m.trans.w != 0.f)));
}

YG_INLINE float power_heuristic(float a, float b) {             // direct_lighting.cpp:18
    if (a == f_inf()) return 1.f;
    return a * a / (a * a + b * b);
}

// ---------------------------------------------------------------------------------------------
// Shading context (shading_context.cpp:17-240, 405-418; math.cpp:21-120)
// ---------------------------------------------------------------------------------------------
struct Shading {
    V3 wo, pos, ng, n, T, B, org_pos, org_neg;
    bool flipped;
    int material, light;
};

YG_INLINE void coordinate_system(V3 v, V3& v1) {                                   // math.cpp:21
    v1 = (fabsf(v.x) > fabsf(v.y)) ? v3(-v.z, 0.f, v.x) : v3(0.f, -v.z, v.y);
    v1 = vdiv(v1, length(v1));
}
YG_INLINE V3 reflect(V3 w, V3 n) { return (2.f * dot(w, n)) * n - w; }             // math.h:111
YG_INLINE V3 origin_dir(const Shading& sc, V3 d) { return dot(d, sc.ng) > 0.f ? sc.org_pos : sc.org_neg; }

YG_INLINE void offset_origins(V3 p, V3 ng, V3& pos, V3& neg) {                      // math.cpp:81-120
    const float int_scale = 256.f, origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f;
    float pa[3] = {p.x, p.y, p.z}, na[3] = {ng.x, ng.y, ng.z}, po[3], ne[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int di = (int)(int_scale * na[a]);
        int ib = (int)fbits(pa[a]);
        int i1 = ib - di, i2 = ib + di;
        po[a] = bitsf((uint32_t)(pa[a] < 0 ? i1 : i2));
        ne[a] = bitsf((uint32_t)(pa[a] < 0 ? i2 : i1));
        if (fabsf(pa[a]) < origin) {
            float dlt = float_scale * na[a];
            po[a] = pa[a] + dlt;
            ne[a] = pa[a] - dlt;
        }
    }
    pos = v3(po[0], po[1], po[2]);
    neg = v3(ne[0], ne[1], ne[2]);
}

YG_INLINE void adjust_normal(V3 wo, V3 ng, V3& n) {                                 // shading_context.cpp:17
    V3 R = reflect(wo, n);
    float a = dot(R, ng);
    if (a >= 0.f) return;
    if (dot(wo, ng) < 0.0017453f) { n = ng; return; }
    float b = dot(n, ng);
    V3 tangent;
    if (b > 1e-4f) {
        float dist = fabsf(a) / b;
        tangent = R + dist * n;
        tangent = vdiv(tangent, length(tangent));
    } else {
        tangent = n;
    }
    tangent = tangent + 1e-4f * ng;
    n = normalized(wo + tangent);
}

YG_INLINE void init_shading(const DevScene& s, V3 ray_d, int obj, uint32_t tri, V3 b, Shading& sc) {
    const int4 oi = s.obj_info[obj];
    const int mflags = s.mesh_info[oi.z].x;
    sc.wo = -normalized(ray_d);
    const V3 p0 = from4(s.tri_verts[3 * (size_t)tri]), p1 = from4(s.tri_verts[3 * (size_t)tri + 1]),
             p2 = from4(s.tri_verts[3 * (size_t)tri + 2]);
    V3 position = b.x * p0 + b.y * p1 + b.z * p2;
    V3 ng = normalized(cross(p1 - p0, p2 - p0));
    if (mflags & 4) ng = -ng;
    V3 n = ng;
    const int4 vi = s.tri_vidx[tri];
    if (mflags & 1) {
        V3 n0 = from4(s.vtx_n[vi.x]), n1 = from4(s.vtx_n[vi.y]), n2 = from4(s.vtx_n[vi.z]);
        V3 in = b.x * n0 + b.y * n1 + b.z * n2;
        if (len_sq(in) != 0.f) n = normalized(in);
    }
    const bool has_uv = (mflags & 2) != 0;
    V3 dpdu = v3(0.f, 0.f, 0.f);
    if (has_uv) {
        const float2 t0 = s.vtx_uv[vi.x], t1 = s.vtx_uv[vi.y], t2 = s.vtx_uv[vi.z];
        float a00 = t1.x - t0.x, a01 = t1.y - t0.y, a10 = t2.x - t0.x, a11 = t2.y - t0.y;
        V3 bp0 = p1 - p0, bp1 = p2 - p0;
        float det = a00 * a11 - a01 * a10;                                            // math.cpp:8
        if (det != 0.f) {
            float inv = 1.f / det;
            dpdu = (a11 * bp0 - a01 * bp1) * inv;
        }
        if (fabsf(det) < 1e-9f) coordinate_system(ng, dpdu);
    }
    const M34 o2w = {s.obj_o2w[3 * obj], s.obj_o2w[3 * obj + 1], s.obj_o2w[3 * obj + 2]};
    const M34 nrm = {s.obj_nrm[3 * obj], s.obj_nrm[3 * obj + 1], s.obj_nrm[3 * obj + 2]};
    sc.pos = xform_point(o2w, position);
    n = xform_vector(nrm, n); n = vdiv(n, length(n));
    ng = xform_vector(nrm, ng); ng = vdiv(ng, length(ng));
    dpdu = xform_vector(o2w, dpdu);
    if (dot(n, ng) == 0.f) ng = n;
    if (dot(ng, sc.wo) < 0.f) ng = -ng;
    sc.flipped = false;
    if (dot(n, ng) < 0.f) { n = -n; sc.flipped = true; }
    offset_origins(sc.pos, ng, sc.org_pos, sc.org_neg);
    adjust_normal(sc.wo, ng, n);
    V3 dpdu_s;
    if (!has_uv) coordinate_system(n, dpdu_s);
    else if (fabsf(dot(n, ng) - 1.f) < 1e-6f) dpdu_s = dpdu;
    else {
        V3 t = cross(dpdu, n);                                                         // math.cpp:34
        dpdu_s = normalized(cross(n, t));
        dpdu_s = dpdu_s * length(dpdu);
    }
    sc.ng = ng;
    sc.n = n;
    sc.B = normalized(cross(n, dpdu_s));                                               // bsdf.cpp:15-21
    sc.T = cross(sc.B, n);
    sc.material = oi.x;
    sc.light = oi.y;
}

// ---------------------------------------------------------------------------------------------
// BSDFs: Diffuse_BRDF (bsdf.cpp:60-84), Metal_BRDF (bsdf.cpp:146-200) with scattering.cpp and
// sampling.cpp:43-53, 311-359
// ---------------------------------------------------------------------------------------------
YG_INLINE V3 to_world(const Shading& sc, V3 l) {
    return v3(sc.T.x * l.x + sc.B.x * l.y + sc.n.x * l.z, sc.T.y * l.x + sc.B.y * l.y + sc.n.y * l.z,
              sc.T.z * l.x + sc.B.z * l.y + sc.n.z * l.z);
}
YG_INLINE V3 to_local(const Shading& sc, V3 w) { return v3(dot(w, sc.T), dot(w, sc.B), dot(w, sc.n)); }

YG_INLINE RGB conductor_fresnel(float cos_i, float eta_i, RGB eta_t, RGB k_t) {
    cos_i = fabsf(fminf(fmaxf(cos_i, -1.f), 1.f));
    float c2 = cos_i * cos_i, s2 = 1.f - c2;
    RGB eta = cdivs(eta_t, eta_i), k = cdivs(k_t, eta_i);
    RGB eta2 = eta * eta, k2 = k * k;
    RGB t0 = eta2 - k2 - splat(s2);
    RGB a2b2 = csqrt(t0 * t0 + 4.f * eta2 * k2);
    RGB t1 = a2b2 + splat(c2);
    RGB a = csqrt(0.5f * (a2b2 + t0));
    RGB t2 = (2.f * cos_i) * a;
    RGB Rs = cdiv(t1 - t2, t1 + t2);
    RGB t3 = c2 * a2b2 + splat(s2 * s2);
    RGB t4 = t2 * s2;
    RGB Rp = cdiv(Rs * (t3 - t4), t3 + t4);
    return 0.5f * (Rp + Rs);
}
YG_INLINE float dielectric_fresnel(float cos_i, float eta) {
    cos_i = fminf(fabsf(cos_i), 1.f);
    float sin_i = sqrtf(1.f - cos_i * cos_i);
    float sin_t = (1.f / eta) * sin_i;
    if (sin_t >= 1.f) return 1.f;
    float cos_t = sqrtf(1.f - sin_t * sin_t);
    float Rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float Rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (Rp * Rp + Rs * Rs);
}
YG_INLINE float D_aniso(V3 wh, float ax, float ay) {
    float c = fminf(fmaxf(wh.z, -1.f), 1.f);
    float c2 = c * c, s2 = 1.f - c2;
    float s2inv = 1.f / s2;
    float cp = (s2 == 0.f) ? 1.f : (wh.x * wh.x) * s2inv;
    float sp = (s2 == 0.f) ? 0.f : (wh.y * wh.y) * s2inv;
    float k = c2 + s2 * (cp / (ax * ax) + sp / (ay * ay));
    return 1.f / (kPi * ax * ay * k * k);
}
YG_INLINE float lambda_aniso(V3 v, float ax, float ay) {
    float c = fminf(fmaxf(v.z, -1.f), 1.f);
    float c2 = c * c, s2 = 1.f - c2;
    float tan2 = s2 / c2;
    float s2inv = 1.f / s2;
    float cp = (s2 == 0.f) ? 1.f : (v.x * v.x) * s2inv;
    float sp = (s2 == 0.f) ? 0.f : (v.y * v.y) * s2inv;
    float alpha2 = cp * ax * ax + sp * ay * ay;
    return 0.5f * (-1.f + sqrtf(1.f + alpha2 * tan2));
}
YG_INLINE float G_aniso(V3 wi, V3 wo, float ax, float ay) {
    return 1.f / (1.f + lambda_aniso(wi, ax, ay) + lambda_aniso(wo, ax, ay));
}
YG_INLINE float G1_aniso(V3 v, float ax, float ay) { return 1.f / (1.f + lambda_aniso(v, ax, ay)); }

YG_INLINE V3 vndf_sample(float u0, float u1, V3 wo, float ax, float ay) {
    V3 N = normalized(v3(ax * wo.x, ay * wo.y, wo.z));
    float lsq = N.x * N.x + N.y * N.y;
    V3 T1 = lsq > 0.f ? vdiv(v3(-N.y, N.x, 0.f), sqrtf(lsq)) : v3(1.f, 0.f, 0.f);
    V3 T2 = cross(N, T1);
    float r = sqrtf(u0);
    float phi = 2.f * kPi * u1;
    float t1 = r * cosf(phi);
    float t2 = r * sinf(phi);
    float sgl = 0.5f * (1.f + N.z);
    t2 = (1.f - sgl) * sqrtf(1.f - t1 * t1) + sgl * t2;
    V3 nh = t1 * T1 + t2 * T2 + sqrtf(fmaxf(0.f, 1.f - t1 * t1 - t2 * t2)) * N;
    return normalized(v3(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}
YG_INLINE float vndf_wi_pdf(V3 wo, V3 wh, float ax, float ay) {
    float G1 = G1_aniso(wo, ax, ay);
    float D = D_aniso(wh, ax, ay);
    float wh_pdf = G1 * D * fmaxf(0.f, dot(wo, wh)) / wo.z;
    return wh_pdf / (4 * dot(wh, wo));
}

// Isotropic GGX against the shading normal (GGX_Distribution::D / lambda / G / G1, scattering.cpp:153-238)
// and the visible-normal reflection pdf (sampling.cpp:337-347, scattering.cpp:107-116).
YG_INLINE float D_iso(V3 wh, V3 n, float alpha) {
    float c = dot(wh, n);
    if (c <= 0.f) return 0.f;
    float c2 = c * c;
    float a2 = alpha * alpha;
    float k = c2 * (a2 - 1.f) + 1.f;
    return a2 / (kPi * k * k);
}
YG_INLINE float lambda_iso(V3 v, V3 n, float alpha) {
    float c = dot(v, n);
    float c2 = c * c;
    float t2 = fmaxf(0.f, (1.f - c2) / c2);
    return 0.5f * (-1.f + sqrtf(1.f + alpha * alpha * t2));
}
YG_INLINE float G_iso(V3 wi, V3 wo, V3 n, float a) { return 1.f / (1.f + lambda_iso(wi, n, a) + lambda_iso(wo, n, a)); }
YG_INLINE float G1_iso(V3 v, V3 n, float a) { return 1.f / (1.f + lambda_iso(v, n, a)); }
YG_INLINE float vndf_wi_pdf_iso(V3 wo, V3 wh, V3 n, float alpha) {
    float wh_pdf = G1_iso(wo, n, alpha) * D_iso(wh, n, alpha) * fmaxf(0.f, dot(wo, wh)) / dot(wo, n);
    return wh_pdf / (4.f * dot(wh, wo));
}

// Plastic_BRDF (bsdf.cpp:214-262) and its pbrt3 form Pbrt3_Plastic_BRDF (bsdf_pbrt.cpp:264-291, Fresnel
// of a 1.5 dielectric with pbrt3's flipped relative ior).  Material: refl = Kd, alpha_x, trans.x = r0,
// trans.y != 0 for the pbrt3 form.
YG_INLINE RGB plastic_eval(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    V3 wh = normalized(wo + wi);
    float cos_i = dot(wi, wh);
    RGB F;
    if (m.trans.y != 0.f) {
        float rel = 1.f / 1.5f;
        V3 original = sc.flipped ? -sc.n : sc.n;
        if (dot(original, wi) < 0.f) rel = 1.5f / 1.f;
        F = splat(dielectric_fresnel(cos_i, rel));
    } else {
        float k = 1.f - fabsf(cos_i);                                                  // scattering.cpp:10
        float k5 = (k * k) * (k * k) * k;
        F = splat(0.04f) + (splat(1.f) - splat(0.04f)) * k5;
    }
    float G = G_iso(wi, wo, sc.n, m.alpha_x);
    float D = D_iso(wh, sc.n, m.alpha_x);
    float won = dot(wo, sc.n), win = dot(wi, sc.n);
    RGB base = F * ((G * D) / (4.f * won * win));
    RGB spec = m.trans.x * base;
    RGB diffuse = rgb4(m.refl) * kPiInv;
    return diffuse + spec;
}
YG_INLINE float plastic_pdf(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    float diffuse_pdf = dot(sc.n, wi) / kPi;
    V3 wh = normalized(wo + wi);
    float spec_pdf = vndf_wi_pdf_iso(wo, wh, sc.n, m.alpha_x);
    return 0.5f * (diffuse_pdf + spec_pdf);
}

// Ashikhmin_Shirley_Phong_BRDF (bsdf.cpp:427-495): Schlick-Fresnel anisotropic GGX specular over a
// diffuse base scaled by (1 - r0).  Material: refl = diffuse reflectance, k = r0, alpha_x/y.
YG_INLINE RGB coated_eval(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    const V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    const V3 wh = normalized(wol + wil);
    const float cos_i = dot(wil, wh);
    const RGB r0 = rgb4(m.k);
    const float k = 1.f - fabsf(cos_i);                                                 // scattering.cpp:10
    const float k5 = (k * k) * (k * k) * k;
    const RGB F = r0 + (splat(1.f) - r0) * k5;
    const float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    const RGB spec = F * (D / (4.f * cos_i * fmaxf(wol.z, wil.z)));
    const float ai = 1.f - 0.5f * wil.z, ao = 1.f - 0.5f * wol.z;
    const float pi5 = (ai * ai) * (ai * ai) * ai, po5 = (ao * ao) * (ao * ao) * ao;
    const RGB diffuse = (rgb4(m.refl) * (splat(1.f) - r0)) * (28.f / (23.f * kPi) * (1.f - pi5) * (1.f - po5));
    return diffuse + spec;
}
YG_INLINE float coated_pdf(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    const float diffuse_pdf = dot(sc.n, wi) / kPi;
    const V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    const V3 wh = normalized(wol + wil);
    const float spec_pdf = vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
    return 0.5f * (diffuse_pdf + spec_pdf);
}

// Rough_Glass_BSDF (bsdf.cpp:268-420; scattering.cpp:65-150): GGX reflection and transmission
// through a dielectric boundary.  Material: refl = Kr, trans = Kt with w = index of refraction,
// alpha_x = GGX alpha.  The side the ray enters from follows the original shading-normal
// orientation (pbrt scenes do not track nested dielectrics).
YG_INLINE void rough_glass_etas(const DevMaterial& m, const Shading& sc, float& eta_o, float& eta_i) {
    const bool enter = !sc.flipped;
    eta_o = enter ? 1.f : m.trans.w;
    eta_i = enter ? m.trans.w : 1.f;
}
YG_INLINE float max_component(RGB c) { return fmaxf(c.r, fmaxf(c.g, c.b)); }
YG_INLINE V3 refraction_half(float eta_o, V3 wo, float eta_i, V3 wi, V3 n) {
    V3 wh = -normalized(eta_o * wo + eta_i * wi);
    if (dot(wh, n) < 0.f) wh = -wh;
    return wh;
}
YG_INLINE bool refract_dir(V3 w, V3 n, float eta, V3& out) {                          // math.h:115
    const float ci = dot(w, n);
    const float st2 = eta * eta * fmaxf(0.f, 1.f - ci * ci);
    if (st2 >= 1.f) return false;
    const float ct = sqrtf(1.f - st2);
    out = (-eta) * w + (eta * ci - ct) * n;
    return true;
}
YG_INLINE float reflection_probability(const DevMaterial& m, float fresnel) {
    const float max_r = max_component(rgb4(m.refl)), max_t = max_component(rgb4(m.trans));
    const float ratio = max_r / (max_r + max_t);
    const float r = fresnel * ratio, t = (1.f - fresnel) * (1 - ratio);
    return ((r + t) == 0) ? 0.f : r / (r + t);
}
YG_INLINE RGB rough_glass_eval(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    float eta_o, eta_i;
    rough_glass_etas(m, sc, eta_o, eta_i);
    const V3 n = sc.n;
    const float a = m.alpha_x;
    if (dot(wo, n) * dot(wi, n) > 0.f) {
        const V3 wh = normalized(wo + wi);
        const float F = dielectric_fresnel(dot(wi, wh), eta_i / eta_o);
        const float G = G_iso(wi, wo, n, a), D = D_iso(wh, n, a);
        const float base = (F * G * D) / (4.f * dot(wo, n) * dot(wi, n));
        return rgb4(m.refl) * base;
    }
    const V3 wh = refraction_half(eta_o, wo, eta_i, wi, n);
    const float wow = dot(wo, wh), wiw = dot(wi, wh);
    if (wow * wiw > 0.f) return rgb(0.f, 0.f, 0.f);
    const float F = dielectric_fresnel(wiw, eta_o / eta_i);
    if (F == 1.f) return rgb(0.f, 0.f, 0.f);
    const float G = G_iso(wi, wo, n, a), D = D_iso(wh, n, a);
    const float k = fabsf((wow * wiw) / (dot(wo, n) * dot(wi, n)));
    const float k2 = eta_o * wow + eta_i * wiw;
    const float base = (eta_o * eta_o * k * G * D * (1.f - F)) / (k2 * k2);
    return rgb4(m.trans) * base;
}
YG_INLINE float rough_glass_pdf(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    float eta_o, eta_i;
    rough_glass_etas(m, sc, eta_o, eta_i);
    const V3 n = sc.n;
    const float a = m.alpha_x;
    if (dot(wo, n) * dot(wi, n) > 0.f) {
        const V3 wh = normalized(wo + wi);
        const float rpdf = vndf_wi_pdf_iso(wo, wh, n, a);
        return rpdf * reflection_probability(m, dielectric_fresnel(dot(wi, wh), eta_i / eta_o));
    }
    const V3 wh = refraction_half(eta_o, wo, eta_i, wi, n);
    if (dot(wo, wh) * dot(wi, wh) > 0.f) return 0.f;
    const float rp = reflection_probability(m, dielectric_fresnel(dot(wo, wh), eta_i / eta_o));
    const float wh_pdf = G1_iso(wo, n, a) * D_iso(wh, n, a) * fmaxf(0.f, dot(wo, wh)) / dot(wo, n);
    const float denom = eta_o * dot(wo, wh) + eta_i * dot(wi, wh);
    const float dwh_dwi = eta_i * eta_i * fabsf(dot(wi, wh)) / (denom * denom);
    return wh_pdf * dwh_dwi * (1.f - rp);
}

// Pbrt3_Uber_BRDF (bsdf_pbrt.cpp:14-91): refl = Kd | ior, trans = Ks, k = opacity; Kd/pi and an
// anisotropic GGX lobe with a dielectric's Fresnel (ior inverted when leaving), both scaled by the opacity.
YG_INLINE RGB uber_eval(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    const RGB diffuse = (kPiInv * rgb4(m.refl)) * rgb4(m.k);
    const V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    const V3 wh = normalized(wol + wil);
    const float cos_i = dot(wil, wh);
    const float F = dielectric_fresnel(cos_i, sc.flipped ? 1.f / m.refl.w : m.refl.w);
    const float G = G_aniso(wil, wol, m.alpha_x, m.alpha_y);
    const float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    const float f = (F * G * D) / (4.f * wol.z * wil.z);                             // scattering.cpp:84
    return diffuse + (rgb4(m.trans) * rgb4(m.k)) * f;
}
YG_INLINE float uber_pdf(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    const float diffuse_pdf = dot(sc.n, wi) / kPi;
    const V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    const V3 wh = normalized(wol + wil);
    return 0.5f * (diffuse_pdf + vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y));
}

YG_INLINE RGB bsdf_eval_base(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    if (m.type == 1) return kPiInv * rgb4(m.refl);
    if (m.type == kRoughGlassType) return rough_glass_eval(m, sc, wo, wi);
    if (m.type == 5) return plastic_eval(m, sc, wo, wi);
    if (m.type == 6) return coated_eval(m, sc, wo, wi);
    if (m.type == kUberType) return uber_eval(m, sc, wo, wi);
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    float ci = dot(wil, wh);
    RGB F = conductor_fresnel(ci, m.eta_i, rgb4(m.eta), rgb4(m.k));
    float G = G_aniso(wil, wol, m.alpha_x, m.alpha_y);
    float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    return F * ((G * D) / (4.f * wol.z * wil.z));
}
YG_INLINE float bsdf_pdf_base(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    if (m.type == 1) return dot(sc.n, wi) / kPi;
    if (m.type == 5) return plastic_pdf(m, sc, wo, wi);
    if (m.type == 6) return coated_pdf(m, sc, wo, wi);
    if (m.type == kRoughGlassType) return rough_glass_pdf(m, sc, wo, wi);
    if (m.type == kUberType) return uber_pdf(m, sc, wo, wi);
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    return vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
}
YG_INLINE V3 sample_cosine(float u0, float u1) {                                    // sampling.cpp:43
    float z = sqrtf(1.f - u0);
    float r = sqrtf(u0);
    float phi = 2.f * kPi * u1;
    return v3(r * cosf(phi), r * sinf(phi), z);
}

// u_scat selects the lobe of layered BSDFs (plastic, coated diffuse: diffuse below 0.5, specular above).
YG_INLINE RGB bsdf_sample_base(const DevMaterial& m, const Shading& sc, float u0, float u1, float u_scat, V3 wo, V3& wi,
                          float& pdf) {
    if (m.type == 1) {
        wi = to_world(sc, sample_cosine(u0, u1));
        pdf = bsdf_pdf_base(m, sc, wo, wi);
        return bsdf_eval_base(m, sc, wo, wi);
    }
    if (m.type == 5) {                                                                 // bsdf.cpp:235-251
        if (u_scat < 0.5f) {
            wi = to_world(sc, sample_cosine(u0, u1));
        } else {
            V3 wh = to_world(sc, vndf_sample(u0, u1, to_local(sc, wo), m.alpha_x, m.alpha_x));   // bsdf.cpp:41-48
            wi = reflect(wo, wh);
        }
        if (dot(sc.n, wi) <= 0.f) return rgb(0.f, 0.f, 0.f);
        pdf = plastic_pdf(m, sc, wo, wi);
        return plastic_eval(m, sc, wo, wi);
    }
    if (m.type == kRoughGlassType) {                                                   // bsdf.cpp:339-376
        float eta_o, eta_i;
        rough_glass_etas(m, sc, eta_o, eta_i);
        const V3 wh = to_world(sc, vndf_sample(u0, u1, to_local(sc, wo), m.alpha_x, m.alpha_x));
        const V3 rwi = reflect(wo, wh);
        if (dot(rwi, sc.n) <= 0.f) return rgb(0.f, 0.f, 0.f);
        const float rp = reflection_probability(m, dielectric_fresnel(dot(wo, wh), eta_i / eta_o));
        if (u_scat < rp) {
            wi = rwi;
        } else {
            if (!refract_dir(wo, wh, eta_o / eta_i, wi)) return rgb(0.f, 0.f, 0.f);
            if (dot(wi, sc.n) >= 0.f) return rgb(0.f, 0.f, 0.f);
        }
        pdf = rough_glass_pdf(m, sc, wo, wi);
        if (pdf == 0.f) return rgb(0.f, 0.f, 0.f);
        return rough_glass_eval(m, sc, wo, wi);
    }
    if (m.type == 6 || m.type == kUberType) {                      // bsdf.cpp:462-479, bsdf_pbrt.cpp:56-74
        if (u_scat < 0.5f) {
            wi = to_world(sc, sample_cosine(u0, u1));
        } else {
            const V3 wol = to_local(sc, wo);
            const V3 wh = vndf_sample(u0, u1, wol, m.alpha_x, m.alpha_y);
            wi = to_world(sc, reflect(wol, wh));
        }
        if (dot(sc.n, wi) <= 0.f) return rgb(0.f, 0.f, 0.f);
        if (m.type == kUberType) {
            pdf = uber_pdf(m, sc, wo, wi);
            return uber_eval(m, sc, wo, wi);
        }
        pdf = coated_pdf(m, sc, wo, wi);
        return coated_eval(m, sc, wo, wi);
    }
    V3 wol = to_local(sc, wo);
    V3 wh = vndf_sample(u0, u1, wol, m.alpha_x, m.alpha_y);
    V3 wil = reflect(wol, wh);
    if (wil.z <= 0.f) return rgb(0.f, 0.f, 0.f);
    wi = to_world(sc, wil);
    pdf = vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
    return bsdf_eval_base(m, sc, wo, wi);
}

// Mix_BSDF (bsdf.cpp:497-570) over two sub-materials that both have a BSDF.  Record: refl = mix_coeff
// (White - amount) | p_coeff = 1 - luminance(mix_coeff), trans = sub-material indices (values) | the
// clamped luminance of amount | whether a BSDF sub-material transmits, eta.x = index of a delta
// sub-material (-1 when none) | eta.y = the Scaled_BSDF factor.  f = (White - mix_coeff) f1 +
// mix_coeff f2 and pdf = (1 - p) pdf1 + p pdf2, each sub-BSDF counted only where it can scatter
// towards wi; a sample comes from the second BSDF with probability p.  With a delta sub-material the
// BSDF is the other sub-material's scaled by eta.y, its pdf unscaled (Scaled_BSDF, bsdf.cpp:572-629).
// With MIX false (no mix material in the scene) these are the base functions; with MIX true one
// copy of the base code serves every sub-material.
template <bool MIX>
YG_INLINE RGB bsdf_eval(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi, const DevMaterial* mats) {
    if (!MIX) return bsdf_eval_base(m, sc, wo, wi);
    const bool mix = m.type == kMixType;
    const bool scaled = mix && m.eta.x >= 0.f;
    const float ndw = dot(sc.n, wi);
    RGB f = rgb(0.f, 0.f, 0.f);
#pragma unroll 1
    for (int j = 0; j < (mix && !scaled ? 2 : 1); j++) {
        const int sub = scaled ? (m.eta.x == 0.f ? 1 : 0) : j;
        const DevMaterial b = mix ? mats[(int)(sub ? m.trans.y : m.trans.x)] : m;
        RGB fj = rgb(0.f, 0.f, 0.f);
        if (!mix || scaled || scattering_possible(b, ndw)) fj = bsdf_eval_base(b, sc, wo, wi);
        const RGB w = !mix ? rgb(1.f, 1.f, 1.f)
                           : scaled ? rgb(m.eta.y, m.eta.y, m.eta.y)
                                    : (j ? rgb4(m.refl) : rgb(1.f, 1.f, 1.f) - rgb4(m.refl));
        f = f + w * fj;
    }
    return f;
}
template <bool MIX>
YG_INLINE float bsdf_pdf(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi, const DevMaterial* mats) {
    if (!MIX) return bsdf_pdf_base(m, sc, wo, wi);
    const bool mix = m.type == kMixType;
    const bool scaled = mix && m.eta.x >= 0.f;
    const float ndw = dot(sc.n, wi);
    float pdf = 0.f;
#pragma unroll 1
    for (int j = 0; j < (mix && !scaled ? 2 : 1); j++) {
        const int sub = scaled ? (m.eta.x == 0.f ? 1 : 0) : j;
        const DevMaterial b = mix ? mats[(int)(sub ? m.trans.y : m.trans.x)] : m;
        float pj = 0.f;
        if (!mix || scaled || scattering_possible(b, ndw)) pj = bsdf_pdf_base(b, sc, wo, wi);
        const float w = !mix || scaled ? 1.f : (j ? m.refl.w : 1.f - m.refl.w);
        pdf = pdf + w * pj;
    }
    return pdf;
}
template <bool MIX>
YG_INLINE RGB bsdf_sample(const DevMaterial& m, const Shading& sc, float u0, float u1, float u_scat, V3 wo, V3& wi,
                          float& pdf, const DevMaterial* mats) {
    if (!MIX || m.type != kMixType) return bsdf_sample_base(m, sc, u0, u1, u_scat, wo, wi, pdf);
    if (m.eta.x >= 0.f) {
        const RGB f = bsdf_sample_base(mats[(int)(m.eta.x == 0.f ? m.trans.y : m.trans.x)], sc, u0, u1, u_scat, wo, wi, pdf);
        return rgb(m.eta.y, m.eta.y, m.eta.y) * f;
    }
    const float p = m.refl.w;
    const bool second = u_scat < p;
    const float us = second ? u_scat / p : (u_scat - p) / (1.f - p);
    float pdf_s = 0.f, pdf_o = 0.f;
    const RGB f_s = bsdf_sample_base(mats[(int)(second ? m.trans.y : m.trans.x)], sc, u0, u1, us, wo, wi, pdf_s);
    RGB f_o = rgb(0.f, 0.f, 0.f);
    const DevMaterial bo = mats[(int)(second ? m.trans.x : m.trans.y)];
    if (scattering_possible(bo, dot(sc.n, wi))) {
        pdf_o = bsdf_pdf_base(bo, sc, wo, wi);
        f_o = bsdf_eval_base(bo, sc, wo, wi);
    }
    const float pdf1 = second ? pdf_o : pdf_s, pdf2 = second ? pdf_s : pdf_o;
    const RGB f1 = second ? f_o : f_s, f2 = second ? f_s : f_o;
    pdf = (1.f - p) * pdf1 + p * pdf2;
    return (rgb(1.f, 1.f, 1.f) - rgb4(m.refl)) * f1 + rgb4(m.refl) * f2;
}

}  // namespace yg

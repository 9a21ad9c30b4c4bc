// scene_gen.cpp - seeded generators for the benchmark scenes, written as pbrt3 files that the
// loader (and any pbrt3 reader) accepts.  All randomness comes from one pcg32 stream seeded with
// (seed, 0x5ce7e5) so a (config, seed, scale) triple always yields the same bytes.
//
//   config 1: 5-quad box room (reflectance U[0.2,0.8] per wall) + 2,000 random triangles
//             (centroid U[0,1]^3, edge lengths U[0.01,0.1]), one 0.3x0.3 ceiling light of radiance
//             10, pinhole 60 deg, 256x256, 16 spp (4x4), maxdepth 5.
//   config 2: 400 icospheres (subdivision 4, radius U[0.05,0.5], radial 3-octave value-noise
//             displacement of amplitude 0.1 r) + a 2 km x 2 km ground grid (256x256 quads);
//             materials uniform over {diffuse, GGX conductor alpha U[0.05,0.5], smooth glass
//             IOR 1.5}; 2 rectangular area lights + 1 point light; 1024x1024, 64 spp, maxdepth 8.
//   config 4: closed 10x10x4 room of seeded furniture boxes and displaced spheres (40% glossy GGX
//             alpha U[0.1,0.3]) lit only by a 0.2x0.2 emitter of radiance 500 behind a 0.05-wide
//             slit in a partition wall; 1024x1024.
//   config 5: 2,000 copies of a displaced frequency-22 geodesic sphere (9,680 triangles) under
//             seeded rigid transforms with scale U[0.1,2], flattened; 60% GGX conductors
//             (alpha U[0.02,0.2]), 20% glass, 20% diffuse; 8 area lights of radiance U[50,500];
//             maxdepth 16, 256 spp.
// Config 3 renders the config-1 scene from many views (yar_generate_views).
// `scale` < 1 shrinks object counts / tessellation for quick tests.
#include "common.hpp"

#include <charconv>
#include <cstdio>
#include <map>
#include <string>

namespace yar {
namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t seed, uint64_t seq) {
        state = 0; inc = (seq << 1u) | 1u; next(); state += seed; next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    int below(int n) { return (int)(((uint64_t)next() * (uint64_t)n) >> 32); }
};

class Writer {
public:
    std::string s;
    void raw(const char* t) { s += t; }
    void num(float v) {
        char buf[32];
        auto r = std::to_chars(buf, buf + sizeof(buf), v);   // shortest round-trip form
        s.append(buf, r.ptr);
        s.push_back(' ');
    }
    void inum(long v) {
        char buf[24];
        auto r = std::to_chars(buf, buf + sizeof(buf), v);
        s.append(buf, r.ptr);
        s.push_back(' ');
    }
    void rgb(const char* name, F3 c) {
        s += "\"rgb "; s += name; s += "\" [ "; num(c.x); num(c.y); num(c.z); s += "] ";
    }
};

struct Mesh {
    std::vector<F3> v, n;
    std::vector<int> idx;
};

static void write_mesh(Writer& w, const Mesh& m) {
    w.raw("Shape \"trianglemesh\"\n  \"integer indices\" [ ");
    for (size_t i = 0; i < m.idx.size(); i++) { w.inum(m.idx[i]); if ((i % 24) == 23) w.raw("\n"); }
    w.raw("]\n  \"point P\" [ ");
    for (size_t i = 0; i < m.v.size(); i++) { w.num(m.v[i].x); w.num(m.v[i].y); w.num(m.v[i].z); if ((i % 8) == 7) w.raw("\n"); }
    w.raw("]\n");
    if (!m.n.empty()) {
        w.raw("  \"normal N\" [ ");
        for (size_t i = 0; i < m.n.size(); i++) { w.num(m.n[i].x); w.num(m.n[i].y); w.num(m.n[i].z); if ((i % 8) == 7) w.raw("\n"); }
        w.raw("]\n");
    }
}

static Mesh quad(F3 p0, F3 p1, F3 p2, F3 p3) {
    Mesh m;
    m.v = {p0, p1, p2, p3};
    m.idx = {0, 1, 2, 0, 2, 3};
    return m;
}

static void header(Writer& w, F3 eye, F3 at, F3 up, float fov, int res_x, int res_y, int spp_side, int maxdepth) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "LookAt %.9g %.9g %.9g  %.9g %.9g %.9g  %.9g %.9g %.9g\n", eye.x, eye.y, eye.z,
                  at.x, at.y, at.z, up.x, up.y, up.z);
    w.raw(buf);
    std::snprintf(buf, sizeof(buf),
                  "Camera \"perspective\" \"float fov\" [ %.9g ]\n"
                  "Film \"image\" \"integer xresolution\" [ %d ] \"integer yresolution\" [ %d ] \"string filename\" \"out.exr\"\n"
                  "Sampler \"stratified\" \"integer xsamples\" [ %d ] \"integer ysamples\" [ %d ] \"bool jitter\" \"true\"\n"
                  "Integrator \"path\" \"integer maxdepth\" [ %d ]\n"
                  "PixelFilter \"box\"\n"
                  "WorldBegin\n",
                  fov, res_x, res_y, spp_side, spp_side, maxdepth);
    w.raw(buf);
}

static void matte(Writer& w, F3 kd) { w.raw("Material \"matte\" "); w.rgb("Kd", kd); w.raw("\n"); }
static void metal(Writer& w, F3 eta, F3 k, float alpha) {
    w.raw("Material \"metal\" "); w.rgb("eta", eta); w.rgb("k", k);
    w.raw("\"float uroughness\" [ "); w.num(alpha); w.raw("] \"float vroughness\" [ "); w.num(alpha);
    w.raw("] \"bool remaproughness\" \"false\"\n");
}
static void glass(Writer& w, float ior) {
    w.raw("Material \"glass\" \"float index\" [ "); w.num(ior); w.raw("]\n");
}
static void area_light_quad(Writer& w, F3 L, F3 c, F3 ex, F3 ey) {
    // quad centred at c spanned by +-ex, +-ey; the loader's rectangle frame (first triangle's
    // edges) has normal cross(ex, ey), the side light sampling treats as emitting
    w.raw("AttributeBegin\nAreaLightSource \"diffuse\" "); w.rgb("L", L); w.raw("\n");
    write_mesh(w, quad(c - ex - ey, c + ex - ey, c + ex + ey, c - ex + ey));
    w.raw("AttributeEnd\n");
}

// 3-octave lattice value noise in [-1, 1] (hash of integer lattice points, smoothstep blend).
static float lattice(int x, int y, int z, uint32_t salt) {
    uint32_t h = (uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^ (uint32_t)z * 0xcb1ab31fu ^ salt;
    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
    return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}
static float value_noise(F3 p, uint32_t salt) {
    int xi = (int)std::floor(p.x), yi = (int)std::floor(p.y), zi = (int)std::floor(p.z);
    float fx = p.x - xi, fy = p.y - yi, fz = p.z - zi;
    auto sm = [](float t) { return t * t * (3.f - 2.f * t); };
    float u = sm(fx), v = sm(fy), q = sm(fz);
    float r = 0;
    for (int c = 0; c < 8; c++) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        float wgt = (dx ? u : 1 - u) * (dy ? v : 1 - v) * (dz ? q : 1 - q);
        r += wgt * lattice(xi + dx, yi + dy, zi + dz, salt);
    }
    return r;
}
static float fbm3(F3 p, uint32_t salt) {
    return 0.5714286f * (value_noise(p, salt) + 0.5f * value_noise(p * 2.f, salt + 1) + 0.25f * value_noise(p * 4.f, salt + 2));
}

// Icosphere: icosahedron with `sub` midpoint subdivisions (20 * 4^sub triangles).
static Mesh icosphere(int sub) {
    const float t = (1.f + std::sqrt(5.f)) / 2.f;
    Mesh m;
    F3 base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                   {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (F3 p : base) m.v.push_back(normalized(p));
    m.idx = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
             3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < sub; s++) {
        std::map<uint64_t, int> mid;
        auto midpoint = [&](int a, int b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            m.v.push_back(normalized((m.v[a] + m.v[b]) * 0.5f));
            int id = (int)m.v.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<int> nidx;
        nidx.reserve(m.idx.size() * 4);
        for (size_t i = 0; i < m.idx.size(); i += 3) {
            int a = m.idx[i], b = m.idx[i + 1], c = m.idx[i + 2];
            int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nidx.insert(nidx.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        m.idx.swap(nidx);
    }
    return m;
}

// Geodesic sphere of frequency f (20 f^2 triangles): each icosahedron face split into an f x f
// triangular grid, vertices projected to the unit sphere (shared edges are welded).
static Mesh geodesic(int f) {
    Mesh ico = icosphere(0);
    Mesh m;
    std::map<std::tuple<int64_t, int64_t, int64_t>, int> weld;
    auto vid = [&](F3 p) {
        F3 q = normalized(p);
        auto key = std::make_tuple((int64_t)std::llround(q.x * 1e6), (int64_t)std::llround(q.y * 1e6), (int64_t)std::llround(q.z * 1e6));
        auto it = weld.find(key);
        if (it != weld.end()) return it->second;
        m.v.push_back(q);
        int id = (int)m.v.size() - 1;
        weld[key] = id;
        return id;
    };
    for (size_t i = 0; i < ico.idx.size(); i += 3) {
        F3 A = ico.v[ico.idx[i]], B = ico.v[ico.idx[i + 1]], C = ico.v[ico.idx[i + 2]];
        std::vector<std::vector<int>> row(f + 1);
        for (int r = 0; r <= f; r++)
            for (int c = 0; c <= r; c++) {
                float a = 1.f - (float)r / f, b = (float)(r - c) / f, cc = (float)c / f;
                row[r].push_back(vid(A * a + B * b + C * cc));
            }
        for (int r = 0; r < f; r++)
            for (int c = 0; c <= r; c++) {
                m.idx.insert(m.idx.end(), {row[r][c], row[r + 1][c], row[r + 1][c + 1]});
                if (c < r) m.idx.insert(m.idx.end(), {row[r][c], row[r + 1][c + 1], row[r][c + 1]});
            }
    }
    return m;
}

static void smooth_normals(Mesh& m) {
    m.n.assign(m.v.size(), F3(0, 0, 0));
    for (size_t i = 0; i < m.idx.size(); i += 3) {
        F3 a = m.v[m.idx[i]], b = m.v[m.idx[i + 1]], c = m.v[m.idx[i + 2]];
        F3 fn = cross(b - a, c - a);
        for (int k = 0; k < 3; k++) m.n[m.idx[i + k]] = m.n[m.idx[i + k]] + fn;
    }
    for (F3& n : m.n) { float l = length(n); n = l > 0 ? n * (1.f / l) : F3(0, 1, 0); }
}

// Unit sphere mesh -> displaced sphere of radius r at center c (radial 3-octave value noise).
static Mesh displaced(const Mesh& unit, F3 c, float r, float amp, uint32_t salt, float freq) {
    Mesh m;
    m.idx = unit.idx;
    m.v.resize(unit.v.size());
    for (size_t i = 0; i < unit.v.size(); i++) {
        F3 d = unit.v[i];
        float h = fbm3(d * freq, salt);
        m.v[i] = c + d * (r + amp * r * h);
    }
    smooth_normals(m);
    return m;
}

static Mesh box_mesh(F3 lo, F3 hi) {
    Mesh m;
    for (int k = 0; k < 8; k++) m.v.push_back(F3((k & 1) ? hi.x : lo.x, (k & 2) ? hi.y : lo.y, (k & 4) ? hi.z : lo.z));
    m.idx = {0, 2, 3, 0, 3, 1, 4, 5, 7, 4, 7, 6, 0, 1, 5, 0, 5, 4, 2, 6, 7, 2, 7, 3, 0, 4, 6, 0, 6, 2, 1, 3, 7, 1, 7, 5};
    return m;
}

static void append(Mesh& dst, const Mesh& src) {
    int base = (int)dst.v.size();
    dst.v.insert(dst.v.end(), src.v.begin(), src.v.end());
    dst.n.insert(dst.n.end(), src.n.begin(), src.n.end());
    for (int i : src.idx) dst.idx.push_back(base + i);
}

static F3 random_unit(Pcg32& rng) {
    float z = rng.uniform(-1.f, 1.f);
    float phi = rng.uniform(0.f, 6.2831853f);
    float s = std::sqrt(std::max(0.f, 1.f - z * z));
    return F3(s * std::cos(phi), s * std::sin(phi), z);
}

// ---------------------------------------------------------------------------------------------
static std::string config1(uint64_t seed, double scale) {
    Pcg32 rng(seed, 0x5ce7e5);
    Writer w;
    header(w, F3(0.5f, 0.5f, 3.2f), F3(0.5f, 0.5f, 0.5f), F3(0, 1, 0), 60.f, 256, 256, 4, 5);
    const float lo = -0.5f, hi = 1.5f;
    F3 c[8];
    for (int k = 0; k < 8; k++) c[k] = F3((k & 1) ? hi : lo, (k & 2) ? hi : lo, (k & 4) ? hi : lo);
    // floor, ceiling, back, left, right (the +z side faces the camera and stays open)
    Mesh walls[5] = {quad(c[0], c[1], c[5], c[4]), quad(c[2], c[6], c[7], c[3]), quad(c[0], c[2], c[3], c[1]),
                     quad(c[0], c[4], c[6], c[2]), quad(c[1], c[3], c[7], c[5])};
    for (const Mesh& q : walls) {
        F3 kd(rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f));
        w.raw("AttributeBegin\n"); matte(w, kd); write_mesh(w, q); w.raw("AttributeEnd\n");
    }
    // ceiling light, 0.3 x 0.3, emitting downwards
    area_light_quad(w, F3(10, 10, 10), F3(0.5f, hi - 1e-3f, 0.5f), F3(0.15f, 0, 0), F3(0, 0, 0.15f));
    int ntri = std::max(1, (int)std::lround(2000 * scale));
    Mesh soup;
    for (int i = 0; i < ntri; i++) {
        F3 centroid(rng.uniform(), rng.uniform(), rng.uniform());
        F3 e1 = random_unit(rng);
        F3 tmp = random_unit(rng);
        F3 e2 = cross(e1, tmp);
        float l2 = length(e2);
        e2 = l2 > 1e-6f ? e2 * (1.f / l2) : F3(0, 1, 0);
        float a = rng.uniform(0.01f, 0.1f), b = rng.uniform(0.01f, 0.1f);
        float th = rng.uniform(0.5235988f, 2.6179939f);
        F3 p0(0, 0, 0), p1 = e1 * a, p2 = (e1 * std::cos(th) + e2 * std::sin(th)) * b;
        F3 g = (p0 + p1 + p2) * (1.f / 3.f);
        int base = (int)soup.v.size();
        soup.v.push_back(centroid + p0 - g); soup.v.push_back(centroid + p1 - g); soup.v.push_back(centroid + p2 - g);
        soup.idx.insert(soup.idx.end(), {base, base + 1, base + 2});
    }
    F3 kd(rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f));
    w.raw("AttributeBegin\n"); matte(w, kd); write_mesh(w, soup); w.raw("AttributeEnd\n");
    w.raw("WorldEnd\n");
    return w.s;
}

static void conductor_eta_k(Pcg32& rng, F3* eta, F3* k) {
    static const F3 etas[3] = {{0.143f, 0.374f, 1.442f}, {0.200f, 0.924f, 1.102f}, {2.746f, 2.958f, 2.863f}};
    static const F3 ks[3] = {{3.983f, 2.385f, 1.603f}, {3.912f, 2.452f, 2.142f}, {3.801f, 3.339f, 3.226f}};
    int i = rng.below(3);
    *eta = etas[i]; *k = ks[i];
}

static std::string config2(uint64_t seed, double scale) {
    Pcg32 rng(seed, 0x5ce7e5);
    Writer w;
    header(w, F3(0.f, 3.5f, 16.f), F3(0.f, 0.6f, 0.f), F3(0, 1, 0), 45.f, 1024, 1024, 8, 8);
    int nsph = std::max(1, (int)std::lround(400 * scale));
    int sub = scale >= 1.0 ? 4 : (scale >= 0.25 ? 3 : 2);
    Mesh unit = icosphere(sub);
    for (int i = 0; i < nsph; i++) {
        float r = rng.uniform(0.05f, 0.5f);
        F3 c(rng.uniform(-10.f, 10.f), 0.f, rng.uniform(-12.f, 6.f));
        c.y = r * rng.uniform(0.9f, 3.0f);
        uint32_t salt = rng.next();
        Mesh m = displaced(unit, c, r, 0.1f, salt, 2.f);
        int kind = rng.below(3);
        w.raw("AttributeBegin\n");
        if (kind == 0) matte(w, F3(rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)));
        else if (kind == 1) { F3 eta, k; conductor_eta_k(rng, &eta, &k); metal(w, eta, k, rng.uniform(0.05f, 0.5f)); }
        else glass(w, 1.5f);
        write_mesh(w, m);
        w.raw("AttributeEnd\n");
    }
    // 2 km x 2 km ground grid at y = 0
    int g = std::max(2, (int)std::lround(256 * std::sqrt(scale)));
    Mesh ground;
    for (int j = 0; j <= g; j++)
        for (int i = 0; i <= g; i++)
            ground.v.push_back(F3(-1000.f + 2000.f * i / g, 0.f, -1000.f + 2000.f * j / g));
    for (int j = 0; j < g; j++)
        for (int i = 0; i < g; i++) {
            int a = j * (g + 1) + i, b = a + 1, c = a + (g + 1), d = c + 1;
            ground.idx.insert(ground.idx.end(), {a, c, d, a, d, b});
        }
    w.raw("AttributeBegin\n"); matte(w, F3(0.5f, 0.5f, 0.5f)); write_mesh(w, ground); w.raw("AttributeEnd\n");
    area_light_quad(w, F3(12, 12, 11), F3(-4.f, 7.f, -2.f), F3(1.f, 0, 0), F3(0, 0, 1.f));
    area_light_quad(w, F3(9, 10, 12), F3(5.f, 6.f, 2.f), F3(0.8f, 0, 0), F3(0, 0, 0.8f));
    w.raw("LightSource \"point\" \"rgb I\" [ 40 38 34 ] \"point from\" [ 0 5 8 ]\n");
    w.raw("WorldEnd\n");
    return w.s;
}

static std::string config4(uint64_t seed, double scale) {
    Pcg32 rng(seed, 0x5ce7e5);
    Writer w;
    // room [0,10] x [0,4] x [0,10] (y up); partition wall at z = 8 with a 0.05-wide vertical slit;
    // the emitter sits behind the partition facing it, so all light enters through the slit.
    header(w, F3(5.f, 1.8f, 0.3f), F3(5.f, 1.4f, 6.f), F3(0, 1, 0), 70.f, 1024, 1024, 8, 8);
    auto wall = [&](F3 a, F3 b, F3 c, F3 d, F3 kd) {
        w.raw("AttributeBegin\n"); matte(w, kd); write_mesh(w, quad(a, b, c, d)); w.raw("AttributeEnd\n");
    };
    F3 wkd(0.7f, 0.7f, 0.7f);
    wall(F3(0, 0, 0), F3(0, 0, 10), F3(10, 0, 10), F3(10, 0, 0), wkd);         // floor (faces +y)
    wall(F3(0, 4, 0), F3(10, 4, 0), F3(10, 4, 10), F3(0, 4, 10), wkd);         // ceiling
    wall(F3(0, 0, 0), F3(10, 0, 0), F3(10, 4, 0), F3(0, 4, 0), wkd);           // z = 0
    wall(F3(0, 0, 10), F3(0, 4, 10), F3(10, 4, 10), F3(10, 0, 10), wkd);       // z = 10
    wall(F3(0, 0, 0), F3(0, 4, 0), F3(0, 4, 10), F3(0, 0, 10), wkd);           // x = 0
    wall(F3(10, 0, 0), F3(10, 0, 10), F3(10, 4, 10), F3(10, 4, 0), wkd);       // x = 10
    const float zw = 8.f, slit0 = 4.975f, slit1 = 5.025f;
    wall(F3(0, 0, zw), F3(slit0, 0, zw), F3(slit0, 4, zw), F3(0, 4, zw), wkd);   // partition, left part
    wall(F3(slit1, 0, zw), F3(10, 0, zw), F3(10, 4, zw), F3(slit1, 4, zw), wkd); // partition, right part
    // emitter behind the partition, facing -z towards the slit
    area_light_quad(w, F3(500, 500, 500), F3(5.f, 2.f, 9.6f), F3(-0.1f, 0, 0), F3(0, 0.1f, 0));
    // furniture boxes and displaced spheres in the front part of the room (z < 7.8)
    long target = (long)(500000 * scale);
    Mesh unit = icosphere(scale >= 0.5 ? 4 : 2);
    long tris = 0;
    int obj = 0;
    while (tris < target) {
        bool sphere = (obj % 3) == 2;
        Mesh m;
        if (sphere) {
            float r = rng.uniform(0.15f, 0.6f);
            F3 c(rng.uniform(0.8f, 9.2f), rng.uniform(r, 3.0f), rng.uniform(0.8f, 7.4f));
            m = displaced(unit, c, r, 0.1f, rng.next(), 2.f);
        } else {
            F3 sz(rng.uniform(0.2f, 1.2f), rng.uniform(0.1f, 1.0f), rng.uniform(0.2f, 1.2f));
            F3 lo(rng.uniform(0.1f, 9.9f - sz.x), rng.uniform(0.f, 2.5f), rng.uniform(0.1f, 7.7f - sz.z));
            // tessellate boxes finely so triangle counts reach the target: 8x8 grid per face
            Mesh b = box_mesh(lo, lo + sz);
            Mesh fine;
            for (size_t t = 0; t < b.idx.size(); t += 6) {
                F3 p0 = b.v[b.idx[t]], p1 = b.v[b.idx[t + 1]], p2 = b.v[b.idx[t + 2]];
                F3 e1 = p1 - p0, e2 = p2 - p1;
                const int n = 8;
                int base = (int)fine.v.size();
                for (int j = 0; j <= n; j++)
                    for (int i = 0; i <= n; i++) fine.v.push_back(p0 + e1 * ((float)i / n) + e2 * ((float)j / n));
                for (int j = 0; j < n; j++)
                    for (int i = 0; i < n; i++) {
                        int a = base + j * (n + 1) + i, bb = a + 1, c = a + n + 1, d = c + 1;
                        fine.idx.insert(fine.idx.end(), {a, bb, d, a, d, c});
                    }
            }
            m = fine;
        }
        bool glossy = rng.uniform() < 0.4f;
        w.raw("AttributeBegin\n");
        if (glossy) { F3 eta, k; conductor_eta_k(rng, &eta, &k); metal(w, eta, k, rng.uniform(0.1f, 0.3f)); }
        else matte(w, F3(rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)));
        write_mesh(w, m);
        w.raw("AttributeEnd\n");
        tris += (long)m.idx.size() / 3;
        obj++;
    }
    w.raw("WorldEnd\n");
    return w.s;
}

static std::string config5(uint64_t seed, double scale) {
    Pcg32 rng(seed, 0x5ce7e5);
    Writer w;
    header(w, F3(0.f, 12.f, 60.f), F3(0.f, 2.f, 0.f), F3(0, 1, 0), 50.f, 1024, 1024, 16, 16);
    int copies = std::max(1, (int)std::lround(2000 * scale));
    int freq = scale >= 1.0 ? 22 : (scale >= 0.1 ? 11 : 5);
    Mesh unit = geodesic(freq);
    Mesh base = displaced(unit, F3(0, 0, 0), 1.f, 0.1f, 0x1234567u, 2.f);
    for (int i = 0; i < copies; i++) {
        float s = rng.uniform(0.1f, 2.f);
        F3 axis = random_unit(rng);
        float ang = rng.uniform(0.f, 6.2831853f);
        F3 t(rng.uniform(-40.f, 40.f), rng.uniform(s, 8.f), rng.uniform(-40.f, 40.f));
        // Rodrigues rotation, then scale and translate; instances are flattened into world space
        float ca = std::cos(ang), sa = std::sin(ang);
        Mesh m;
        m.idx = base.idx;
        m.v.resize(base.v.size());
        m.n.resize(base.n.size());
        auto rot = [&](F3 v) { return v * ca + cross(axis, v) * sa + axis * (dot(axis, v) * (1.f - ca)); };
        for (size_t k = 0; k < base.v.size(); k++) { m.v[k] = t + rot(base.v[k]) * s; m.n[k] = rot(base.n[k]); }
        float u = rng.uniform();
        w.raw("AttributeBegin\n");
        if (u < 0.6f) { F3 eta, k; conductor_eta_k(rng, &eta, &k); metal(w, eta, k, rng.uniform(0.02f, 0.2f)); }
        else if (u < 0.8f) glass(w, 1.5f);
        else matte(w, F3(rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)));
        write_mesh(w, m);
        w.raw("AttributeEnd\n");
    }
    w.raw("AttributeBegin\n"); matte(w, F3(0.5f, 0.5f, 0.5f));
    write_mesh(w, quad(F3(-200, 0, -200), F3(-200, 0, 200), F3(200, 0, 200), F3(200, 0, -200)));
    w.raw("AttributeEnd\n");
    for (int i = 0; i < 8; i++) {
        float L = rng.uniform(50.f, 500.f);
        F3 c(rng.uniform(-30.f, 30.f), rng.uniform(14.f, 20.f), rng.uniform(-30.f, 30.f));
        float hs = rng.uniform(0.3f, 1.0f);
        area_light_quad(w, F3(L, L, L), c, F3(hs, 0, 0), F3(0, 0, hs));
    }
    w.raw("WorldEnd\n");
    return w.s;
}

}  // namespace
}  // namespace yar

extern "C" int yar_scene_generate(int32_t config, uint64_t seed, double scale, const char* pbrt_path) {
    using namespace yar;
    if (!pbrt_path || !(scale > 0.0) || scale > 1.0) {
        set_last_error("yar_scene_generate: invalid argument");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    if (config < 1 || config > 5) {
        set_last_error("yar_scene_generate: config must be 1..5");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    return abi_guard("yar_scene_generate", [&] {
        std::string text;
        switch (config) {
            case 1: case 3: text = config1(seed, scale); break;
            case 2: text = config2(seed, scale); break;
            case 4: text = config4(seed, scale); break;
            default: text = config5(seed, scale); break;
        }
        FILE* f = std::fopen(pbrt_path, "wb");
        if (!f) throw Error{YAR_ERR_IO, std::string("cannot write ") + pbrt_path};
        bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw Error{YAR_ERR_IO, std::string("write failed: ") + pbrt_path};
    });
}

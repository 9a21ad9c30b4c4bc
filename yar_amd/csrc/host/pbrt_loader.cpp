// pbrt_loader.cpp - loader for the pbrt3 subset consumed by the RAY path tracer.
//
// Import semantics follow the reference (kennyalive/YAR):
//   load_pbrt_scene             src/lib/pbrt_scene.cpp:1050-1183
//   import_pbrt_triangle_mesh   src/lib/pbrt_scene.cpp:676-743 (default UVs for 3/4-vertex meshes,
//                               remove_degenerate_triangles src/lib/triangle_mesh.cpp:5)
//   check_if_mesh_is_rectangle  src/lib/pbrt_scene.cpp:45-121 (rectangular area lights)
//   import_pbrt_material        src/lib/pbrt_scene.cpp:418-501 (matte, metal, glass)
//   import_pbrt_non_area_light  src/lib/pbrt_scene.cpp:903-956 (point, spot, distant lights)
//   import_pbrt_sphere          src/lib/pbrt_scene.cpp:745-758, 853-887 (tessellated spheres, sphere lights;
//                               create_sphere_mesh src/lib/tessellation.cpp:43-171)
//   import_pbrt_camera          src/lib/pbrt_scene.cpp:985-1045
//   finalize_scene              src/lib/scene_loader.cpp:56-72
// and the parameter defaults of the vendored pbrt-parser fork (third-party/pbrt-parser:
// impl/semantic/{Camera,Sampler,Integrator,Materials}.cpp, impl/syntactic/Parser.inl:695 LookAt).
#include "common.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <unordered_map>

namespace yar {
namespace {

struct ParseError {
    std::string msg;
};

// ---- tokenizer over a memory buffer ---------------------------------------------------------
struct Lexer {
    const char* p;
    const char* end;
    int line = 1;

    bool skip_ws() {
        while (p < end) {
            char c = *p;
            if (c == '\n') { line++; p++; }
            else if (c == ' ' || c == '\t' || c == '\r') p++;
            else if (c == '#') { while (p < end && *p != '\n') p++; }
            else return true;
        }
        return false;
    }
    // Token kinds: 's' quoted string, '[' ']', 'n' number, 'i' identifier, 0 end.
    char peek_kind() {
        if (!skip_ws()) return 0;
        char c = *p;
        if (c == '"') return 's';
        if (c == '[' || c == ']') return c;
        if (c == '-' || c == '+' || c == '.' || std::isdigit((unsigned char)c)) return 'n';
        return 'i';
    }
    std::string ident() {
        skip_ws();
        const char* s = p;
        while (p < end && !std::isspace((unsigned char)*p) && *p != '"' && *p != '[' && *p != ']' && *p != '#') p++;
        return std::string(s, p);
    }
    std::string string() {
        skip_ws();
        if (p >= end || *p != '"') throw ParseError{"expected quoted string at line " + std::to_string(line)};
        const char* s = ++p;
        while (p < end && *p != '"') { if (*p == '\n') line++; p++; }
        if (p >= end) throw ParseError{"unterminated string at line " + std::to_string(line)};
        std::string r(s, p);
        p++;
        return r;
    }
    double number() { float f; return number2(&f); }
    // Parses one number both as double (exact for integers) and directly as float (strtof, the
    // correctly rounded conversion the pbrt-parser's std::stof performs).
    double number2(float* f) {
        skip_ws();
        char* e = nullptr;
        double v = std::strtod(p, &e);
        if (e == p) throw ParseError{"expected number at line " + std::to_string(line)};
        *f = std::strtof(p, nullptr);
        p = e;
        return v;
    }
    void expect(char c) {
        skip_ws();
        if (p >= end || *p != c) throw ParseError{std::string("expected '") + c + "' at line " + std::to_string(line)};
        p++;
    }
};

// One "type name" [values] parameter.
struct Param {
    std::string type, name;
    std::vector<double> nums;     // exact for integer parameters
    std::vector<float> fnums;     // float parameters, parsed with strtof
    std::vector<std::string> strs;
};
using ParamList = std::vector<Param>;

static ParamList parse_params(Lexer& lx) {
    ParamList out;
    while (lx.peek_kind() == 's') {
        // A parameter declaration is a quoted "type name"; a statement argument is never
        // followed by a value list in this subset, so any quoted token here starts a parameter.
        const char* save = lx.p;
        int save_line = lx.line;
        std::string decl = lx.string();
        std::istringstream ss(decl);
        Param prm;
        ss >> prm.type >> prm.name;
        if (prm.name.empty()) { lx.p = save; lx.line = save_line; break; }
        char k = lx.peek_kind();
        bool bracket = (k == '[');
        if (bracket) lx.expect('[');
        auto read_one = [&]() {
            char kk = lx.peek_kind();
            if (kk == 's') prm.strs.push_back(lx.string());
            else if (kk == 'n') { float f; prm.nums.push_back(lx.number2(&f)); prm.fnums.push_back(f); }
            else if (kk == 'i') {   // bare true/false
                std::string id = lx.ident();
                if (id == "true") { prm.nums.push_back(1); prm.fnums.push_back(1); }
                else if (id == "false") { prm.nums.push_back(0); prm.fnums.push_back(0); }
                else throw ParseError{"unexpected token '" + id + "' in parameter " + prm.name};
            } else throw ParseError{"bad parameter value for " + prm.name};
        };
        if (bracket) {
            while (lx.peek_kind() != ']') {
                if (lx.peek_kind() == 0) throw ParseError{"unterminated parameter list for " + prm.name};
                read_one();
            }
            lx.expect(']');
        } else {
            read_one();
        }
        out.push_back(std::move(prm));
    }
    return out;
}

static const Param* find(const ParamList& ps, const char* name) {
    for (const Param& p : ps) if (p.name == name) return &p;
    return nullptr;
}
static float get_f(const ParamList& ps, const char* name, float def) {
    const Param* p = find(ps, name);
    return (p && !p->fnums.empty()) ? p->fnums[0] : def;
}
static int get_i(const ParamList& ps, const char* name, int def) {
    const Param* p = find(ps, name);
    return (p && !p->nums.empty()) ? (int)p->nums[0] : def;
}
static bool get_b(const ParamList& ps, const char* name, bool def) {
    const Param* p = find(ps, name);
    if (!p) return def;
    if (!p->strs.empty()) return p->strs[0] == "true";
    return !p->nums.empty() ? p->nums[0] != 0 : def;
}
static bool get_rgb(const ParamList& ps, const char* name, float out[3]) {
    const Param* p = find(ps, name);
    if (!p) return false;
    if (p->type == "spectrum" || p->type == "blackbody")
        throw ParseError{std::string("spectral parameter '") + name + "' is not supported (use rgb)"};
    if (p->fnums.size() == 1) { out[0] = out[1] = out[2] = p->fnums[0]; return true; }
    if (p->fnums.size() < 3) throw ParseError{std::string("parameter '") + name + "' needs 3 values"};
    for (int i = 0; i < 3; i++) out[i] = p->fnums[i];
    return true;
}

struct MaterialState {
    std::string type = "";        // "" = pbrt default material (diffuse 0.5)
    ParamList params;
    int id = -1;                  // index into SceneData::materials once imported
};

struct GraphicsState {
    M34 ctm = M34::identity();
    std::shared_ptr<MaterialState> material;
    std::shared_ptr<ParamList> area_light;   // AreaLightSource "diffuse" params
    bool reverse_orientation = false;
};

// pbrt-parser bakes a shape's current transform into its mesh (third-party/pbrt-parser/impl/semantic/
// Geometry.cpp:248-252): points by xfmPoint = l * p + t, normals by xfmNormal = inverse_transpose(l) * n
// with inverse_transpose(l) = adjoint_transpose(l) * (1 / det l), where l * v sums the columns of l
// weighted by v.x, v.y, v.z in that order (include/pbrtParser/math.h:146-168).  Normals stay unnormalised.
static void bake_transform(const M34& m, std::vector<float>& P, std::vector<float>& N) {
    for (size_t i = 0; i + 2 < P.size(); i += 3) {
        const F3 r = transform_point(m, F3(P[i], P[i + 1], P[i + 2]));
        P[i] = r.x; P[i + 1] = r.y; P[i + 2] = r.z;
    }
    if (N.empty()) return;
    const F3 vx = m.column(0), vy = m.column(1), vz = m.column(2);
    const float inv_det = 1.f / dot(vx, cross(vy, vz));
    const F3 c0 = cross(vy, vz) * inv_det, c1 = cross(vz, vx) * inv_det, c2 = cross(vx, vy) * inv_det;
    for (size_t i = 0; i + 2 < N.size(); i += 3) {
        const F3 n = (c0 * N[i] + c1 * N[i + 1]) + c2 * N[i + 2];
        N[i] = n.x; N[i + 1] = n.y; N[i + 2] = n.z;
    }
}

// pbrt-parser's affine3f::rotate (include/pbrtParser/math.h:170-177), angle in radians.
static M34 rotation(F3 axis, float r) {
    const F3 u = normalized(axis);
    const float s = std::sin(r), c = std::cos(r);
    const F3 vx(u.x * u.x + (1 - u.x * u.x) * c, u.x * u.y * (1 - c) + u.z * s, u.x * u.z * (1 - c) - u.y * s);
    const F3 vy(u.x * u.y * (1 - c) - u.z * s, u.y * u.y + (1 - u.y * u.y) * c, u.y * u.z * (1 - c) + u.x * s);
    const F3 vz(u.x * u.z * (1 - c) + u.y * s, u.y * u.z * (1 - c) - u.x * s, u.z * u.z + (1 - u.z * u.z) * c);
    M34 m{};
    m.set_column(0, vx); m.set_column(1, vy); m.set_column(2, vz); m.set_column(3, F3(0, 0, 0));
    return m;
}

// check_if_mesh_is_rectangle, src/lib/pbrt_scene.cpp:45-121
static bool mesh_is_rectangle(const MeshData& mesh, float size[2], M34& xf);
static bool mesh_is_rectangle_light(const MeshData& mesh) {
    float size[2];
    M34 xf{};
    return mesh_is_rectangle(mesh, size, xf);
}
static bool mesh_is_rectangle(const MeshData& mesh, float size[2], M34& xf) {
    size_t nv = mesh.vertices.size() / 3;
    if (!((nv == 4 || nv == 6) && mesh.indices.size() == 6)) return false;
    auto V = [&](int i) { const float* q = &mesh.vertices[3 * (size_t)mesh.indices[i]]; return F3(q[0], q[1], q[2]); };
    F3 p[3] = {V(0), V(1), V(2)};
    bool found = false;
    F3 p3;
    for (int i = 3; i < 6; i++) {
        F3 pp = V(i);
        if (length(pp - p[0]) < 1e-4f || length(pp - p[1]) < 1e-4f || length(pp - p[2]) < 1e-4f) continue;
        if (found) return false;
        found = true;
        p3 = pp;
    }
    if (!found) return false;
    F3 v[3] = {p[1] - p[0], p[2] - p[1], p[0] - p[2]};
    F3 d[3] = {normalized(v[0]), normalized(v[1]), normalized(v[2])};
    int k = 0;
    for (; k < 3; k++) if (std::fabs(dot(d[k], d[(k + 1) % 3])) < 1e-4f) break;
    if (k == 3) return false;
    F3 mid = (p[0] + p[1] + p[2] + p3) * 0.25f;
    F3 test = (p[k] + p[(k + 2) % 3]) * 0.5f;
    if (length(mid - test) > 1e-4f) return false;
    F3 x_axis = d[k], y_axis = d[(k + 1) % 3];
    F3 z_axis = cross(x_axis, y_axis);
    if (!mesh.normals.empty()) {
        F3 n0(mesh.normals[0], mesh.normals[1], mesh.normals[2]);
        if (dot(n0, z_axis) < 0.f) { x_axis = -x_axis; z_axis = -z_axis; }
    }
    size[0] = length(v[k]);
    size[1] = length(v[(k + 1) % 3]);
    xf.set_column(0, x_axis); xf.set_column(1, y_axis); xf.set_column(2, z_axis); xf.set_column(3, mid);
    return true;
}

// The sphere of create_sphere_mesh (src/lib/tessellation.cpp:43-171).  A geodesic grid: the
// unit-circumradius icosahedron (vertex and face tables of geometrictools.com PlatonicSolids.pdf,
// in the reference's order), each face split `levels` times into four with the edge midpoints
// pushed out to the unit sphere.  Every level renumbers vertices by first use while it walks the
// faces in order, visiting each face's corners and then the midpoints of edges 01, 12 and 02, and
// identical positions share one vertex.  Texture coordinates come from the unit position (v = 0 at
// the -z pole), then positions are scaled by the radius and get area-weighted vertex normals
// (calculate_normals with default parameters, triangle_mesh.cpp:285-345).
static MeshData geodesic_sphere(float radius, int levels) {
    const float g = (1.f + std::sqrt(5.f)) / 2.f;
    const float k = 1.f / std::sqrt(1.f + g * g);
    std::vector<F3> pos = {{g, 1, 0}, {-g, 1, 0}, {g, -1, 0}, {-g, -1, 0}, {1, 0, g}, {1, 0, -g},
                           {-1, 0, g}, {-1, 0, -g}, {0, g, 1}, {0, -g, 1}, {0, g, -1}, {0, -g, -1}};
    for (F3& p : pos) p = p * k;
    std::vector<int32_t> idx = {0, 8, 4, 1, 10, 7, 2, 9, 11, 7, 3, 1, 0, 5, 10, 3, 9, 6, 3, 11, 9, 8, 6, 4,
                                2, 4, 9, 3, 7, 11, 4, 2, 0, 9, 4, 6, 2, 11, 5, 0, 10, 8, 5, 0, 2, 10, 5, 7,
                                1, 6, 8, 1, 8, 10, 6, 1, 3, 11, 7, 5};
    struct Key {
        uint32_t x, y, z;
        bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash {
        size_t operator()(const Key& q) const { return ((size_t)q.x * 0x9E3779B97F4A7C15ull) ^ ((size_t)q.y << 21) ^ q.z; }
    };
    auto bits = [](float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; };
    for (int level = 0; level < levels; level++) {
        std::vector<F3> next_pos;
        std::vector<int32_t> next_idx;
        next_idx.reserve(idx.size() * 4);
        // positions are never -0 or NaN here, so equal bit patterns are equal positions
        std::unordered_map<Key, int32_t, KeyHash> first_use;
        auto vertex = [&](F3 p) {
            auto [it, fresh] = first_use.try_emplace(Key{bits(p.x), bits(p.y), bits(p.z)}, (int32_t)next_pos.size());
            if (fresh) next_pos.push_back(p);
            return it->second;
        };
        for (size_t f = 0; f < idx.size(); f += 3) {
            const F3 a = pos[idx[f]], b = pos[idx[f + 1]], c = pos[idx[f + 2]];
            const F3 ab = normalized(a + b), bc = normalized(b + c), ac = normalized(a + c);
            const int32_t ia = vertex(a);
            const int32_t ib = vertex(b);
            const int32_t ic = vertex(c);
            const int32_t iab = vertex(ab);
            const int32_t ibc = vertex(bc);
            const int32_t iac = vertex(ac);
            next_idx.insert(next_idx.end(), {ia, iab, iac, ib, ibc, iab, ic, iac, ibc, iab, ibc, iac});
        }
        pos.swap(next_pos);
        idx.swap(next_idx);
    }
    MeshData m;
    m.uvs.reserve(2 * pos.size());
    for (const F3& p : pos) {
        const float cos_theta = std::clamp(-p.z, -1.f, 1.f);
        const float v = std::clamp(std::acos(cos_theta) / kPi, 0.f, kOneMinusEpsilon);
        float phi = std::atan2(p.y, p.x);
        if (phi < 0) phi += kPi2;
        const float u = std::clamp(phi / kPi2, 0.f, kOneMinusEpsilon);
        m.uvs.push_back(u);
        m.uvs.push_back(v);
    }
    for (F3& p : pos) p = F3(p.x * radius, p.y * radius, p.z * radius);
    std::vector<F3> nrm(pos.size());
    for (size_t f = 0; f < idx.size(); f += 3) {
        const F3 a = pos[idx[f]], b = pos[idx[f + 1]], c = pos[idx[f + 2]];
        const F3 n = cross(b - a, c - a);
        for (int j = 0; j < 3; j++) nrm[idx[f + j]] = nrm[idx[f + j]] + n;
    }
    m.vertices.reserve(3 * pos.size());
    m.normals.reserve(3 * pos.size());
    for (size_t i = 0; i < pos.size(); i++) {
        F3 n = nrm[i];
        if (n.x == 0.f && n.y == 0.f && n.z == 0.f) n = F3(0, 0, 1);
        const float l = length(n);
        m.vertices.insert(m.vertices.end(), {pos[i].x, pos[i].y, pos[i].z});
        m.normals.insert(m.normals.end(), {n.x / l, n.y / l, n.z / l});
    }
    m.indices = std::move(idx);
    return m;
}

struct Loader {
    SceneData& sd;
    std::vector<GraphicsState> stack;
    GraphicsState gs;
    std::map<std::string, std::shared_ptr<MaterialState>> named;
    std::map<float, int32_t> sphere_meshes;       // tessellation per radius (pbrt_scene.cpp:748-756)
    std::vector<std::pair<int32_t, int32_t>> sphere_objects;   // (object, sphere light)
    // camera/film/sampler/integrator state
    bool have_camera = false;
    M34 camera_ctm = M34::identity();
    float fov = 90.f;
    bool z_is_up = false;
    bool have_film = false;
    int xres = 640, yres = 480;      // pbrt defaults
    float crop[4] = {0, 1, 0, 1};
    bool have_crop = false;
    float film_scale = 1.f, max_component = kInf;
    int xs = 1, ys = 1;
    bool have_integrator = false;
    int maxdepth = 5;
    float rr_threshold = 1.f;
    float filter_radius = 0.5f;

    explicit Loader(SceneData& s) : sd(s) {}

    void transform(const M34& x) { gs.ctm = mul(gs.ctm, x); }

    int import_material(const std::shared_ptr<MaterialState>& ms) {
        if (ms && ms->id >= 0) return ms->id;
        yar_material_desc m{};
        m.eta_i = 1.f;
        if (!ms || ms->type.empty()) {                                     // default pbrt material
            m.type = YAR_MATERIAL_DIFFUSE;
            m.reflectance[0] = m.reflectance[1] = m.reflectance[2] = 0.5f;
        } else if (ms->type == "matte") {
            m.type = YAR_MATERIAL_DIFFUSE;
            m.reflectance[0] = m.reflectance[1] = m.reflectance[2] = 0.5f;
            if (find(ms->params, "Kd") && find(ms->params, "Kd")->type == "texture")
                throw ParseError{"textured matte Kd is not supported"};
            get_rgb(ms->params, "Kd", m.reflectance);
        } else if (ms->type == "metal") {
            m.type = YAR_MATERIAL_METAL;
            float r = get_f(ms->params, "roughness", 0.01f);
            m.u_roughness = get_f(ms->params, "uroughness", r);
            m.v_roughness = get_f(ms->params, "vroughness", find(ms->params, "roughness") ? r : 0.01f);
            m.roughness_is_alpha = get_b(ms->params, "remaproughness", true) ? 0 : 1;
            if (!get_rgb(ms->params, "eta", m.eta) || !get_rgb(ms->params, "k", m.k))
                throw ParseError{"metal without rgb eta/k (copper spectrum default) is not supported"};
        } else if (ms->type == "glass") {
            m.type = YAR_MATERIAL_GLASS;
            for (int i = 0; i < 3; i++) m.reflectance[i] = m.transmittance[i] = 1.f;
            get_rgb(ms->params, "Kr", m.reflectance);
            get_rgb(ms->params, "Kt", m.transmittance);
            m.index_of_refraction = get_f(ms->params, "index", 1.5f);
            m.u_roughness = get_f(ms->params, "uroughness", 0.f);
            m.v_roughness = get_f(ms->params, "vroughness", 0.f);
            m.roughness_is_alpha = get_b(ms->params, "remaproughness", true) ? 0 : 1;
            if (m.u_roughness != m.v_roughness)                         // pbrt_scene.cpp:450
                throw ParseError{"glass with different u and v roughness is not supported"};
        } else if (ms->type == "mirror") {                                 // pbrt_scene.cpp:443-447
            m.type = YAR_MATERIAL_MIRROR;
            m.reflectance[0] = m.reflectance[1] = m.reflectance[2] = 0.9f;    // pbrt-parser MirrorMaterial kr
            if (find(ms->params, "Kr") && find(ms->params, "Kr")->type == "texture")
                throw ParseError{"textured mirror Kr is not supported"};
            get_rgb(ms->params, "Kr", m.reflectance);
        } else if (ms->type == "plastic") {                                // pbrt_scene.cpp:503-525
            m.type = YAR_MATERIAL_PLASTIC;
            float kd[3] = {0.25f, 0.25f, 0.25f}, ks[3] = {0.25f, 0.25f, 0.25f};   // pbrt-parser PlasticMaterial
            for (const char* name : {"Kd", "Ks", "roughness"})
                if (find(ms->params, name) && find(ms->params, name)->type == "texture")
                    throw ParseError{std::string("textured plastic ") + name + " is not supported"};
            get_rgb(ms->params, "Kd", kd);
            get_rgb(ms->params, "Ks", ks);
            for (int i = 0; i < 3; i++) m.reflectance[i] = kd[i];
            m.specular_r0 = 0.2126729f * ks[0] + 0.7151522f * ks[1] + 0.0721750f * ks[2];   // Y of sRGB_to_XYZ(Ks)
            m.u_roughness = m.v_roughness = get_f(ms->params, "roughness", 0.1f);
            m.roughness_is_alpha = get_b(ms->params, "remaproughness", true) ? 0 : 1;
        } else if (ms->type == "substrate") {                              // pbrt_scene.cpp:531-561
            m.type = YAR_MATERIAL_COATED_DIFFUSE;
            float kd[3] = {0.5f, 0.5f, 0.5f}, ks[3] = {0.5f, 0.5f, 0.5f};     // pbrt-parser SubstrateMaterial
            for (const char* name : {"Kd", "Ks", "uroughness", "vroughness", "bumpmap"})
                if (find(ms->params, name) && find(ms->params, name)->type == "texture")
                    throw ParseError{std::string("textured substrate ") + name + " is not supported"};
            get_rgb(ms->params, "Kd", kd);
            get_rgb(ms->params, "Ks", ks);
            for (int i = 0; i < 3; i++) { m.reflectance[i] = kd[i]; m.specular[i] = ks[i]; }
            m.u_roughness = get_f(ms->params, "uroughness", 0.1f);
            m.v_roughness = get_f(ms->params, "vroughness", 0.1f);
            m.roughness_is_alpha = get_b(ms->params, "remaproughness", true) ? 0 : 1;
        } else if (ms->type == "uber") {                                   // pbrt_scene.cpp:571-636
            m.type = YAR_MATERIAL_PBRT3_UBER;
            for (const Param& e : ms->params) {                               // Materials.cpp createMaterial_uber
                static const char* known[] = {"type", "Kd", "Ks", "Kr", "Kt", "opacity", "index", "eta", "roughness",
                                              "uroughness", "vroughness", "alpha", "shadowalpha"};
                bool ok = false;
                for (const char* k : known) ok = ok || e.name == k;
                if (!ok) throw ParseError{"unsupported uber parameter '" + e.name + "'"};
                if (e.type == "texture") throw ParseError{"textured uber " + e.name + " is not supported"};
            }
            float kd[3] = {0.25f, 0.25f, 0.25f}, ks[3] = {0.25f, 0.25f, 0.25f};   // pbrt-parser UberMaterial
            float kr[3] = {0, 0, 0}, kt[3] = {0, 0, 0}, opacity[3] = {1, 1, 1};
            get_rgb(ms->params, "Kd", kd);
            get_rgb(ms->params, "Ks", ks);
            get_rgb(ms->params, "Kr", kr);
            get_rgb(ms->params, "Kt", kt);
            get_rgb(ms->params, "opacity", opacity);
            if (kt[0] != 0.f || kt[1] != 0.f || kt[2] != 0.f)                 // delta_scattering.cpp:132-134
                throw ParseError{"uber delta transmission (Kt) is not supported"};
            if (get_f(ms->params, "alpha", 0.f) != 0.f || get_f(ms->params, "shadowalpha", 0.f) != 0.f)
                throw ParseError{"uber alpha is not supported"};
            for (int i = 0; i < 3; i++) {
                m.reflectance[i] = kd[i]; m.specular[i] = ks[i];
                m.delta_reflectance[i] = kr[i]; m.opacity[i] = opacity[i];
            }
            m.index_of_refraction = find(ms->params, "index") ? get_f(ms->params, "index", 1.5f)
                                                              : get_f(ms->params, "eta", 1.5f);
            const float r = get_f(ms->params, "roughness", 0.1f);
            m.u_roughness = get_f(ms->params, "uroughness", r);
            m.v_roughness = find(ms->params, "vroughness") ? get_f(ms->params, "vroughness", r) : m.u_roughness;
            m.roughness_is_alpha = 0;
            auto black = [](const float* c) { return c[0] == 0.f && c[1] == 0.f && c[2] == 0.f; };
            if (black(kd) && black(ks) && black(kr) && opacity[0] == 1.f && opacity[1] == 1.f && opacity[2] == 1.f)
                throw ParseError{"uber material with no scattering component"};
        } else if (ms->type == "mix") {                                    // pbrt_scene.cpp:563-569
            m.type = YAR_MATERIAL_MIX;                                         // Materials.cpp:376-402
            m.mix_amount[0] = m.mix_amount[1] = m.mix_amount[2] = 0.5f;
            if (find(ms->params, "amount") && find(ms->params, "amount")->type == "texture")
                throw ParseError{"textured mix amount is not supported"};
            if (find(ms->params, "amount") && find(ms->params, "amount")->fnums.size() != 3)
                throw ParseError{"mix amount needs 3 components"};
            get_rgb(ms->params, "amount", m.mix_amount);
            int delta_subs = 0;
            for (int j = 0; j < 2; j++) {
                const char* key = j == 0 ? "namedmaterial1" : "namedmaterial2";
                const Param* pn = find(ms->params, key);
                if (!pn || pn->strs.empty()) throw ParseError{std::string("mix material without '") + key + "'"};
                auto it = named.find(pn->strs[0]);
                if (it == named.end()) throw ParseError{"unknown named material '" + pn->strs[0] + "'"};
                const std::string& t = it->second->type;
                if (t == "mix" || t == "uber")
                    throw ParseError{"mix of a '" + t + "' sub-material (uber or nested mix) is not supported"};
                if (t == "mirror" || (t == "glass" && get_f(it->second->params, "uroughness", 0.f) == 0.f)) delta_subs++;
                m.mix_materials[j] = import_material(it->second);
            }
            if (delta_subs == 2)                                               // bsdf.cpp:615, delta_scattering.cpp:164
                throw ParseError{"mix of two delta sub-materials is not supported"};
        } else {
            throw ParseError{"unsupported material type '" + ms->type + "'"};
        }
        sd.materials.push_back(m);
        int id = (int)sd.materials.size() - 1;
        if (ms) ms->id = id;
        return id;
    }

    // pbrt-parser's createLightSource_point/spot/distant (impl/semantic/Lights.cpp:50-168): "from" and
    // "to" are transformed by the CTM when given; import_pbrt_non_area_light (pbrt_scene.cpp:903-956).
    F3 light_point(const ParamList& ps, const char* name, F3 dflt) {
        const Param* p = find(ps, name);
        if (!p) return dflt;
        if (p->fnums.size() != 3) throw ParseError{std::string("light '") + name + "' needs 3 values"};
        return transform_point(gs.ctm, F3(p->fnums[0], p->fnums[1], p->fnums[2]));
    }
    void light_source(const std::string& type, const ParamList& ps) {
        for (const char* spectral : {"I", "L"})
            if (find(ps, spectral) && find(ps, spectral)->type != "rgb" && find(ps, spectral)->type != "color")
                throw ParseError{std::string("light ") + spectral + " must be an rgb value"};
        float sc[3] = {1, 1, 1};
        get_rgb(ps, "scale", sc);
        if (type == "point") {
            yar_point_light_desc L{};
            float I[3] = {1, 1, 1};
            get_rgb(ps, "I", I);
            const F3 from = light_point(ps, "from", F3(0, 0, 0));
            L.position[0] = from.x; L.position[1] = from.y; L.position[2] = from.z;
            for (int i = 0; i < 3; i++) L.intensity[i] = I[i] * sc[i];
            sd.point_lights.push_back(L);
        } else if (type == "spot") {
            yar_spot_light_desc L{};
            float I[3] = {1, 1, 1};
            get_rgb(ps, "I", I);
            const F3 from = light_point(ps, "from", F3(0, 0, 0)), to = light_point(ps, "to", F3(0, 0, 0));
            const F3 dir = normalized(to - from);
            L.position[0] = from.x; L.position[1] = from.y; L.position[2] = from.z;
            L.direction[0] = dir.x; L.direction[1] = dir.y; L.direction[2] = dir.z;
            constexpr float deg_2_rad = kPi / 180.f;                        // math.h:22 radians()
            L.cone_angle = get_f(ps, "coneangle", 30.f) * deg_2_rad;
            L.penumbra_angle = get_f(ps, "conedeltaangle", 5.f) * deg_2_rad;
            for (int i = 0; i < 3; i++) L.intensity[i] = I[i] * sc[i];
            sd.spot_lights.push_back(L);
        } else if (type == "distant") {
            if (find(ps, "L") && find(ps, "L")->fnums.size() == 2)
                throw ParseError{"blackbody distant lights are not supported"};
            yar_directional_light_desc L{};
            float E[3] = {1, 1, 1};
            get_rgb(ps, "L", E);
            const F3 from = light_point(ps, "from", F3(0, 0, 0)), to = light_point(ps, "to", F3(0, 0, 1));
            const F3 dir = normalized(from - to);                          // instance transform: identity
            L.direction[0] = dir.x; L.direction[1] = dir.y; L.direction[2] = dir.z;
            for (int i = 0; i < 3; i++) L.irradiance[i] = E[i] * sc[i];
            sd.directional_lights.push_back(L);
        } else if (type == "infinite") {                                    // pbrt_scene.cpp:958-980
            for (const Param& e : ps)                                          // Lights.cpp:15-48
                if (e.name != "L" && e.name != "scale" && e.name != "nsamples" && e.name != "mapname")
                    throw ParseError{"unknown 'infinite' light source param '" + e.name + "'"};
            if (find(ps, "mapname")) throw ParseError{"environment maps (infinite light mapname) are not supported"};
            if (!sd.environment_lights.empty()) throw ParseError{"more than one infinite light"};
            yar_environment_light_desc L{};
            float Lrgb[3] = {1, 1, 1};
            get_rgb(ps, "L", Lrgb);
            const M34 w2l = inverse_transform(gs.ctm);
            std::memcpy(L.light_to_world, gs.ctm.a, sizeof(gs.ctm.a));
            std::memcpy(L.world_to_light, w2l.a, sizeof(w2l.a));
            for (int i = 0; i < 3; i++) L.radiance[i] = sc[i] * Lrgb[i];
            L.sample_count = get_i(ps, "nsamples", 1);
            sd.environment_lights.push_back(L);
        } else {
            throw ParseError{"unsupported light source '" + type + "'"};
        }
    }

    // import_pbrt_sphere and the sphere branch of import_pbrt_shape (pbrt_scene.cpp:745-758, 853-893):
    // the unit-radius-scaled tessellation in object space under the current transform, which
    // pbrt-parser keeps as the shape's transform (Geometry.cpp:323-332); an emitter becomes a sphere
    // light centred at the transform's translation with the unscaled radius.
    void sphere(const ParamList& ps) {
        const float radius = get_f(ps, "radius", 1.f);
        auto it = sphere_meshes.find(radius);
        if (it == sphere_meshes.end()) {
            sd.meshes.push_back(geodesic_sphere(radius, 6));
            it = sphere_meshes.emplace(radius, (int32_t)sd.meshes.size() - 1).first;
        }
        yar_object_desc obj{};
        obj.mesh = it->second;
        obj.material = -1;
        obj.area_light = -1;
        const M34 w2o = inverse_transform(gs.ctm);
        std::memcpy(obj.object_to_world, gs.ctm.a, sizeof(gs.ctm.a));
        std::memcpy(obj.world_to_object, w2o.a, sizeof(w2o.a));
        if (gs.area_light) {
            if (find(*gs.area_light, "L") && find(*gs.area_light, "L")->fnums.size() == 2)
                throw ParseError{"blackbody area lights are not supported"};
            yar_sphere_light_desc L{};
            L.object = (int32_t)sd.objects.size();
            const F3 c = gs.ctm.column(3);
            L.position[0] = c.x; L.position[1] = c.y; L.position[2] = c.z;
            L.radius = radius;
            float Lrgb[3] = {1, 1, 1};
            get_rgb(*gs.area_light, "L", Lrgb);
            for (int i = 0; i < 3; i++) L.emitted_radiance[i] = Lrgb[i];
            L.sample_count = get_i(*gs.area_light, "nsamples", 1);
            sphere_objects.emplace_back(L.object, (int32_t)sd.sphere_lights.size());
            sd.sphere_lights.push_back(L);
        } else {
            obj.material = import_material(gs.material);
        }
        sd.objects.push_back(obj);
    }

    void shape(const std::string& type, const ParamList& ps) {
        if (type == "sphere") { sphere(ps); return; }
        if (type != "trianglemesh") throw ParseError{"unsupported shape '" + type + "'"};
        const Param* P = find(ps, "P");
        const Param* I = find(ps, "indices");
        if (!P) throw ParseError{"trianglemesh without P"};
        MeshData mesh;
        mesh.vertices = P->fnums;
        size_t nv = mesh.vertices.size() / 3;
        if (I) {
            mesh.indices.resize(I->nums.size());
            for (size_t i = 0; i < I->nums.size(); i++) mesh.indices[i] = (int32_t)I->nums[i];
        } else if (nv == 3) {
            mesh.indices = {0, 1, 2};
        } else {
            throw ParseError{"trianglemesh without indices"};
        }
        if (mesh.indices.size() % 3) throw ParseError{"trianglemesh index count is not a multiple of 3"};
        for (int32_t ix : mesh.indices) if (ix < 0 || (size_t)ix >= nv) throw ParseError{"trianglemesh index out of range"};
        if (const Param* N = find(ps, "N")) {
            if (N->nums.size() != mesh.vertices.size()) throw ParseError{"N count differs from P count"};
            mesh.normals = N->fnums;
        }
        bake_transform(gs.ctm, mesh.vertices, mesh.normals);
        const Param* UV = find(ps, "uv");
        if (!UV) UV = find(ps, "st");
        if (UV) {
            if (UV->nums.size() != 2 * nv) throw ParseError{"uv count differs from vertex count"};
            mesh.uvs = UV->fnums;
        } else if (nv == 4) {
            mesh.uvs = {0, 0, 1, 0, 1, 1, 1, 0};       // pbrt_scene.cpp:704-712 ("not a typo")
        } else if (nv == 3) {
            mesh.uvs = {0, 0, 1, 0, 1, 1};             // pbrt_scene.cpp:714-720
        }
        mesh.reverse = gs.reverse_orientation;
        // remove_degenerate_triangles (triangle_mesh.cpp:5): drop triangles with coincident vertices
        {
            std::vector<int32_t> good;
            good.reserve(mesh.indices.size());
            for (size_t t = 0; t + 2 < mesh.indices.size(); t += 3) {
                const float* a = &mesh.vertices[3 * (size_t)mesh.indices[t]];
                const float* b = &mesh.vertices[3 * (size_t)mesh.indices[t + 1]];
                const float* c = &mesh.vertices[3 * (size_t)mesh.indices[t + 2]];
                auto eq = [](const float* u, const float* v) { return u[0] == v[0] && u[1] == v[1] && u[2] == v[2]; };
                if (eq(a, b) || eq(a, c) || eq(b, c)) continue;
                good.insert(good.end(), {mesh.indices[t], mesh.indices[t + 1], mesh.indices[t + 2]});
            }
            mesh.indices.swap(good);
        }
        if (mesh.indices.empty()) return;

        yar_object_desc obj{};
        obj.mesh = (int32_t)sd.meshes.size();
        obj.material = -1;
        obj.area_light = -1;
        // the shape's transform is baked into its vertices: object space is world space
        // (pbrt_scene.cpp:1084, instance transform and shape transform both identity)
        M34 o2w = M34::identity();
        M34 w2o = inverse_transform(o2w);
        std::memcpy(obj.object_to_world, o2w.a, sizeof(o2w.a));
        std::memcpy(obj.world_to_object, w2o.a, sizeof(w2o.a));
        if (gs.area_light && !mesh_is_rectangle_light(mesh)) {
            // import_pbrt_shape (pbrt_scene.cpp:786-794): any other emitter is a triangle-mesh light
            yar_mesh_light_desc L{};
            L.object = (int32_t)sd.objects.size();
            float Lrgb[3] = {1, 1, 1};
            get_rgb(*gs.area_light, "L", Lrgb);
            for (int i = 0; i < 3; i++) L.emitted_radiance[i] = Lrgb[i];
            sd.mesh_lights.push_back(L);
            obj.area_light = -2 - (int32_t)(sd.mesh_lights.size() - 1);   // resolved in finish()
        } else if (gs.area_light) {
            float size[2];
            M34 rect{};
            mesh_is_rectangle(mesh, size, rect);
            if (gs.reverse_orientation) {
                rect.set_column(0, -rect.column(0));
                rect.set_column(2, -rect.column(2));
            }
            yar_rect_light_desc L{};
            std::memcpy(L.light_to_world, rect.a, sizeof(rect.a));
            float Lrgb[3] = {1, 1, 1};
            get_rgb(*gs.area_light, "L", Lrgb);
            for (int i = 0; i < 3; i++) L.emitted_radiance[i] = Lrgb[i];
            L.size[0] = size[0]; L.size[1] = size[1];
            L.sample_count = get_i(*gs.area_light, "nsamples", 1);
            sd.rect_lights.push_back(L);
            obj.area_light = (int32_t)sd.rect_lights.size() - 1;
        } else {
            obj.material = import_material(gs.material);
        }
        sd.meshes.push_back(std::move(mesh));
        sd.objects.push_back(obj);
    }

    void run(Lexer& lx) {
        bool in_world = false;
        while (true) {
            char k = lx.peek_kind();
            if (k == 0) break;
            if (k != 'i') throw ParseError{"expected directive at line " + std::to_string(lx.line)};
            std::string d = lx.ident();
            if (d == "LookAt") {                                             // Parser.inl:695
                float v[9];
                for (float& x : v) lx.number2(&x);
                F3 eye(v[0], v[1], v[2]), at(v[3], v[4], v[5]), up(v[6], v[7], v[8]);
                F3 vz = normalized(at - eye);
                F3 vx = normalized(cross(up, vz));
                F3 vy = cross(vz, vx);
                M34 x;
                x.set_column(0, vx); x.set_column(1, vy); x.set_column(2, vz); x.set_column(3, eye);
                transform(inverse_transform(x));
                z_is_up = std::fabs(up.z) > std::fabs(up.y);
            } else if (d == "Translate") {
                M34 x = M34::identity();
                for (int i = 0; i < 3; i++) lx.number2(&x.a[i][3]);
                transform(x);
            } else if (d == "Rotate") {                                      // Parser.inl:335-339
                float angle, ax[3];
                lx.number2(&angle);
                for (float& x : ax) lx.number2(&x);
                transform(rotation(F3(ax[0], ax[1], ax[2]), angle * kPi / 180.f));
            } else if (d == "Scale") {
                M34 x = M34::identity();
                for (int i = 0; i < 3; i++) lx.number2(&x.a[i][i]);
                transform(x);
            } else if (d == "Transform" || d == "ConcatTransform") {
                lx.expect('[');
                float m[16];
                for (float& x : m) lx.number2(&x);
                lx.expect(']');
                M34 x;
                for (int c = 0; c < 4; c++) for (int r = 0; r < 3; r++) x.a[r][c] = m[4 * c + r];
                if (d == "Transform") gs.ctm = x; else transform(x);
            } else if (d == "Identity") {
                gs.ctm = M34::identity();
            } else if (d == "Camera") {
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                if (type != "perspective") throw ParseError{"unsupported camera '" + type + "'"};
                fov = get_f(ps, "fov", 30.f);                               // Camera.cpp findCameraFov
                camera_ctm = gs.ctm;
                have_camera = true;
            } else if (d == "Film") {
                lx.string();
                ParamList ps = parse_params(lx);
                have_film = true;
                xres = get_i(ps, "xresolution", xres);
                yres = get_i(ps, "yresolution", yres);
                if (const Param* c = find(ps, "cropwindow")) {
                    if (c->nums.size() != 4) throw ParseError{"cropwindow needs 4 values"};
                    for (int i = 0; i < 4; i++) crop[i] = c->fnums[i];
                    have_crop = true;
                }
                film_scale = get_f(ps, "scale", 1.f);
                max_component = get_f(ps, "maxsampleluminance", kInf);
            } else if (d == "Sampler") {                                     // Sampler.cpp
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                if (type == "stratified") {
                    xs = get_i(ps, "xsamples", 1);
                    ys = get_i(ps, "ysamples", 1);
                } else {
                    int n = get_i(ps, "pixelsamples", 1);
                    int kk = (int)std::ceil(std::sqrt((double)n));
                    xs = ys = kk;
                }
            } else if (d == "Integrator") {                                  // Integrator.cpp
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                if (type != "path" && type != "bdpt" && type != "volpath")
                    throw ParseError{"unsupported integrator '" + type + "'"};
                have_integrator = true;
                maxdepth = get_i(ps, "maxdepth", 5);
                rr_threshold = get_f(ps, "rrthreshold", 1.f);
            } else if (d == "PixelFilter") {
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                if (type != "box") throw ParseError{"unsupported pixel filter '" + type + "'"};
                filter_radius = get_f(ps, "xwidth", 0.5f);
            } else if (d == "WorldBegin") {
                in_world = true;
                gs.ctm = M34::identity();
            } else if (d == "WorldEnd") {
                in_world = false;
            } else if (d == "AttributeBegin" || d == "TransformBegin") {
                stack.push_back(gs);
            } else if (d == "AttributeEnd" || d == "TransformEnd") {
                if (stack.empty()) throw ParseError{"unbalanced " + d};
                if (d == "AttributeEnd") gs = stack.back();
                else gs.ctm = stack.back().ctm;
                stack.pop_back();
            } else if (d == "Material") {
                auto ms = std::make_shared<MaterialState>();
                ms->type = lx.string();
                ms->params = parse_params(lx);
                gs.material = ms;
            } else if (d == "MakeNamedMaterial") {
                std::string name = lx.string();
                auto ms = std::make_shared<MaterialState>();
                ms->params = parse_params(lx);
                const Param* t = find(ms->params, "type");
                if (!t || t->strs.empty()) throw ParseError{"MakeNamedMaterial without type"};
                ms->type = t->strs[0];
                named[name] = ms;
            } else if (d == "NamedMaterial") {
                std::string name = lx.string();
                auto it = named.find(name);
                if (it == named.end()) throw ParseError{"unknown named material '" + name + "'"};
                gs.material = it->second;
            } else if (d == "AreaLightSource") {
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                if (type != "diffuse") throw ParseError{"unsupported area light '" + type + "'"};
                gs.area_light = std::make_shared<ParamList>(ps);
            } else if (d == "LightSource") {
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                light_source(type, ps);
            } else if (d == "ReverseOrientation") {
                gs.reverse_orientation = !gs.reverse_orientation;
            } else if (d == "Shape") {
                std::string type = lx.string();
                ParamList ps = parse_params(lx);
                if (!in_world) throw ParseError{"Shape outside WorldBegin/WorldEnd"};
                shape(type, ps);
            } else {
                throw ParseError{"unsupported directive '" + d + "' at line " + std::to_string(lx.line)};
            }
        }
    }

    void finish() {
        if (!have_camera) throw ParseError{"scene has no Camera"};
        yar_render_config& c = sd.desc.config;
        c.film_width = xres;
        c.film_height = yres;
        int cx0 = 0, cy0 = 0, cx1 = xres, cy1 = yres;                    // Camera.cpp createFilm
        if (have_crop) {
            cx0 = (int)std::ceil(xres * crop[0]);
            cx1 = (int)std::ceil(xres * crop[1]);
            cy0 = (int)std::ceil(yres * crop[2]);
            cy1 = (int)std::ceil(yres * crop[3]);
        }
        c.region_x0 = xres - cx1;                                          // pbrt_scene.cpp:1104-1112
        c.region_x1 = xres - cx0;
        c.region_y0 = cy0;
        c.region_y1 = cy1;
        c.x_pixel_sample_count = xs;
        c.y_pixel_sample_count = ys;
        c.max_light_bounces = have_integrator ? maxdepth : 32;             // raytracer_config.h:21
        c.russian_roulette_bounce_count_threshold = 3;
        c.russian_roulette_threshold = have_integrator ? rr_threshold : 1.f;
        c.film_radiance_scale = film_scale;
        c.max_rgb_component_value_of_film_sample = max_component;
        c.pixel_filter_radius = filter_radius;
        c.rng_seed_offset = 0;
        c.pbrt3_scene = 1;
        // import_pbrt_camera (pbrt_scene.cpp:985-1045): frame = inverse(CTM at Camera)
        M34 frame = inverse_transform(camera_ctm);
        F3 right = frame.column(0), up = frame.column(1), forward = frame.column(2), pos = frame.column(3);
        M34 vp;
        if (z_is_up) { vp.set_column(0, -right); vp.set_column(1, forward); vp.set_column(2, up); }
        else { vp.set_column(0, -right); vp.set_column(1, up); vp.set_column(2, -forward); }
        vp.set_column(3, pos);
        std::memcpy(sd.desc.camera.camera_to_world, vp.a, sizeof(vp.a));
        sd.desc.camera.z_is_up = z_is_up ? 1 : 0;
        if (yres > xres) {
            constexpr float deg2rad = kPi / 180.f, rad2deg = 180.f / kPi;
            float tx = std::tan((fov / 2.f) * deg2rad);
            float ty = tx * float(yres) / float(xres);
            sd.desc.camera.fov_y_degrees = (2.f * std::atan(ty)) * rad2deg;
        } else {
            sd.desc.camera.fov_y_degrees = fov;
        }
        if (sd.point_lights.empty() && sd.rect_lights.empty() && sd.spot_lights.empty() &&
            sd.directional_lights.empty() && sd.mesh_lights.empty() && sd.sphere_lights.empty() &&
            sd.environment_lights.empty())
            throw ParseError{"scene has no supported lights (the default directional light is not supported)"};
        // emitter objects index the concatenation rect lights, sphere lights, mesh lights
        const int32_t n_rects = (int32_t)sd.rect_lights.size(), n_spheres = (int32_t)sd.sphere_lights.size();
        for (yar_object_desc& o : sd.objects)
            if (o.area_light <= -2) o.area_light = n_rects + n_spheres + (-2 - o.area_light);
        for (const auto& [object, light] : sphere_objects) sd.objects[object].area_light = n_rects + light;
        sd.finalize_desc();
    }
};

}  // namespace
}  // namespace yar

extern "C" int yar_scene_load_pbrt(const char* path, yar_scene** out_scene) {
    using namespace yar;
    if (!path || !out_scene) { set_last_error("yar_scene_load_pbrt: null argument"); return YAR_ERR_INVALID_ARGUMENT; }
    *out_scene = nullptr;
    return abi_guard(path, [&] {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw Error{YAR_ERR_IO, "cannot open file"};
        std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        std::unique_ptr<yar_scene> scene(new yar_scene);
        try {
            Lexer lx{text.data(), text.data() + text.size()};
            Loader ld(scene->data);
            ld.run(lx);
            ld.finish();
        } catch (const ParseError& e) {
            throw Error{YAR_ERR_PARSE, e.msg};
        }
        *out_scene = scene.release();
    });
}

extern "C" const yar_scene_desc* yar_scene_get_desc(const yar_scene* scene) {
    return scene ? &scene->data.desc : nullptr;
}

extern "C" void yar_scene_free(yar_scene* scene) { delete scene; }

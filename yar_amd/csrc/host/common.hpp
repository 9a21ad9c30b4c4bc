// common.hpp - shared host-side helpers of the path-tracing core: error reporting, float3 math in
// the reference renderer's operation order, and the owned scene representation behind
// yar_scene_desc.
#pragma once

#include "../../../include/yar_gpu.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <new>
#include <string>
#include <vector>

namespace yar {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr float kPi = 3.14159265f;   // src/lib/math.h:9
constexpr float kPi2 = 6.2831853f;   // src/lib/math.h:10
constexpr float kOneMinusEpsilon = 0x1.fffffep-1;   // src/lib/math.h:14

// Thread-local last-error text for calls that have no context (scene, kd-tree, generator).
void set_last_error(const std::string& msg);
const char* get_last_error();

struct Error {
    int code;
    std::string msg;
};

// Runs the body of a context-free extern "C" entry point: every C++ exception becomes a status code
// and the thread's last-error text ("name: message"), so none crosses the C ABI.
template <class Body>
int abi_guard(const std::string& name, Body&& body) {
    try {
        body();
        return YAR_OK;
    } catch (const Error& e) {
        set_last_error(name + ": " + e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error(name + ": out of host memory");
        return YAR_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(name + ": " + e.what());
        return YAR_ERR_INTERNAL;
    } catch (...) {
        set_last_error(name + ": unknown exception");
        return YAR_ERR_INTERNAL;
    }
}

struct F3 {
    float x = 0, y = 0, z = 0;
    F3() = default;
    F3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return (&x)[i]; }
    float& operator[](int i) { return (&x)[i]; }
};
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
inline F3 operator*(F3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }
inline F3 operator*(float t, F3 a) { return a * t; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(F3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline F3 normalized(F3 a) { float l = length(a); return {a.x / l, a.y / l, a.z / l}; }

// Row-major 3x4 affine matrix (Matrix3x4, src/lib/matrix.h:7).
struct M34 {
    float a[3][4];
    static M34 identity() {
        M34 m{};
        m.a[0][0] = m.a[1][1] = m.a[2][2] = 1.f;
        return m;
    }
    void set_column(int c, F3 v) { a[0][c] = v.x; a[1][c] = v.y; a[2][c] = v.z; }
    F3 column(int c) const { return {a[0][c], a[1][c], a[2][c]}; }
};
M34 mul(const M34& m1, const M34& m2);                       // matrix.cpp operator*
M34 inverse_transform(const M34& m);                         // get_inverse_transform, matrix.cpp:288
inline F3 transform_point(const M34& m, F3 p) {
    return {m.a[0][0] * p.x + m.a[0][1] * p.y + m.a[0][2] * p.z + m.a[0][3],
            m.a[1][0] * p.x + m.a[1][1] * p.y + m.a[1][2] * p.z + m.a[1][3],
            m.a[2][0] * p.x + m.a[2][1] * p.y + m.a[2][2] * p.z + m.a[2][3]};
}
inline F3 transform_vector(const M34& m, F3 v) {
    return {m.a[0][0] * v.x + m.a[0][1] * v.y + m.a[0][2] * v.z,
            m.a[1][0] * v.x + m.a[1][1] * v.y + m.a[1][2] * v.z,
            m.a[2][0] * v.x + m.a[2][1] * v.y + m.a[2][2] * v.z};
}

// Owned scene data; yar_scene_desc points into it.
struct MeshData {
    std::vector<float> vertices, normals, uvs;
    std::vector<int32_t> indices;
    bool reverse = false;
};

struct SceneData {
    std::vector<MeshData> meshes;
    std::vector<yar_mesh_desc> mesh_descs;
    std::vector<yar_material_desc> materials;
    std::vector<yar_object_desc> objects;
    std::vector<yar_rect_light_desc> rect_lights;
    std::vector<yar_point_light_desc> point_lights;
    std::vector<yar_spot_light_desc> spot_lights;
    std::vector<yar_directional_light_desc> directional_lights;
    std::vector<yar_mesh_light_desc> mesh_lights;
    std::vector<yar_sphere_light_desc> sphere_lights;
    std::vector<yar_environment_light_desc> environment_lights;   // at most one
    yar_scene_desc desc{};
    void finalize_desc();
};

}  // namespace yar

struct yar_scene {
    yar::SceneData data;
};

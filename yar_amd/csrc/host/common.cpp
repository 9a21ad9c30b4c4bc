// common.cpp - error text and matrix helpers shared by the host components.
#include "common.hpp"

#include <algorithm>
#include <utility>

namespace yar {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* get_last_error() { return g_last_error.c_str(); }

// Matrix3x4 product (src/lib/matrix.cpp, operator*): affine composition m1 * m2.
M34 mul(const M34& m1, const M34& m2) {
    M34 r{};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 4; j++) {
            r.a[i][j] = m1.a[i][0] * m2.a[0][j] + m1.a[i][1] * m2.a[1][j] + m1.a[i][2] * m2.a[2][j];
        }
        r.a[i][3] += m1.a[i][3];
    }
    return r;
}

// Gauss-Jordan elimination with partial pivoting on the augmented [M | t | I] system; follows
// get_inverse_transform (src/lib/matrix.cpp:288-343) so object transforms invert identically.
M34 inverse_transform(const M34& m) {
    float a[3][7];
    for (int i = 0; i < 3; i++) {
        for (int k = 0; k < 4; k++) a[i][k] = m.a[i][k];
        for (int k = 0; k < 3; k++) a[i][4 + k] = (i == k) ? 1.f : 0.f;
    }
    for (int i = 0; i < 3; i++) {
        float big = std::fabs(a[i][i]);
        int piv = i;
        for (int i2 = i + 1; i2 < 3; i2++) {
            if (std::fabs(a[i2][i]) > big) { big = std::fabs(a[i2][i]); piv = i2; }
        }
        if (piv != i)
            for (int k = i; k < 7; k++) std::swap(a[i][k], a[piv][k]);
        float inv_pivot = 1.f / a[i][i];
        for (int k = i + 1; k < 7; k++) a[i][k] *= inv_pivot;
        for (int i2 = 0; i2 < 3; i2++) {
            if (i2 == i) continue;
            float coeff = -a[i2][i];
            for (int k = i + 1; k < 7; k++) a[i2][k] += a[i][k] * coeff;
        }
    }
    M34 inv{};
    for (int i = 0; i < 3; i++) {
        for (int k = 0; k < 3; k++) inv.a[i][k] = a[i][k + 4];
        inv.a[i][3] = -a[i][3];
    }
    return inv;
}

void SceneData::finalize_desc() {
    mesh_descs.resize(meshes.size());
    for (size_t i = 0; i < meshes.size(); i++) {
        MeshData& m = meshes[i];
        yar_mesh_desc& d = mesh_descs[i];
        d.vertices = m.vertices.data();
        d.normals = m.normals.empty() ? nullptr : m.normals.data();
        d.uvs = m.uvs.empty() ? nullptr : m.uvs.data();
        d.indices = m.indices.data();
        d.vertex_count = (int32_t)(m.vertices.size() / 3);
        d.triangle_count = (int32_t)(m.indices.size() / 3);
        d.reverse_geometric_normal_orientation = m.reverse ? 1 : 0;
        d.reserved = 0;
    }
    desc.meshes = mesh_descs.data();
    desc.mesh_count = (int32_t)mesh_descs.size();
    desc.materials = materials.data();
    desc.material_count = (int32_t)materials.size();
    desc.objects = objects.data();
    desc.object_count = (int32_t)objects.size();
    desc.rect_lights = rect_lights.data();
    desc.rect_light_count = (int32_t)rect_lights.size();
    desc.point_lights = point_lights.data();
    desc.point_light_count = (int32_t)point_lights.size();
    desc.spot_lights = spot_lights.data();
    desc.spot_light_count = (int32_t)spot_lights.size();
    desc.directional_lights = directional_lights.data();
    desc.directional_light_count = (int32_t)directional_lights.size();
    desc.mesh_lights = mesh_lights.data();
    desc.mesh_light_count = (int32_t)mesh_lights.size();
    desc.sphere_lights = sphere_lights.data();
    desc.sphere_light_count = (int32_t)sphere_lights.size();
    desc.environment_light = environment_lights.empty() ? nullptr : environment_lights.data();
    desc.environment_light_count = (int32_t)environment_lights.size();
}

}  // namespace yar

extern "C" const char* yar_last_error(void) { return yar::get_last_error(); }

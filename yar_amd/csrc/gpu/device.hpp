// device.hpp - device-side data layout, float math and two-level kd-tree traversal for gfx950.
//
// All arithmetic keeps the reference renderer's operation order and the file is compiled with
// -ffp-contract=off (no FMA contraction), so hit records are bit-identical to a contraction-free
// build of the reference traversal (kennyalive/YAR src/ref/kdtree.cpp, intersection.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace yg {

#define YG_INLINE __device__ __forceinline__

constexpr float kPi = 3.14159265f;          // src/lib/math.h:9
constexpr float kPiInv = 1.f / kPi;         // src/lib/math.h:12
constexpr float kOneMinusEps = 0x1.fffffep-1f;

YG_INLINE float f_inf() { return __builtin_huge_valf(); }
YG_INLINE uint32_t fbits(float f) { return __float_as_uint(f); }
YG_INLINE float bitsf(uint32_t u) { return __uint_as_float(u); }

// ---- float3 with explicit operation order --------------------------------------------------
struct V3 {
    float x, y, z;
    YG_INLINE float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
YG_INLINE V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
YG_INLINE V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
YG_INLINE V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
YG_INLINE V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
YG_INLINE V3 operator*(V3 a, float t) { return v3(a.x * t, a.y * t, a.z * t); }
YG_INLINE V3 operator*(float t, V3 a) { return a * t; }
YG_INLINE V3 vdiv(V3 a, float t) { return v3(a.x / t, a.y / t, a.z / t); }
YG_INLINE float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
YG_INLINE V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
YG_INLINE float len_sq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
YG_INLINE float length(V3 a) { return sqrtf(len_sq(a)); }
YG_INLINE V3 normalized(V3 a) { return vdiv(a, length(a)); }
YG_INLINE V3 from4(float4 f) { return v3(f.x, f.y, f.z); }

// ColorRGB semantics: division by a scalar multiplies by its reciprocal (src/lib/color.h:104).
struct RGB { float r, g, b; };
YG_INLINE RGB rgb(float r, float g, float b) { RGB c; c.r = r; c.g = g; c.b = b; return c; }
YG_INLINE RGB operator*(RGB a, RGB b) { return rgb(a.r * b.r, a.g * b.g, a.b * b.b); }
YG_INLINE RGB operator*(RGB a, float k) { return rgb(a.r * k, a.g * k, a.b * k); }
YG_INLINE RGB operator*(float k, RGB a) { return a * k; }
YG_INLINE RGB operator+(RGB a, RGB b) { return rgb(a.r + b.r, a.g + b.g, a.b + b.b); }
YG_INLINE RGB operator-(RGB a, RGB b) { return rgb(a.r - b.r, a.g - b.g, a.b - b.b); }
YG_INLINE RGB cdiv(RGB a, RGB b) { return rgb(a.r / b.r, a.g / b.g, a.b / b.b); }
YG_INLINE RGB cdivs(RGB a, float k) { float inv = 1.f / k; return a * inv; }
YG_INLINE RGB csqrt(RGB a) { return rgb(sqrtf(a.r), sqrtf(a.g), sqrtf(a.b)); }
YG_INLINE RGB splat(float v) { return rgb(v, v, v); }
YG_INLINE bool is_black(RGB a) { return a.r == 0 && a.g == 0 && a.b == 0; }
YG_INLINE RGB rgb4(float4 f) { return rgb(f.x, f.y, f.z); }
YG_INLINE float4 f4(RGB c, float w) { return make_float4(c.r, c.g, c.b, w); }
YG_INLINE float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// 3x4 row-major matrix stored as three float4 rows.
struct M34 { float4 r0, r1, r2; };
YG_INLINE V3 xform_point(const M34& m, V3 p) {
    return v3(m.r0.x * p.x + m.r0.y * p.y + m.r0.z * p.z + m.r0.w,
              m.r1.x * p.x + m.r1.y * p.y + m.r1.z * p.z + m.r1.w,
              m.r2.x * p.x + m.r2.y * p.y + m.r2.z * p.z + m.r2.w);
}
YG_INLINE V3 xform_vector(const M34& m, V3 v) {
    return v3(m.r0.x * v.x + m.r0.y * v.y + m.r0.z * v.z,
              m.r1.x * v.x + m.r1.y * v.y + m.r1.z * v.z,
              m.r2.x * v.x + m.r2.y * v.y + m.r2.z * v.z);
}

// ---- scene layout in HBM ---------------------------------------------------------------
// Object record for traversal (80 bytes): world_to_object rows, mesh-tree bounds with the
// mesh's root node index and triangle base packed into the w lanes.
struct ObjTrav { float4 w2o0, w2o1, w2o2, lo_node, hi_tri; };

// Device-only material codes: rough glass (yar_material_desc type GLASS with roughness > 0), the
// pbrt3 uber material and the mix material.
constexpr int kRoughGlassType = 7;
constexpr int kUberType = 8;
constexpr int kMixType = 9;
struct DevMaterial {      // 64 bytes
    int type; float eta_i, alpha_x, alpha_y;
    float4 refl;          // diffuse reflectance / glass Kr
    float4 trans;         // glass Kt (w = index of refraction)
    float4 eta;           // metal eta
    float4 k;             // metal k / coated diffuse Fresnel r0
};
// light_to_world rows, radiance, size (xy); size.z != 0 when the analytic MIS-ray test may not be used
// for this light (its emitter geometry is not known to lie inside the rectangle)
struct DevRect { float4 l0, l1, l2; float4 L; float4 size; };
struct DevPoint { float4 p; float4 I; };
// Spot_Light with cos(cone_angle) in dir.w and cos(max(0, cone_angle - penumbra_angle)) in I.w, both
// computed on the host as the reference computes them (direct_lighting.cpp:58-63)
struct DevSpot { float4 p; float4 dir; float4 I; };
struct DevDirLight { float4 dir; float4 E; };          // direction towards the light, irradiance
// Diffuse_Triangle_Mesh_Light: emitter object, its triangles' global range, area CDF offset, area
struct DevMeshLight { int obj, tri_base, tri_count, cdf; float area, pad0, pad1, pad2; float4 L; };
// Diffuse_Sphere_Light: world-space centre with the radius in p.w, radiance
struct DevSphere { float4 p; float4 L; };
// Environment_Light with a constant map: light_to_world and world_to_light rows, radiance
struct DevEnv { float4 l2w0, l2w1, l2w2, w2l0, w2l1, w2l2, L; };

struct DevScene {
    const uint2* scene_nodes;
    const uint2* mesh_nodes;       // all mesh trees; interior above_child and leaf ids rebased
    const ObjTrav* objs;
    const float4* tri_verts;       // 3 float4 per triangle (object space)
    float4 scene_lo, scene_hi;
    // shading
    const float4* obj_o2w;         // 3 rows per object
    const float4* obj_nrm;         // 3 rows per object (transpose of world_to_object 3x3)
    const int4* obj_info;          // material, area light, mesh, triangle base
    const int4* mesh_info;         // flags (1 normals, 2 uvs, 4 reverse), -, -, -
    const int4* tri_vidx;          // global vertex indices per triangle, w = its mesh's flags
    const float4* vtx_n;           // per vertex normal
    const float2* vtx_uv;          // per vertex uv
    const DevMaterial* materials;
    const DevRect* rects;
    const DevPoint* points;
    const DevSpot* spots;
    const DevDirLight* dirs;
    const DevMeshLight* mesh_lights;
    const float* mesh_light_cdf;
    const DevSphere* spheres;
    const DevEnv* env;
    // area lights are numbered rectangles, spheres, meshes (an object's area-light index); the
    // environment light, when there is one, comes last in the light order
    int n_rects, n_points, n_spots, n_dirs, n_mesh_lights, n_spheres, n_env, n_lights;
    int has_mix;                   // some material is a mix
};

// ---- ray/box slab test (Bounding_Box::intersect_by_ray_without_NaNs, bounding_box.h:63) ----
YG_INLINE bool slab(float4 lo, float4 hi, V3 o, V3 d, float& tmin, float& tmax) {
    float t0 = 0.f, t1 = f_inf();
    float los[3] = {lo.x, lo.y, lo.z}, his[3] = {hi.x, hi.y, hi.z};
    float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (ds[i] != 0.f) {
            float inv = 1.f / ds[i];
            float s0 = (los[i] - os[i]) * inv;
            float s1 = (his[i] - os[i]) * inv;
            if (inv < 0.f) { float t = s0; s0 = s1; s1 = t; }
            t0 = s0 > t0 ? s0 : t0;
            t1 = s1 < t1 ? s1 : t1;
            if (t0 > t1) return false;
        } else if (os[i] < los[i] || os[i] > his[i]) {
            return false;
        }
    }
    tmin = t0; tmax = t1;
    return true;
}

// ---- watertight ray/triangle test (intersection.cpp:53-113) ------------------------------
YG_INLINE float watertight(V3 o, V3 d, int kx, int ky, int kz, float sx, float sy, float sz,
                           float4 p0, float4 p1, float4 p2, V3& b) {
    V3 a0 = v3(p0.x - o.x, p0.y - o.y, p0.z - o.z);
    V3 a1 = v3(p1.x - o.x, p1.y - o.y, p1.z - o.z);
    V3 a2 = v3(p2.x - o.x, p2.y - o.y, p2.z - o.z);
    float p0x = a0[kx], p0y = a0[ky], p0z = a0[kz];
    float p1x = a1[kx], p1y = a1[ky], p1z = a1[kz];
    float p2x = a2[kx], p2y = a2[ky], p2z = a2[kz];
    float x0 = p0x + sx * p0z, y0 = p0y + sy * p0z;
    float x1 = p1x + sx * p1z, y1 = p1y + sy * p1z;
    float x2 = p2x + sx * p2z, y2 = p2y + sy * p2z;
    float e0 = x1 * y2 - y1 * x2;
    float e1 = x2 * y0 - y2 * x0;
    float e2 = x0 * y1 - y0 * x1;
    if (e0 == 0.f || e1 == 0.f || e2 == 0.f) {
        e0 = (float)((double)x1 * (double)y2 - (double)y1 * (double)x2);
        e1 = (float)((double)x2 * (double)y0 - (double)y2 * (double)x0);
        e2 = (float)((double)x0 * (double)y1 - (double)y0 * (double)x1);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return f_inf();
    float det = e0 + e1 + e2;
    if (det == 0.f) return f_inf();
    float z0 = sz * p0z, z1 = sz * p1z, z2 = sz * p2z;
    float ts = e0 * z0 + e1 * z1 + e2 * z2;
    if ((fbits(det) ^ fbits(ts)) >> 31) return f_inf();
    float inv_det = 1.f / det;
    float t = inv_det * ts;
    b = v3(e0 * inv_det, e1 * inv_det, e2 * inv_det);
    return t;
}

YG_INLINE bool degenerate(float4 p0, float4 p1, float4 p2) {
    V3 e1 = v3(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z);
    V3 e2 = v3(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z);
    return len_sq(cross(e1, e2)) == 0.f;
}

struct TravCounters { uint32_t nodes, leaf_refs, tri_tests, objects; };

struct HitRec { float t; int obj; uint32_t tri; V3 b; };

// Short traversal stack in LDS: entries [slot][lane] so a wave's accesses hit distinct banks.
// When a push finds the stack full the oldest entry is dropped; when the stack then runs empty
// the traversal restarts from the root at the current exit distance (kd-restart), visiting the
// remaining leaves in the same front-to-back order.  With depth >= tree depth it is the plain
// stack traversal of kdtree.cpp:219-439.
struct LdsStack {
    uint32_t* node;
    float* tmax;
    int stride;
    YG_INLINE void put(int slot, uint32_t n, float t) { node[slot * stride] = n; tmax[slot * stride] = t; }
    YG_INLINE void get(int slot, uint32_t& n, float& t) const { n = node[slot * stride]; t = tmax[slot * stride]; }
};

// The same stack in per-lane private memory (scratch): no LDS footprint, so occupancy is bounded by
// registers only; accesses are lane-interleaved by the hardware and stay in the vector L1.
template <int N>
struct PrivStack {
    uint32_t node[N];
    float tmax[N];
    YG_INLINE void put(int slot, uint32_t n, float t) { node[slot] = n; tmax[slot] = t; }
    YG_INLINE void get(int slot, uint32_t& n, float& t) const { n = node[slot]; t = tmax[slot]; }
};

// Leaf id j of a leaf whose first id is in word1 of node `nd` (kdtree.h:53: ids packed inline).
YG_INLINE uint32_t leaf_id(const uint2* __restrict__ nodes, uint32_t nd, uint32_t first, uint32_t j) {
    if (j == 0) return first;
    uint2 w = nodes[nd + ((j + 1) >> 1)];
    return (j & 1) ? w.x : w.y;
}

// Short-stack bookkeeping of one tree level packed in one word: push pointer in bits 0-7 (used
// modulo the stack size, a power of two dividing 256), live entries in bits 8-15, bit 16 set
// when entries were dropped since the last restart, bit 17 set while a descent is suspended.
constexpr uint32_t kSsCount = 1u << 8, kSsDropped = 1u << 16, kSsResume = 1u << 17;

// Interior nodes a lane descends per traversal step (a constant, so the descent is unrolled).
constexpr int kDescentBudget = 6;
YG_INLINE uint32_t ss_count(uint32_t ss) { return (ss >> 8) & 0xffu; }
YG_INLINE uint32_t ss_push(uint32_t ss) { return (ss & ~0xffu) | ((ss + 1u) & 0xffu); }
YG_INLINE uint32_t ss_pop(uint32_t ss) { return ((ss & ~0xffu) | ((ss - 1u) & 0xffu)) - kSsCount; }

// One step of a kd-tree descent shared by both levels: interior node -> child (pushing the far
// child when the ray straddles the split), as KdTree::intersect's inner loop (kdtree.cpp:245-290).
template <class Stack>
YG_INLINE void kd_interior(uint2 w, uint32_t& node, V3 o, V3 d, V3 inv, float& tmin, float& tmax, Stack& st,
                           int mask, int depth, uint32_t& ss) {
    // Written as selects so a wave runs one straight-line sequence whatever mix of cases its lanes
    // are in; only the stack write is a branch.
    const uint32_t axis = w.x & 3u;
    const float oa = axis == 0u ? o.x : (axis == 1u ? o.y : o.z);
    const float da = axis == 0u ? d.x : (axis == 1u ? d.y : d.z);
    const float ia = axis == 0u ? inv.x : (axis == 1u ? inv.y : inv.z);
    const float dist = bitsf(w.y) - oa;
    const uint32_t below = node + 1, above = w.x >> 2;
    const bool pos = dist > 0.f;
    const uint32_t first = pos ? below : above;
    const uint32_t second = pos ? above : below;
    const float ts = dist * ia;
    const bool zero = dist == 0.f;
    const bool only_first = ts > tmax || ts < 0.f;
    const bool only_second = !only_first && ts < tmin;
    const bool push = !zero && !only_first && !only_second;
    const uint32_t next = only_second ? second : first;
    const uint32_t on_plane = da > 0.f ? above : below;
    if (push) st.put((int)(ss & (uint32_t)mask), second, tmax);
    const uint32_t pushed = ss_push(ss);
    const uint32_t ss_new = ss_count(ss) == (uint32_t)depth ? (pushed | kSsDropped) : (pushed + kSsCount);
    ss = push ? ss_new : ss;
    tmax = push ? ts : tmax;
    node = zero ? on_plane : next;
}

// After a leaf: pop the next subtree, or restart behind the last leaf when the short stack dropped
// entries (kd-restart; a restart that would not advance steps past the degenerate point), or finish.
// Returns false when the level's traversal is complete.
template <class Stack>
YG_INLINE bool kd_next(uint32_t& node, uint32_t root, float& tmin, float& tmax, float t1, Stack& st, int mask, uint32_t& ss,
                       float& restart_t) {
    if (ss_count(ss) > 0) {
        ss = ss_pop(ss);
        uint32_t nn; float tm;
        st.get((int)(ss & (uint32_t)mask), nn, tm);
        node = nn;
        tmin = tmax;
        tmax = tm;
        return true;
    }
    if ((ss & kSsDropped) && tmax < t1) {
        ss &= ~kSsDropped;
        node = root;
        tmin = tmax > restart_t ? tmax : nextafterf(restart_t, f_inf());
        restart_t = tmin;
        tmax = t1;
        return true;
    }
    return false;
}

// Two-level closest-hit / any-hit query over the scene tree and the per-mesh trees
// (KdTree::intersect / intersect_any with intersect_scene_geometry_data and
// intersect_triangle_mesh_geometry_data, kdtree.cpp:12-109), as a resumable per-lane state
// machine.  Both levels run in one loop: each step() advances a lane by one node of whichever tree
// it is in (scene tree, or the mesh tree of the object being visited), so lanes inside a mesh tree
// and lanes still descending the scene tree make progress together, and a persistent kernel can
// hand a lane a new ray as soon as its query ends.  The visit order and hit updates are those of
// the nested traversal, so results are identical.
template <bool ANY, bool COUNT>
struct Tracer {
    V3 o, d, inv;
    float any_t, tmax_in, st1;
    HitRec hit;
    bool found_any;
    // scene level
    uint32_t snode;
    float stmin, stmax, srestart;
    uint32_t sss;                   // packed short-stack state
    // objects of the current scene leaf still to visit
    uint32_t lnode, lfirst, ln, lj;
    // mesh level
    bool in_mesh;
    uint32_t mroot, mnode, oi;
    float mtmin, mtmax, mt1, mrestart;
    uint32_t mss;                   // packed short-stack state
    V3 lo, ld, linv;
    uint32_t kxyz;                  // watertight axis permutation kx | ky << 2 | kz << 4
    float sx, sy, sz;
    int iter;

    // Returns false when the ray misses the scene bounds (the query is then already complete).
    YG_INLINE bool start(const DevScene& s, V3 o_, V3 d_, float tmax_) {
        o = o_; d = d_; tmax_in = tmax_; any_t = tmax_;
        hit.t = tmax_; hit.obj = -1; hit.tri = 0; hit.b = v3(0, 0, 0);
        found_any = false;
        float t0, t1;
        if (!slab(s.scene_lo, s.scene_hi, o, d, t0, t1)) return false;
        inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);
        st1 = t1;
        snode = 0; stmin = t0; stmax = t1; srestart = -1.f; sss = 0u;
        lnode = lfirst = ln = lj = 0;
        in_mesh = false;
        iter = 0;
        return true;
    }

    // closest hit: hit.t < tmax_in; any hit: found_any
    YG_INLINE bool result() const { return ANY ? found_any : hit.t < tmax_in; }

    // One traversal step; returns true when the query is complete.  A lane outside a mesh tree
    // takes the scene-level actions in traversal order, each at most once per step: finish the
    // current scene leaf, descend to the next one, set up its next object; a lane inside a mesh
    // tree (possibly just entered) then processes one mesh leaf.  A descent stops after
    // kDescentBudget interior nodes and resumes from the same node in the next step, so a lane far
    // from its next leaf does not hold the wave while the other lanes wait at theirs.
    template <class SStack, class MStack>
    YG_INLINE bool step(const DevScene& s, SStack& sst, int sdepth, MStack& mst, int mdepth, TravCounters& cnt) {
        // The iteration cap only guarantees that every wave terminates; real trees need far fewer.
        if (++iter > (1 << 24)) return true;
        if (!in_mesh) {
            if (ln > 0 && lj >= ln) {
                // the leaf's objects are done: continue the scene traversal behind it
                ln = 0;
                if (!kd_next(snode, 0u, stmin, stmax, st1, sst, sdepth - 1, sss, srestart)) return true;
            }
            if (ln == 0) {
                if (!((ANY ? any_t : hit.t) > stmin)) return true;
                if (COUNT && !(sss & kSsResume)) cnt.nodes++;
                sss &= ~kSsResume;
                uint2 w = s.scene_nodes[snode];
                for (int b = kDescentBudget; (w.x & 3u) != 3u;) {
                    if (b-- == 0) {
                        sss |= kSsResume;
                        return false;
                    }
                    kd_interior(w, snode, o, d, inv, stmin, stmax, sst, sdepth - 1, sdepth, sss);
                    if (COUNT) cnt.nodes++;
                    w = s.scene_nodes[snode];
                }
                const uint32_t n = w.x >> 2;
                if (COUNT) cnt.leaf_refs += n;
                if (n == 0) return !kd_next(snode, 0u, stmin, stmax, st1, sst, sdepth - 1, sss, srestart);
                lnode = snode; lfirst = w.y; ln = n; lj = 0;
            }
            // next object of the scene leaf: ray to object space, mesh bounds, watertight setup
            oi = leaf_id(s.scene_nodes, lnode, lfirst, lj);
            lj++;
            const ObjTrav ob = s.objs[oi];
            if (COUNT) cnt.objects++;
            const M34 w2o = {ob.w2o0, ob.w2o1, ob.w2o2};
            lo = xform_point(w2o, o);
            ld = xform_vector(w2o, d);
            float mt0;
            if (!slab(ob.lo_node, ob.hi_tri, lo, ld, mt0, mt1)) return false;
            linv = v3(1.f / ld.x, 1.f / ld.y, 1.f / ld.z);
            mroot = fbits(ob.lo_node.w);
            // watertight-test setup hoisted per mesh ray (intersection.cpp:55-61)
            const float ax = fabsf(ld.x), ay = fabsf(ld.y), az = fabsf(ld.z);
            const int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
            const int kx = (kz == 2 ? 0 : kz + 1);
            const int ky = (kz == 0 ? 2 : kz - 1);
            kxyz = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
            const float dx = ld[kx], dy = ld[ky], dz = ld[kz];
            sx = -dx / dz; sy = -dy / dz; sz = linv[kz];           // linv[kz] is 1.f / dz
            mnode = mroot; mtmin = mt0; mtmax = mt1; mrestart = -1.f;
            mss = 0u;
            in_mesh = true;
        }
        // one leaf of the current mesh tree
        bool done = !((ANY ? any_t : hit.t) > mtmin);
        if (!done) {
            // descend to the next leaf: neither the loop bound nor tmin changes on the way down
            if (COUNT && !(mss & kSsResume)) cnt.nodes++;
            mss &= ~kSsResume;
            uint2 w = s.mesh_nodes[mnode];
            for (int b = kDescentBudget; (w.x & 3u) != 3u;) {
                if (b-- == 0) {
                    mss |= kSsResume;
                    return false;
                }
                kd_interior(w, mnode, lo, ld, linv, mtmin, mtmax, mst, mdepth - 1, mdepth, mss);
                if (COUNT) cnt.nodes++;
                w = s.mesh_nodes[mnode];
            }
            const uint32_t n = w.x >> 2;
            if (COUNT) cnt.leaf_refs += n;
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t ti = leaf_id(s.mesh_nodes, mnode, w.y, k);
                const float4 p0 = s.tri_verts[3 * (size_t)ti], p1 = s.tri_verts[3 * (size_t)ti + 1],
                             p2 = s.tri_verts[3 * (size_t)ti + 2];
                if (COUNT) cnt.tri_tests++;
                V3 b;
                const float t = watertight(lo, ld, (int)(kxyz & 3u), (int)((kxyz >> 2) & 3u), (int)(kxyz >> 4), sx, sy,
                                           sz, p0, p1, p2, b);
                if (ANY) {
                    if (t >= any_t) continue;
                    if (degenerate(p0, p1, p2)) continue;
                    found_any = true;
                    return true;
                } else if (t < hit.t) {
                    if (degenerate(p0, p1, p2)) continue;
                    hit.t = t; hit.tri = ti; hit.b = b; hit.obj = (int)oi;
                }
            }
            done = !kd_next(mnode, mroot, mtmin, mtmax, mt1, mst, mdepth - 1, mss, mrestart);
        }
        if (done) {
            in_mesh = false;
        }
        return false;
    }
};

// Runs one query to completion.
template <bool ANY, bool COUNT, class SStack, class MStack>
YG_INLINE bool trace(const DevScene& s, V3 o, V3 d, float tmax_in, HitRec& hit, SStack& sst, int sdepth,
                     MStack& mst, int mdepth, TravCounters& cnt) {
    Tracer<ANY, COUNT> tr;
    if (tr.start(s, o, d, tmax_in))
        while (!tr.step(s, sst, sdepth, mst, mdepth, cnt)) {}
    hit = tr.hit;
    return tr.result();
}

// ---- PCG32 (third-party/pcg/pcg_basic.c:46-75) ------------------------------------------
constexpr uint64_t kPcgMult = 6364136223846793005ULL;
YG_INLINE uint32_t pcg_output(uint64_t old) {
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
YG_INLINE uint32_t pcg_next(uint64_t& state, uint64_t inc) {
    uint64_t old = state;
    state = old * kPcgMult + inc;
    return pcg_output(old);
}
YG_INLINE float pcg_float(uint64_t& state, uint64_t inc) {
    return bitsf((pcg_next(state, inc) >> 9) | 0x3f800000u) - 1.0f;
}
// State after `delta` further steps of the LCG (Brown 1994 jump-ahead, as pcg32_advance_r).
YG_INLINE uint64_t pcg_advance(uint64_t state, uint64_t inc, uint64_t delta) {
    uint64_t acc_mult = 1, acc_plus = 0, cur_mult = kPcgMult, cur_plus = inc;
    while (delta > 0) {
        if (delta & 1) { acc_mult *= cur_mult; acc_plus = acc_plus * cur_mult + cur_plus; }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}
// Table-driven jump-ahead: the state n steps after s is A_n * s + G_n * inc (mod 2^64) with
// A_n = M^n and G_n = M^(n-1) + ... + M + 1.  n = hi * kJumpLo + lo is looked up in two tables
// (lo[r] = (A_r, G_r), hi[q] = (A_qL, G_qL)), so any position of a stream costs two affine steps.
constexpr int kJumpLoBits = 12;
constexpr uint32_t kJumpLo = 1u << kJumpLoBits;
constexpr uint32_t kJumpHi = 4096;            // offsets below 2^24
struct JumpTables { const ulonglong2* lo; const ulonglong2* hi; };
YG_INLINE uint64_t pcg_jump(const JumpTables& J, uint64_t s, uint64_t inc, uint32_t n) {
    const ulonglong2 b = J.hi[n >> kJumpLoBits];
    const ulonglong2 a = J.lo[n & (kJumpLo - 1u)];
    s = b.x * s + b.y * inc;
    return a.x * s + a.y * inc;
}

// pcg32_srandom_r(rng, 0, stream): seeded state of a pixel stream.
YG_INLINE uint64_t pcg_seed_state(uint64_t stream, uint64_t& inc) {
    inc = (stream << 1u) | 1u;
    uint64_t st = 0;
    pcg_next(st, inc);
    st += 0;
    pcg_next(st, inc);
    return st;
}

}  // namespace yg

// kdtree_builder.cpp - host SAH kd-tree builder.
//
// Produces exactly the trees of the reference builder (kennyalive/YAR src/ref/kdtree_builder.cpp):
// SAH cost with empty-node bonus 0.3 (:8), leaf at <= 2 primitives (:9), depth limit
// round(8 + 1.3 floor(log2 N)) capped at 40 (kdtree.cpp:199-203), split clipping of straddling
// triangles (Soupikov et al. 2008, :17-93), edges ordered by position with end events before
// start events and ties kept in primitive order (stable sort, :118-123), planar primitives sent
// below (:249-262), sorted leaf index lists (:321), and the 8-byte node encoding (kdtree.h:18-77).
//
// What differs is the implementation: edge events are packed into 64-bit sort keys
// (orderable position | start flag | local index), so an unstable sort reproduces the stable
// order, scratch buffers are reused per depth level instead of reallocated, and independent
// meshes are built on worker threads.
//
// A binned fast path (not in the reference) replaces the sorted edge sweep of nodes with more than
// kBinnedMin primitives by the same SAH cost evaluated at kBins uniform planes per axis, with the
// same partition and clipping rules below the chosen plane; smaller nodes use the exact sweep.  Its
// trees differ from the reference's but satisfy the same invariants (every triangle overlapping a
// leaf's box is listed in the leaf), so traversal returns the same closest hits.
#include "kdtree_builder.hpp"

#include <memory>

#include <algorithm>
#include <atomic>
#include <thread>

namespace yar {

static inline uint32_t float_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
// Monotone map float -> uint32 with -0 and +0 mapped to the same key (they compare equal).
static inline uint32_t order_key(float f) {
    uint32_t u = float_bits(f);
    if (u == 0x80000000u) u = 0;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

class KdBuilder {
public:
    struct Prim { uint32_t id; Box3 b; };

    KdBuilder(const float* verts, const int32_t* idx, bool binned = false) : verts_(verts), idx_(idx), binned_(binned) {}

    KdBuildResult build(std::vector<Prim>& prims) {
        KdBuildResult r;
        uint32_t n = (uint32_t)prims.size();
        for (const Prim& p : prims) r.bounds = box_union(r.bounds, p.b);
        int max_depth = std::min((int)std::lround(8.0 + 1.3 * std::floor(std::log2((double)n))), 40);
        words_.clear();
        words_.reserve(4 * (size_t)n + 16);
        deepest_ = 0;
        levels_.resize(max_depth + 2);
        build_node(r.bounds, prims, max_depth, 0);
        r.words.swap(words_);
        r.depth = deepest_;
        return r;
    }

private:
    static constexpr uint64_t START = 1ull << 31;

    struct Level {   // per-depth scratch, reused across the nodes of one depth
        std::vector<uint64_t> keys[3];
        std::vector<Prim> below, above;
    };

    static constexpr int kBins = 32;
    static constexpr uint32_t kBinnedMin = 1024;

    const float* verts_;
    const int32_t* idx_;
    bool binned_ = false;
    std::vector<uint32_t> words_;
    std::vector<Level> levels_;
    int deepest_ = 0;

    static Box3 box_union(const Box3& a, const Box3& b) {
        Box3 r;
        for (int i = 0; i < 3; i++) { r.lo[i] = std::min(a.lo[i], b.lo[i]); r.hi[i] = std::max(a.hi[i], b.hi[i]); }
        return r;
    }

    void tri(uint32_t t, F3 p[3]) const {
        for (int k = 0; k < 3; k++) {
            const float* q = verts_ + 3 * (size_t)idx_[3 * (size_t)t + k];
            p[k] = F3(q[0], q[1], q[2]);
        }
    }

    // clip_bounds, kdtree_builder.cpp:17-93
    void clip(uint32_t t, float split, int axis, bool left, Box3& bounds) const {
        if (left) bounds.hi[axis] = split; else bounds.lo[axis] = split;
        F3 p[3];
        tri(t, p);
        if (p[1][axis] < p[0][axis]) std::swap(p[1], p[0]);
        if (p[2][axis] < p[0][axis]) std::swap(p[2], p[0]);
        if (p[2][axis] < p[1][axis]) std::swap(p[2], p[1]);
        bool middle_left = p[1][axis] < split;
        F3 a, c, b = p[1];
        if (middle_left) { a = p[2]; c = p[0]; } else { a = p[0]; c = p[2]; }
        float eps = ((left && !middle_left) || (!left && middle_left)) ? 1e-5f : -1e-5f;
        F3 ab = b - a;
        F3 iab = a + ab * ((split - a[axis]) / ab[axis] + eps);
        F3 ac = c - a;
        F3 iac = a + ac * ((split - a[axis]) / ac[axis] + eps);
        Box3 b2;
        auto add = [&](F3 q) { for (int i = 0; i < 3; i++) { b2.lo[i] = std::min(b2.lo[i], q[i]); b2.hi[i] = std::max(b2.hi[i], q[i]); } };
        add(iab); add(iac);
        if (left) { add(p[0]); if (middle_left) add(p[1]); }
        else { add(p[2]); if (!middle_left) add(p[1]); }
        for (int i = 0; i < 3; i++) { bounds.lo[i] = std::max(bounds.lo[i], b2.lo[i]); bounds.hi[i] = std::min(bounds.hi[i], b2.hi[i]); }
    }

    static inline float key_pos(uint64_t k, const std::vector<Prim>& prims, int axis) {
        uint32_t i = (uint32_t)(k & 0x7fffffffu);
        return (k & START) ? prims[i].b.lo[axis] : prims[i].b.hi[axis];
    }

    // select_split_for_axis, kdtree_builder.cpp:359-423
    float split_cost(const Box3& nb, const std::vector<Prim>& prims, const std::vector<uint64_t>& e, int axis,
                     uint32_t* split_edge) const {
        static const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        const int o0 = other[axis][0], o1 = other[axis][1];
        const uint32_t n = (uint32_t)prims.size();
        F3 diag = nb.hi - nb.lo;
        const float s0 = 2.0f * (diag[o0] * diag[o1]);
        const float d0 = 2.0f * (diag[o0] + diag[o1]);
        const float inv_total_area = 1.0f / (2.0f * (diag.x * diag.y + diag.x * diag.z + diag.y * diag.z));
        const uint32_t ne = 2 * n;
        float best_cost = (float)n;
        uint32_t best_edge = 0xffffffffu;
        uint32_t nbelow = 0, nabove = n;
        uint32_t i = 0;
        const float lo = nb.lo[axis], hi = nb.hi[axis];
        while (i < ne) {
            const float t = key_pos(e[i], prims, axis);
            uint32_t gend = i + 1;
            while (gend < ne && t == key_pos(e[gend], prims, axis)) gend++;
            uint32_t mid = i;
            while (mid != gend && !(e[mid] & START)) mid++;
            nabove -= mid - i;
            if (t > lo && t < hi) {
                float below_area = s0 + d0 * (t - lo);
                float above_area = s0 + d0 * (hi - t);
                float pb = below_area * inv_total_area;
                float pa = above_area * inv_total_area;
                float bonus = (nbelow == 0 || nabove == 0) ? 0.3f : 0.0f;
                float expected = pb * nbelow + pa * nabove;
                float cost = (1.f - bonus) * expected;
                if (cost < best_cost) { best_edge = (mid == gend) ? mid - 1 : mid; best_cost = cost; }
            }
            nbelow += gend - mid;
            i = gend;
        }
        if (best_edge == 0xffffffffu) return kInf;
        *split_edge = best_edge;
        return best_cost;
    }

    void make_leaf(const std::vector<Prim>& prims, int level) {
        deepest_ = std::max(deepest_, level);
        size_t node = words_.size() / 2;
        uint32_t n = (uint32_t)prims.size();
        words_.push_back(3u | (n << 2));
        words_.push_back(0xffffffffu);
        if (n > 1) words_.insert(words_.end(), 2 * (size_t)(n / 2), 0u);
        uint32_t* ids = &words_[2 * node + 1];
        for (uint32_t i = 0; i < n; i++) ids[i] = prims[i].id;
        std::sort(ids, ids + n);
    }

    // SAH at kBins - 1 uniform planes per axis (the cost formula of split_cost): primitives reaching
    // below a plane t are those with lo < t, above it those with hi > t, counted from per-bin
    // histograms of lo and hi.  Returns the best (axis, plane) whose cost beats the leaf cost n.
    bool binned_split(const Box3& nb, const std::vector<Prim>& prims, int* out_axis, float* out_split) const {
        static const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        const uint32_t n = (uint32_t)prims.size();
        F3 diag = nb.hi - nb.lo;
        const float inv_total_area = 1.0f / (2.0f * (diag.x * diag.y + diag.x * diag.z + diag.y * diag.z));
        float best_cost = (float)n;
        bool found = false;
        for (int axis = 0; axis < 3; axis++) {
            const float lo = nb.lo[axis], hi = nb.hi[axis];
            if (!(hi > lo)) continue;
            const int o0 = other[axis][0], o1 = other[axis][1];
            const float s0 = 2.0f * (diag[o0] * diag[o1]);
            const float d0 = 2.0f * (diag[o0] + diag[o1]);
            const float scale = (float)kBins / (hi - lo);
            uint32_t starts[kBins] = {}, ends[kBins] = {};
            for (const Prim& p : prims) {
                const int b0 = std::min(kBins - 1, std::max(0, (int)((p.b.lo[axis] - lo) * scale)));
                const int b1 = std::min(kBins - 1, std::max(0, (int)((p.b.hi[axis] - lo) * scale)));
                starts[b0]++;
                ends[b1]++;
            }
            uint32_t nbelow = 0, nabove = n;
            for (int j = 1; j < kBins; j++) {
                nbelow += starts[j - 1];
                nabove -= ends[j - 1];
                const float t = lo + (hi - lo) * ((float)j / (float)kBins);
                if (!(t > lo && t < hi)) continue;
                const float pb = (s0 + d0 * (t - lo)) * inv_total_area;
                const float pa = (s0 + d0 * (hi - t)) * inv_total_area;
                const float bonus = (nbelow == 0 || nabove == 0) ? 0.3f : 0.0f;
                const float cost = (1.f - bonus) * (pb * nbelow + pa * nabove);
                if (cost < best_cost) { best_cost = cost; *out_axis = axis; *out_split = t; found = true; }
            }
        }
        return found;
    }

    void build_node(const Box3& nb, const std::vector<Prim>& prims, int depth, int level) {
        const uint32_t n = (uint32_t)prims.size();
        if (n <= 2u || depth == 0) { make_leaf(prims, level); return; }
        Level& L = levels_[level];
        if (binned_ && n > kBinnedMin) {
            int axis = -1;
            float split = 0.f;
            if (!binned_split(nb, prims, &axis, &split)) { make_leaf(prims, level); return; }
            L.below.clear();
            L.above.clear();
            for (const Prim& src : prims) {
                const bool planar_on = src.b.lo[axis] == src.b.hi[axis] && src.b.hi[axis] == split;
                if (src.b.lo[axis] < split || planar_on) {
                    Prim p = src;
                    if (p.b.hi[axis] > split && verts_) clip(p.id, split, axis, true, p.b);
                    L.below.push_back(p);
                }
                if (src.b.hi[axis] > split) {
                    Prim p = src;
                    if (p.b.lo[axis] < split && verts_) clip(p.id, split, axis, false, p.b);
                    L.above.push_back(p);
                }
            }
            emit_interior(nb, axis, split, depth, level);
            return;
        }
        float best_cost = kInf;
        int best_axis = -1;
        uint32_t best_edge = 0;
        for (int axis = 0; axis < 3; axis++) {
            std::vector<uint64_t>& e = L.keys[axis];
            e.resize(2 * (size_t)n);
            for (uint32_t i = 0; i < n; i++) {
                e[2 * i] = ((uint64_t)order_key(prims[i].b.lo[axis]) << 32) | START | i;
                e[2 * i + 1] = ((uint64_t)order_key(prims[i].b.hi[axis]) << 32) | i;
            }
            std::sort(e.begin(), e.end());
            uint32_t ed = 0;
            float cost = split_cost(nb, prims, e, axis, &ed);
            if (cost < best_cost) { best_cost = cost; best_axis = axis; best_edge = ed; }
        }
        if (best_axis == -1) { make_leaf(prims, level); return; }
        const int axis = best_axis;
        const std::vector<uint64_t>& e = L.keys[axis];
        const float split = key_pos(e[best_edge], prims, axis);
        L.below.clear();
        L.above.clear();
        for (uint32_t i = 0; i < best_edge; i++) {
            uint64_t k = e[i];
            const Prim& src = prims[(uint32_t)(k & 0x7fffffffu)];
            bool planar = src.b.lo[axis] == src.b.hi[axis];
            if ((k & START) || (planar && src.b.hi[axis] == split)) {
                Prim p = src;
                if (p.b.hi[axis] > split && verts_) clip(p.id, split, axis, true, p.b);
                L.below.push_back(p);
            }
        }
        for (uint32_t i = best_edge + 1; i < 2 * n; i++) {
            uint64_t k = e[i];
            if (!(k & START)) {
                Prim p = prims[(uint32_t)(k & 0x7fffffffu)];
                if (p.b.lo[axis] < split && verts_) clip(p.id, split, axis, false, p.b);
                L.above.push_back(p);
            }
        }
        emit_interior(nb, axis, split, depth, level);
    }

    // Interior node over this level's child lists, then both subtrees (below first).
    void emit_interior(const Box3& nb, int axis, float split, int depth, int level) {
        Level& L = levels_[level];
        // The recursion below reuses deeper levels only, so this level's child lists stay valid;
        // the above list is moved out because the below subtree does not touch it.
        std::vector<Prim> above;
        above.swap(L.above);
        size_t self = words_.size() / 2;
        words_.push_back(0);
        words_.push_back(0);
        Box3 b0 = nb; b0.hi[axis] = split;
        build_node(b0, L.below, depth - 1, level + 1);
        uint32_t above_child = (uint32_t)(words_.size() / 2);
        words_[2 * self] = (uint32_t)axis | (above_child << 2);
        words_[2 * self + 1] = float_bits(split);
        Box3 b1 = nb; b1.lo[axis] = split;
        build_node(b1, above, depth - 1, level + 1);
        L.above.swap(above);   // return the buffer for reuse
    }
};

// Triangle bounds (Triangle_Mesh::get_triangle_bounds, triangle_mesh.h:100)
static std::vector<KdBuilder::Prim> triangle_prims(const float* v, const int32_t* idx, int32_t ntri) {
    std::vector<KdBuilder::Prim> prims(ntri);
    for (int32_t t = 0; t < ntri; t++) {
        Box3 b;
        for (int k = 0; k < 3; k++) {
            const float* q = v + 3 * (size_t)idx[3 * (size_t)t + k];
            for (int a = 0; a < 3; a++) {
                if (k == 0) { b.lo[a] = q[a]; b.hi[a] = q[a]; }
                else { b.lo[a] = std::min(b.lo[a], q[a]); b.hi[a] = std::max(b.hi[a], q[a]); }
            }
        }
        prims[t] = {(uint32_t)t, b};
    }
    return prims;
}

KdBuildResult build_mesh_kdtree(const float* v, const int32_t* idx, int32_t ntri, bool binned) {
    std::vector<KdBuilder::Prim> prims = triangle_prims(v, idx, ntri);
    KdBuilder b(v, idx, binned);
    return b.build(prims);
}

KdBuildResult build_box_kdtree(const float* boxes, int32_t n) {
    std::vector<KdBuilder::Prim> prims(n);
    for (int32_t i = 0; i < n; i++) {
        Box3 b;
        for (int a = 0; a < 3; a++) { b.lo[a] = boxes[6 * (size_t)i + a]; b.hi[a] = boxes[6 * (size_t)i + 3 + a]; }
        prims[i] = {(uint32_t)i, b};
    }
    KdBuilder bld(nullptr, nullptr);
    return bld.build(prims);
}

// Builds one tree per mesh on `threads` workers (load_geometry_kdtrees, scene_context.cpp:33-61).
std::vector<KdBuildResult> build_mesh_kdtrees(const yar_scene_desc* d, int threads, bool binned) {
    std::vector<KdBuildResult> out(d->mesh_count);
    std::atomic<int> next{0};
    // Largest meshes first so one big mesh does not finish last on a single worker.
    std::vector<int> order(d->mesh_count);
    for (int i = 0; i < d->mesh_count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return d->meshes[a].triangle_count > d->meshes[b].triangle_count;
    });
    auto work = [&]() {
        for (int j = next++; j < d->mesh_count; j = next++) {
            int i = order[j];
            const yar_mesh_desc& m = d->meshes[i];
            out[i] = build_mesh_kdtree(m.vertices, m.indices, m.triangle_count, binned);
        }
    };
    int nth = std::max(1, std::min(threads, d->mesh_count));
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return out;
}

}  // namespace yar

struct yar_kdtree {
    yar::KdBuildResult r;
};

extern "C" {

static int build_mesh_entry(const char* name, const float* vertices, const int32_t* indices, int32_t triangle_count,
                            yar_kdtree** out_tree, bool binned) {
    if (!vertices || !indices || triangle_count <= 0 || !out_tree) {
        yar::set_last_error("yar_kdtree_build_mesh: invalid argument");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    if ((uint32_t)triangle_count > 0x3fffffffu) {
        yar::set_last_error("yar_kdtree_build_mesh: more than 2^30-1 primitives");
        return YAR_ERR_CAPACITY;
    }
    return yar::abi_guard(name, [&] {
        std::unique_ptr<yar_kdtree> t(new yar_kdtree);
        t->r = yar::build_mesh_kdtree(vertices, indices, triangle_count, binned);
        *out_tree = t.release();
    });
}

int yar_kdtree_build_mesh(const float* vertices, const int32_t* indices, int32_t triangle_count, yar_kdtree** out_tree) {
    return build_mesh_entry("yar_kdtree_build_mesh", vertices, indices, triangle_count, out_tree, false);
}

int yar_kdtree_build_mesh_binned(const float* vertices, const int32_t* indices, int32_t triangle_count,
                                 yar_kdtree** out_tree) {
    return build_mesh_entry("yar_kdtree_build_mesh_binned", vertices, indices, triangle_count, out_tree, true);
}

int yar_kdtree_build_boxes(const float* boxes, int32_t count, yar_kdtree** out_tree) {
    if (!boxes || count <= 0 || !out_tree) {
        yar::set_last_error("yar_kdtree_build_boxes: invalid argument");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    return yar::abi_guard("yar_kdtree_build_boxes", [&] {
        std::unique_ptr<yar_kdtree> t(new yar_kdtree);
        t->r = yar::build_box_kdtree(boxes, count);
        *out_tree = t.release();
    });
}

uint32_t yar_kdtree_node_count(const yar_kdtree* tree) { return tree ? (uint32_t)(tree->r.words.size() / 2) : 0; }
const uint32_t* yar_kdtree_nodes(const yar_kdtree* tree) { return tree ? tree->r.words.data() : nullptr; }
void yar_kdtree_bounds(const yar_kdtree* tree, float out_bounds[6]) {
    if (!tree || !out_bounds) return;
    for (int a = 0; a < 3; a++) { out_bounds[a] = tree->r.bounds.lo[a]; out_bounds[3 + a] = tree->r.bounds.hi[a]; }
}
int32_t yar_kdtree_max_depth(const yar_kdtree* tree) { return tree ? tree->r.depth : -1; }

// KdTree::save layout (kdtree.cpp:140-160): Bounding_Box (24 B), geometry hash (u64, always 0 in
// the reference, :205-215), node count (u32), nodes.
int yar_kdtree_save(const yar_kdtree* tree, const char* path) {
    if (!tree || !path) { yar::set_last_error("yar_kdtree_save: invalid argument"); return YAR_ERR_INVALID_ARGUMENT; }
    FILE* f = std::fopen(path, "wb");
    if (!f) { yar::set_last_error(std::string("cannot write ") + path); return YAR_ERR_IO; }
    float b[6];
    yar_kdtree_bounds(tree, b);
    uint64_t hash = 0;
    uint32_t count = yar_kdtree_node_count(tree);
    bool ok = std::fwrite(b, 4, 6, f) == 6 && std::fwrite(&hash, 8, 1, f) == 1 && std::fwrite(&count, 4, 1, f) == 1 &&
              std::fwrite(tree->r.words.data(), 8, count, f) == count;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { yar::set_last_error(std::string("write failed: ") + path); return YAR_ERR_IO; }
    return YAR_OK;
}

void yar_kdtree_free(yar_kdtree* tree) { delete tree; }

}  // extern "C"

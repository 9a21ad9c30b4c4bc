// kdtree_builder.hpp - host kd-tree builder interface shared by the C ABI and the GPU context.
#pragma once

#include "common.hpp"

namespace yar {

struct Box3 {
    F3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
};

struct KdBuildResult {
    std::vector<uint32_t> words;   // 2 words per KdNode (kdtree.h:18-77)
    Box3 bounds;                   // union of primitive bounds (KdTree::bounds)
    int depth = 0;                 // level of the deepest leaf, root = 0
};

KdBuildResult build_mesh_kdtree(const float* v, const int32_t* idx, int32_t ntri, bool binned = false);
KdBuildResult build_box_kdtree(const float* boxes, int32_t n);
std::vector<KdBuildResult> build_mesh_kdtrees(const yar_scene_desc* d, int threads, bool binned = false);

}  // namespace yar

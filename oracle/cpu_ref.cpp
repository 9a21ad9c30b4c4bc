// cpu_ref.cpp - CPU ORACLE for the RAY path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain scalar restatement of the reference renderer's
// algorithm (kennyalive/YAR src/ref + src/lib) used as the checker for the HIP path and as the
// host-CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product library never links or calls it.
//
// Every function cites the reference file:line it follows.  Floating-point expressions keep the
// reference's operation order, and the file is compiled with -ffp-contract=off, so results are
// bit-identical to a contraction-free build of the reference wherever the same libm is used.
//
// Pinning (what checks each part, see tests/):
//   pcg32 ............ pcg-c-basic demo vectors + the renderer's pixel-(0,0) stream + oracle/_ref
//                      (the reference's own third-party/pcg/pcg_basic.c compiled unchanged)
//   sampler .......... stratification/permutation invariants (each stratum hit exactly once per
//                      dimension), exact RNG draw counts, independent numpy restatement
//   watertight test .. closed-form fixtures of test_triangle_intersection.cpp:139-279, point-on-ray
//                      consistency, watertightness on a closed sphere
//   kd-tree builder .. structural invariants (every triangle overlapping a leaf's cell is listed,
//                      depth limit), and kd traversal == brute force exactly
//   traversal ........ kd result == brute-force result (t bit-exact, test_kdtree.cpp:138-194)
//   BSDF / sampling .. white furnace, Fresnel closed forms, VNDF pdf normalisation, sample/pdf
//                      consistency
//   whole estimator .. closed-box furnace convergence (analytic radiance)
//   film ............. box-filter mean closed form, jitter-0 neighbour rule
// The reference renderer itself is unbuildable here (src/lib/common.h:9 expands ASSERT to the
// MSVC intrinsic __debugbreak and src/std.h targets MSVC); only pcg_basic.c builds unchanged.

#include "../include/yar_gpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace oref {

constexpr float Pi = 3.14159265f;                   // src/lib/math.h:9
constexpr float Pi_Inv = 1.f / Pi;                  // src/lib/math.h:12
constexpr float One_Minus_Epsilon = 0x1.fffffep-1;  // src/lib/math.h:14
constexpr float Pi2 = 6.2831853f;                   // src/lib/math.h:10
constexpr float Pi2_Inv = 1.f / Pi2;                 // src/lib/math.h:13
constexpr float Inf = std::numeric_limits<float>::infinity();

static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ---------------------------------------------------------------------------------------------
// Vector math in the reference's operation order (src/lib/vector.h:12-180)
// ---------------------------------------------------------------------------------------------
struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return (&x)[i]; }
    float& operator[](int i) { return (&x)[i]; }
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }
static inline V3 operator*(float t, V3 a) { return a * t; }
static inline V3 vdiv(V3 a, float t) { return {a.x / t, a.y / t, a.z / t}; }       // vector.h:85
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // vector.h:170
static inline V3 cross(V3 a, V3 b) {                                                 // vector.h:174
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float len_sq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float length(V3 a) { return std::sqrt(len_sq(a)); }
static inline V3 normalized(V3 a) { return vdiv(a, length(a)); }                     // vector.h:100

// ColorRGB (src/lib/color.h:14-125): scalar division multiplies by the reciprocal.
struct RGB { float r = 0, g = 0, b = 0; };
static inline RGB operator*(RGB a, RGB b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
static inline RGB operator*(RGB a, float k) { return {a.r * k, a.g * k, a.b * k}; }
static inline RGB operator*(float k, RGB a) { return a * k; }
static inline RGB operator+(RGB a, RGB b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
static inline RGB operator-(RGB a, RGB b) { return {a.r - b.r, a.g - b.g, a.b - b.b}; }
static inline RGB cdiv(RGB a, RGB b) { return {a.r / b.r, a.g / b.g, a.b / b.b}; }
static inline RGB cdivs(RGB a, float k) { float inv = 1.f / k; return a * inv; }   // color.h:104
static inline RGB csqrt(RGB a) { return {std::sqrt(a.r), std::sqrt(a.g), std::sqrt(a.b)}; }
static inline RGB splat(float v) { return {v, v, v}; }
static inline bool is_black(RGB a) { return a.r == 0 && a.g == 0 && a.b == 0; }
static inline float luminance(RGB c) {                                               // color.cpp:18
    return 0.2126729f * c.r + 0.7151522f * c.g + 0.0721750f * c.b;
}

struct M34 { float a[3][4]; };
static inline V3 xform_point(const M34& m, V3 p) {                                  // matrix.cpp:371
    return {m.a[0][0] * p.x + m.a[0][1] * p.y + m.a[0][2] * p.z + m.a[0][3],
            m.a[1][0] * p.x + m.a[1][1] * p.y + m.a[1][2] * p.z + m.a[1][3],
            m.a[2][0] * p.x + m.a[2][1] * p.y + m.a[2][2] * p.z + m.a[2][3]};
}
static inline V3 xform_vector(const M34& m, V3 v) {                                 // matrix.cpp:379
    return {m.a[0][0] * v.x + m.a[0][1] * v.y + m.a[0][2] * v.z,
            m.a[1][0] * v.x + m.a[1][1] * v.y + m.a[1][2] * v.z,
            m.a[2][0] * v.x + m.a[2][1] * v.y + m.a[2][2] * v.z};
}
static inline M34 load_m34(const float* f) { M34 m; std::memcpy(m.a, f, sizeof(m.a)); return m; }
static inline V3 column(const M34& m, int c) { return {m.a[0][c], m.a[1][c], m.a[2][c]}; }

struct Ray { V3 o, d; };

// ---------------------------------------------------------------------------------------------
// PCG32 (third-party/pcg/pcg_basic.c:46-75) and RNG (src/lib/random.h:7-74)
// ---------------------------------------------------------------------------------------------
struct Pcg { uint64_t state = 0, inc = 1; };
static inline uint32_t pcg_next(Pcg& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}
static inline void pcg_seed(Pcg& r, uint64_t initstate, uint64_t initseq) {
    r.state = 0;
    r.inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r.state += initstate;
    pcg_next(r);
}
static inline float rng_float(Pcg& r) { return bitsf((pcg_next(r) >> 9) | 0x3f800000u) - 1.0f; }
static inline uint32_t rng_bounded_fast(Pcg& r, uint32_t bound) {                  // random.h:54
    uint64_t m = (uint64_t)pcg_next(r) * bound;
    return (uint32_t)(m >> 32);
}

// ---------------------------------------------------------------------------------------------
// Stratified pixel sampler (src/ref/pixel_sampling.cpp:54-173, sampling.cpp:79-96, math.h:138)
// ---------------------------------------------------------------------------------------------
struct V2 { float x = 0, y = 0; };

struct PixelSampler {
    int nx = 1, ny = 1, spp = 1, d1 = 0, d2 = 0;
    std::vector<V2> image;    // [spp]
    std::vector<float> s1;    // [spp * d1], sample-major as in the reference
    std::vector<V2> s2;       // [spp * d2]
    int cur = 0, cur1 = 0, cur2 = 0;
    Pcg* rng = nullptr;

    void init(int nx_, int ny_, int d1_, int d2_, Pcg* r) {
        nx = nx_; ny = ny_; spp = nx * ny; d1 = d1_; d2 = d2_; rng = r;
        image.resize(spp); s1.resize((size_t)spp * d1); s2.resize((size_t)spp * d2);
    }
    void strat1d(int n, float* out) {
        float dx = 1.f / float(n);
        for (int x = 0; x < n; x++) out[x] = std::min((float(x) + rng_float(*rng)) * dx, One_Minus_Epsilon);
    }
    void strat2d(int sx, int sy, V2* out) {
        float dx = 1.f / float(sx), dy = 1.f / float(sy);
        int k = 0;
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++, k++) {
                out[k].x = std::min((float(x) + rng_float(*rng)) * dx, One_Minus_Epsilon);
                out[k].y = std::min((float(y) + rng_float(*rng)) * dy, One_Minus_Epsilon);
            }
    }
    template <class T> void shuffle(T* data, int n) {
        for (int i = n; i > 1; i--) {
            uint32_t k = rng_bounded_fast(*rng, (uint32_t)i);
            std::swap(data[i - 1], data[k]);
        }
    }
    void next_pixel() {
        cur = 0; cur1 = 0; cur2 = 0;
        strat2d(nx, ny, image.data());
        std::vector<float> t1(spp);
        for (int i = 0; i < d1; i++) {
            strat1d(spp, t1.data());
            shuffle(t1.data(), spp);
            for (int k = 0; k < spp; k++) s1[(size_t)k * d1 + i] = t1[k];
        }
        std::vector<V2> t2(spp);
        for (int i = 0; i < d2; i++) {
            strat2d(nx, ny, t2.data());
            shuffle(t2.data(), spp);
            for (int k = 0; k < spp; k++) s2[(size_t)k * d2 + i] = t2[k];
        }
    }
    bool next_sample() {
        if (cur < spp) { cur++; cur1 = 0; cur2 = 0; }
        return cur < spp;
    }
    float next1d() {
        if (cur1 < d1) return s1[(size_t)cur * d1 + cur1++];
        return rng_float(*rng);
    }
    V2 next2d() {
        if (cur2 < d2) return s2[(size_t)cur * d2 + cur2++];
        V2 v; v.x = rng_float(*rng); v.y = rng_float(*rng);
        return v;
    }
};

// ---------------------------------------------------------------------------------------------
// Watertight ray/triangle test (src/ref/intersection.cpp:53-113)
// ---------------------------------------------------------------------------------------------
static inline int max_dim(V3 v) { return (v.x > v.y) ? (v.x > v.z ? 0 : 2) : (v.y > v.z ? 1 : 2); }
static inline V3 perm(V3 v, int kx, int ky, int kz) { return {v[kx], v[ky], v[kz]}; }

static float intersect_watertight(const Ray& ray, V3 p0, V3 p1, V3 p2, V3* b) {
    V3 ad{std::fabs(ray.d.x), std::fabs(ray.d.y), std::fabs(ray.d.z)};
    const int kz = max_dim(ad);
    const int kx = (kz == 2 ? 0 : kz + 1);
    const int ky = (kz == 0 ? 2 : kz - 1);
    V3 d = perm(ray.d, kx, ky, kz);
    float sx = -d.x / d.z;
    float sy = -d.y / d.z;
    float sz = 1.f / d.z;
    const V3 p0t = perm(p0 - ray.o, kx, ky, kz);
    const V3 p1t = perm(p1 - ray.o, kx, ky, kz);
    const V3 p2t = perm(p2 - ray.o, kx, ky, kz);
    float x0 = p0t.x + sx * p0t.z, y0 = p0t.y + sy * p0t.z;
    float x1 = p1t.x + sx * p1t.z, y1 = p1t.y + sy * p1t.z;
    float x2 = p2t.x + sx * p2t.z, y2 = p2t.y + sy * p2t.z;
    float e0 = x1 * y2 - y1 * x2;
    float e1 = x2 * y0 - y2 * x0;
    float e2 = x0 * y1 - y0 * x1;
    if (e0 == 0.f || e1 == 0.f || e2 == 0.f) {
        e0 = float(double(x1) * double(y2) - double(y1) * double(x2));
        e1 = float(double(x2) * double(y0) - double(y2) * double(x0));
        e2 = float(double(x0) * double(y1) - double(y0) * double(x1));
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return Inf;
    float det = e0 + e1 + e2;
    if (det == 0.f) return Inf;
    float z0 = sz * p0t.z, z1 = sz * p1t.z, z2 = sz * p2t.z;
    float t_scaled = e0 * z0 + e1 * z1 + e2 * z2;
    if ((fbits(det) ^ fbits(t_scaled)) >> 31) return Inf;
    float inv_det = 1.f / det;
    float t = inv_det * t_scaled;
    b->x = e0 * inv_det; b->y = e1 * inv_det; b->z = e2 * inv_det;
    return t;
}

// ---------------------------------------------------------------------------------------------
// kd-tree builder (src/ref/kdtree_builder.cpp:8-463, kdtree.cpp:199-203)
// ---------------------------------------------------------------------------------------------
struct Box { V3 lo{Inf, Inf, Inf}, hi{-Inf, -Inf, -Inf}; };
static inline void box_add(Box& b, V3 p) {
    b.lo.x = std::min(b.lo.x, p.x); b.lo.y = std::min(b.lo.y, p.y); b.lo.z = std::min(b.lo.z, p.z);
    b.hi.x = std::max(b.hi.x, p.x); b.hi.y = std::max(b.hi.y, p.y); b.hi.z = std::max(b.hi.z, p.z);
}
static inline Box box_union(const Box& a, const Box& b) {
    Box r;
    r.lo = {std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z)};
    r.hi = {std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z)};
    return r;
}
static inline Box box_intersection(const Box& a, const Box& b) {
    Box r;
    r.lo = {std::max(a.lo.x, b.lo.x), std::max(a.lo.y, b.lo.y), std::max(a.lo.z, b.lo.z)};
    r.hi = {std::min(a.hi.x, b.hi.x), std::min(a.hi.y, b.hi.y), std::min(a.hi.z, b.hi.z)};
    return r;
}

static int depth_limit(uint32_t n) {                                                // kdtree.cpp:199
    int depth = (int)std::lround(8.0 + 1.3 * std::floor(std::log2((double)n)));
    return std::min(depth, 40);
}

struct TriSource {   // triangle positions for clipping (nullptr source for the object-level tree)
    const float* v = nullptr; const int32_t* idx = nullptr;
    void get(uint32_t t, V3 p[3]) const {
        for (int k = 0; k < 3; k++) { const float* q = v + 3 * (size_t)idx[3 * (size_t)t + k]; p[k] = {q[0], q[1], q[2]}; }
    }
};

struct Builder {
    static constexpr uint32_t END = 0x80000000u, PLANAR = 0x40000000u, MASK = 0x3fffffffu;
    struct Edge { float pos; uint32_t bits; };
    struct Prim { uint32_t prim; Box b; };
    const TriSource* tris = nullptr;
    int max_depth = 0;
    std::vector<uint32_t> words;   // 2 per node
    int deepest = 0;

    // Soupikov et al. 2008 split clipping, kdtree_builder.cpp:17-93
    void clip(uint32_t tri, float split, int axis, bool left, Box& bounds) const {
        if (left) bounds.hi[axis] = split; else bounds.lo[axis] = split;
        V3 p[3]; tris->get(tri, p);
        if (p[1][axis] < p[0][axis]) std::swap(p[1], p[0]);
        if (p[2][axis] < p[0][axis]) std::swap(p[2], p[0]);
        if (p[2][axis] < p[1][axis]) std::swap(p[2], p[1]);
        bool middle_left = p[1][axis] < split;
        V3 a, c, b = p[1];
        if (middle_left) { a = p[2]; c = p[0]; } else { a = p[0]; c = p[2]; }
        float eps = ((left && !middle_left) || (!left && middle_left)) ? 1e-5f : -1e-5f;
        V3 ab = b - a;
        V3 iab = a + ab * ((split - a[axis]) / ab[axis] + eps);
        V3 ac = c - a;
        V3 iac = a + ac * ((split - a[axis]) / ac[axis] + eps);
        Box b2;
        box_add(b2, iab); box_add(b2, iac);
        if (left) { box_add(b2, p[0]); if (middle_left) box_add(b2, p[1]); }
        else { box_add(b2, p[2]); if (!middle_left) box_add(b2, p[1]); }
        bounds = box_intersection(bounds, b2);
    }

    static bool edge_less(const Edge& a, const Edge& b) {                           // kdtree_builder.cpp:118
        if (a.pos == b.pos) return (a.bits & END) && !(b.bits & END);
        return a.pos < b.pos;
    }

    // kdtree_builder.cpp:359-423
    float split_for_axis(const Box& nb, const std::vector<Edge>& e, uint32_t n, int axis, uint32_t* split_edge) const {
        static const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        const int o0 = other[axis][0], o1 = other[axis][1];
        V3 diag = nb.hi - nb.lo;
        const float s0 = 2.0f * (diag[o0] * diag[o1]);
        const float d0 = 2.0f * (diag[o0] + diag[o1]);
        const float inv_total_area = 1.0f / (2.0f * (diag.x * diag.y + diag.x * diag.z + diag.y * diag.z));
        const uint32_t ne = 2 * n;
        float best_cost = (float)n;
        uint32_t best_edge = 0xffffffffu;
        uint32_t nbelow = 0, nabove = n;
        uint32_t i = 0;
        while (i < ne) {
            const float t = e[i].pos;
            uint32_t gend = i + 1;
            while (gend < ne && t == e[gend].pos) gend++;
            uint32_t mid = i;
            while (mid != gend && (e[mid].bits & END)) mid++;
            nabove -= mid - i;
            if (t > nb.lo[axis] && t < nb.hi[axis]) {
                float below_area = s0 + d0 * (t - nb.lo[axis]);
                float above_area = s0 + d0 * (nb.hi[axis] - t);
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
        if (best_edge == 0xffffffffu) return Inf;
        *split_edge = best_edge;
        return best_cost;
    }

    void make_leaf(const std::vector<Prim>& prims, int level) {                   // kdtree_builder.cpp:310
        deepest = std::max(deepest, level);
        size_t node = words.size() / 2;
        uint32_t n = (uint32_t)prims.size();
        words.push_back(3u | (n << 2));
        words.push_back(0xffffffffu);
        if (n > 1) words.insert(words.end(), 2 * (size_t)(n / 2), 0u);
        std::vector<uint32_t> ids(n);
        for (uint32_t i = 0; i < n; i++) ids[i] = prims[i].prim;
        std::sort(ids.begin(), ids.end());
        for (uint32_t i = 0; i < n; i++) words[2 * node + 1 + i] = ids[i];
    }

    void build_node(const Box& nb, const std::vector<Prim>& prims, int depth, int level) {   // :207
        uint32_t n = (uint32_t)prims.size();
        if (n <= 2u || depth == 0) { make_leaf(prims, level); return; }
        std::vector<Edge> edges[3];
        float best_cost = Inf; int best_axis = -1; uint32_t best_edge = 0;
        for (int axis = 0; axis < 3; axis++) {                                     // select_split :331
            edges[axis].resize(2 * (size_t)n);
            for (uint32_t i = 0; i < n; i++) {
                const Box& b = prims[i].b;
                edges[axis][2 * i] = {b.lo[axis], i};
                edges[axis][2 * i + 1] = {b.hi[axis], i | END};
                if (b.lo[axis] == b.hi[axis]) { edges[axis][2 * i].bits |= PLANAR; edges[axis][2 * i + 1].bits |= PLANAR; }
            }
            std::stable_sort(edges[axis].begin(), edges[axis].end(), edge_less);
            uint32_t ed = 0;
            float cost = split_for_axis(nb, edges[axis], n, axis, &ed);
            if (cost < best_cost) { best_cost = cost; best_axis = axis; best_edge = ed; }
        }
        if (best_axis == -1) { make_leaf(prims, level); return; }
        const int axis = best_axis;
        const std::vector<Edge>& e = edges[axis];
        const float split = e[best_edge].pos;
        std::vector<Prim> below, above;
        for (uint32_t i = 0; i < best_edge; i++) {
            const Edge ed = e[i];
            if (!(ed.bits & END) || (ed.pos == split && (ed.bits & PLANAR))) {
                Prim p = prims[ed.bits & MASK];
                if (p.b.hi[axis] > split && tris) clip(p.prim, split, axis, true, p.b);
                below.push_back(p);
            }
        }
        for (uint32_t i = best_edge + 1; i < 2 * n; i++) {
            const Edge ed = e[i];
            if (ed.bits & END) {
                Prim p = prims[ed.bits & MASK];
                if (p.b.lo[axis] < split && tris) clip(p.prim, split, axis, false, p.b);
                above.push_back(p);
            }
        }
        size_t self = words.size() / 2;
        words.push_back(0); words.push_back(0);
        Box b0 = nb; b0.hi[axis] = split;
        build_node(b0, below, depth - 1, level + 1);
        uint32_t above_child = (uint32_t)(words.size() / 2);
        words[2 * self] = (uint32_t)axis | (above_child << 2);
        words[2 * self + 1] = fbits(split);
        Box b1 = nb; b1.lo[axis] = split;
        build_node(b1, above, depth - 1, level + 1);
    }
};

struct Tree {
    std::vector<uint32_t> words;
    Box bounds;
    int depth = 0;
    bool leaf(uint32_t n) const { return (words[2 * n] & 3u) == 3u; }
};

static Tree build_tree(const std::vector<Box>& boxes, const TriSource* tris) {
    Builder bld;
    bld.tris = tris;
    uint32_t n = (uint32_t)boxes.size();
    bld.max_depth = depth_limit(n);
    std::vector<Builder::Prim> prims(n);
    Box total;
    for (uint32_t i = 0; i < n; i++) { prims[i] = {i, boxes[i]}; total = box_union(total, boxes[i]); }
    bld.build_node(total, prims, bld.max_depth, 0);
    Tree t; t.words = std::move(bld.words); t.bounds = total; t.depth = bld.deepest;
    return t;
}

// ---------------------------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------------------------
struct Mesh {
    std::vector<float> v, n, uv;
    std::vector<int32_t> idx;
    bool reverse = false;
    Tree tree;
    void pos(uint32_t t, V3 p[3]) const {
        for (int k = 0; k < 3; k++) { const float* q = &v[3 * (size_t)idx[3 * (size_t)t + k]]; p[k] = {q[0], q[1], q[2]}; }
    }
};
struct Object { int mesh, material, light; M34 o2w, w2o, nrm; };
struct Material {
    int type; RGB refl, trans, eta, k; float eta_i, alpha_x, alpha_y, ior, roughness;
    float r0 = 0.f;           // plastic specular scale
    bool pbrt3 = false;       // plastic: the pbrt3 variant (Fresnel of a 1.5 dielectric, pbrt3's flipped ior)
    // pbrt3 uber (material_pbrt.h:51-84): Ks, Kr, opacity and the component list in the order
    // diffuse 0, specular 1, delta reflection 2, (delta transmission 3), opacity 4
    RGB ks, kr, opacity;
    int ncomp = 0, comps[5] = {};
    // mix (material.h:153, Mix_BSDF bsdf.cpp:497-570): sub-materials, amount
    int sub[2] = {-1, -1};
    RGB amount;
};
struct RectLight { M34 l2w; RGB L; V2 size; };
struct PointLight { V3 p; RGB I; };
struct SpotLight { V3 p, dir; float cone_angle, penumbra_angle; RGB I; };     // light.h:40
struct DirectionalLight { V3 dir; RGB E; };                                   // light.h:48
// Diffuse_Triangle_Mesh_Light + its sampler (light.h:85, light_sampling.h:48): the emitter mesh,
// its area and the area CDF of Distribution_1D (sampling.cpp:149-166)
struct MeshLight { int object, mesh; RGB L; float area; std::vector<float> cdf; };
struct SphereLight { V3 p; float radius; RGB L; };                           // light.h:78
struct EnvLight { bool present = false; M34 l2w, w2l; RGB L; };             // light.h:91, constant map

struct Counters { int64_t nodes = 0, leaf_refs = 0, tri_tests = 0, objects = 0; };

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Object> objects;
    std::vector<Material> materials;
    std::vector<RectLight> rects;
    std::vector<PointLight> points;
    std::vector<SpotLight> spots;
    std::vector<DirectionalLight> dirs;
    std::vector<MeshLight> mesh_lights;
    std::vector<SphereLight> spheres;
    EnvLight env;
    Tree scene_tree;
    yar_camera_desc camera;
    yar_render_config config;
    int total_lights = 0;
};

// Hit record: Intersection subset (src/ref/intersection.h)
struct Hit { float t; int obj; uint32_t tri; V3 b; };

// Bounding_Box::intersect_by_ray_without_NaNs (src/lib/bounding_box.h:63-85)
static bool slab(const Box& b, const Ray& r, float* tmin, float* tmax) {
    float t0 = 0.f, t1 = Inf;
    for (int i = 0; i < 3; i++) {
        if (r.d[i] != 0.f) {
            float inv = 1.f / r.d[i];
            float s0 = (b.lo[i] - r.o[i]) * inv;
            float s1 = (b.hi[i] - r.o[i]) * inv;
            if (inv < 0.f) std::swap(s0, s1);
            t0 = s0 > t0 ? s0 : t0;
            t1 = s1 < t1 ? s1 : t1;
            if (t0 > t1) return false;
        } else if (r.o[i] < b.lo[i] || r.o[i] > b.hi[i]) {
            return false;
        }
    }
    *tmin = t0; *tmax = t1;
    return true;
}

// KdTree::intersect / intersect_any (src/ref/kdtree.cpp:219-439).  `Leaf` visits one leaf's
// primitive list and returns true to stop (any-hit).  `closest_t` points at the running
// Intersection::t for closest-hit queries (the loop condition), or at the fixed ray_tmax.
template <class Leaf>
static void traverse(const Tree& tree, const Ray& ray, const float* loop_t, Counters* cnt, Leaf&& leaf) {
    float tmin, tmax;
    if (!slab(tree.bounds, ray, &tmin, &tmax)) return;
    V3 inv{1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z};
    struct Entry { uint32_t node; float tmax; };
    Entry stack[40];
    int sp = 0;
    uint32_t node = 0;
    const uint32_t* w = tree.words.data();
    while (*loop_t > tmin) {
        if (cnt) cnt->nodes++;
        uint32_t w0 = w[2 * node];
        if ((w0 & 3u) != 3u) {
            int axis = (int)(w0 & 3u);
            float dist = bitsf(w[2 * node + 1]) - ray.o[axis];
            uint32_t below = node + 1, above = w0 >> 2;
            if (dist != 0.0) {
                uint32_t first, second;
                if (dist > 0.0) { first = below; second = above; } else { first = above; second = below; }
                float ts = dist * inv[axis];
                if (ts > tmax || ts < 0.0) node = first;
                else if (ts < tmin) node = second;
                else { stack[sp].node = second; stack[sp].tmax = tmax; sp++; node = first; tmax = ts; }
            } else {
                node = (ray.d[axis] > 0.0) ? above : below;
            }
        } else {
            uint32_t n = w0 >> 2;
            const uint32_t* ids = &w[2 * node + 1];
            if (cnt) cnt->leaf_refs += n;
            if (leaf(ids, n)) return;
            if (sp == 0) break;
            --sp;
            node = stack[sp].node;
            tmin = tmax;
            tmax = stack[sp].tmax;
        }
    }
}

// Closest hit over the two-level structure (kdtree.cpp:12-43, 81-95)
static bool intersect(const Scene& s, const Ray& ray, Hit& hit, Counters* cnt) {
    float tmax0 = hit.t;
    traverse(s.scene_tree, ray, &hit.t, cnt, [&](const uint32_t* ids, uint32_t n) {
        for (uint32_t i = 0; i < n; i++) {
            const Object& o = s.objects[ids[i]];
            const Mesh& m = s.meshes[o.mesh];
            if (cnt) cnt->objects++;
            Ray lr{xform_point(o.w2o, ray.o), xform_vector(o.w2o, ray.d)};
            float mesh_tmax = hit.t;
            traverse(m.tree, lr, &hit.t, cnt, [&](const uint32_t* tids, uint32_t tn) {
                for (uint32_t j = 0; j < tn; j++) {
                    V3 p[3]; m.pos(tids[j], p);
                    V3 b;
                    if (cnt) cnt->tri_tests++;
                    float t = intersect_watertight(lr, p[0], p[1], p[2], &b);
                    if (t < hit.t) {
                        if (len_sq(cross(p[1] - p[0], p[2] - p[0])) == 0.f) continue;
                        hit.t = t; hit.tri = tids[j]; hit.b = b;
                        hit.obj = -2;   // marks "found in this mesh"
                    }
                }
                return false;
            });
            if (hit.t < mesh_tmax) hit.obj = (int)ids[i];
        }
        return false;
    });
    return hit.t < tmax0;
}

static bool intersect_any(const Scene& s, const Ray& ray, float ray_tmax, Counters* cnt) {
    bool found = false;
    traverse(s.scene_tree, ray, &ray_tmax, cnt, [&](const uint32_t* ids, uint32_t n) {
        for (uint32_t i = 0; i < n; i++) {
            const Object& o = s.objects[ids[i]];
            const Mesh& m = s.meshes[o.mesh];
            if (cnt) cnt->objects++;
            Ray lr{xform_point(o.w2o, ray.o), xform_vector(o.w2o, ray.d)};
            bool hit = false;
            traverse(m.tree, lr, &ray_tmax, cnt, [&](const uint32_t* tids, uint32_t tn) {
                for (uint32_t j = 0; j < tn; j++) {
                    V3 p[3]; m.pos(tids[j], p);
                    V3 b;
                    if (cnt) cnt->tri_tests++;
                    float t = intersect_watertight(lr, p[0], p[1], p[2], &b);
                    if (t >= ray_tmax) continue;
                    if (len_sq(cross(p[1] - p[0], p[2] - p[0])) == 0.f) continue;
                    hit = true;
                    return true;
                }
                return false;
            });
            if (hit) { found = true; return true; }
        }
        return false;
    });
    return found;
}

// Brute force over every object and triangle (BRUTE_FORCE_INTERSECTION, kdtree.cpp:217-231)
static bool intersect_brute(const Scene& s, const Ray& ray, Hit& hit) {
    float t0 = hit.t;
    for (size_t oi = 0; oi < s.objects.size(); oi++) {
        const Object& o = s.objects[oi];
        const Mesh& m = s.meshes[o.mesh];
        Ray lr{xform_point(o.w2o, ray.o), xform_vector(o.w2o, ray.d)};
        uint32_t nt = (uint32_t)(m.idx.size() / 3);
        for (uint32_t j = 0; j < nt; j++) {
            V3 p[3]; m.pos(j, p); V3 b;
            float t = intersect_watertight(lr, p[0], p[1], p[2], &b);
            if (t < hit.t && len_sq(cross(p[1] - p[0], p[2] - p[0])) != 0.f) {
                hit.t = t; hit.tri = j; hit.b = b; hit.obj = (int)oi;
            }
        }
    }
    return hit.t < t0;
}

// ---------------------------------------------------------------------------------------------
// Shading context (src/ref/shading_context.cpp:17-240, 405-475; src/lib/math.cpp:21-120)
// ---------------------------------------------------------------------------------------------
static void coordinate_system(V3 v, V3* v1, V3* v2) {                                // math.cpp:21
    *v1 = (std::fabs(v.x) > std::fabs(v.y)) ? V3(-v.z, 0, v.x) : V3(0, -v.z, v.y);
    *v1 = vdiv(*v1, length(*v1));
    *v2 = cross(v, *v1);
}
static V3 project_dir(V3 v, V3 n) { V3 t = cross(v, n); V3 d = cross(n, t); return normalized(d); }   // math.cpp:34
static inline V3 reflect(V3 w, V3 n) { return (2.f * dot(w, n)) * n - w; }          // math.h:111

static void offset_origins(V3 p, V3 ng, V3* pos, V3* neg) {                         // math.cpp:81-120
    const float int_scale = 256.f, origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f;
    int di[3] = {int(int_scale * ng.x), int(int_scale * ng.y), int(int_scale * ng.z)};
    for (int a = 0; a < 3; a++) {
        int ib = (int)fbits(p[a]);
        int i1 = ib - di[a], i2 = ib + di[a];
        (*pos)[a] = bitsf((uint32_t)(p[a] < 0 ? i1 : i2));
        (*neg)[a] = bitsf((uint32_t)(p[a] < 0 ? i2 : i1));
        if (std::fabs(p[a]) < origin) {
            float d = float_scale * ng[a];
            (*pos)[a] = p[a] + d;
            (*neg)[a] = p[a] - d;
        }
    }
}

static bool adjust_normal(V3 wo, V3 ng, V3* n) {                                      // shading_context.cpp:17
    V3 R = reflect(wo, *n);
    float a = dot(R, ng);
    if (a >= 0.f) return false;
    if (dot(wo, ng) < 0.0017453f) { *n = ng; return true; }
    float b = dot(*n, ng);
    V3 tangent;
    if (b > 1e-4f) {
        float dist = std::fabs(a) / b;
        tangent = R + dist * (*n);
        tangent = vdiv(tangent, length(tangent));
    } else {
        tangent = *n;
    }
    tangent = tangent + 1e-4f * ng;
    *n = normalized(wo + tangent);
    return true;
}

struct Shading {
    V3 wo, pos, ng, n, dpdu_s;
    V3 org_pos, org_neg;
    bool flipped = false;
    int material = -1, light = -1;
    V3 T, B;   // BSDF frame (bsdf.cpp:15-21)
    V3 origin_dir(V3 d) const { return dot(d, ng) > 0.f ? org_pos : org_neg; }        // :405
    V3 origin_point(V3 p) const { return origin_dir(p - pos); }                      // :413
};

static void init_shading(const Scene& s, const Ray& ray, const Hit& hit, Shading& sc) {
    const Object& o = s.objects[hit.obj];
    const Mesh& m = s.meshes[o.mesh];
    sc = Shading{};
    sc.wo = -normalized(ray.d);
    V3 p[3]; m.pos(hit.tri, p);
    V3 b = hit.b;
    V3 position = b.x * p[0] + b.y * p[1] + b.z * p[2];
    V3 ng = normalized(cross(p[1] - p[0], p[2] - p[0]));
    if (m.reverse) ng = -ng;
    V3 n = ng;
    const int32_t* vi = &m.idx[3 * (size_t)hit.tri];
    if (!m.n.empty()) {
        V3 nn[3];
        for (int k = 0; k < 3; k++) { const float* q = &m.n[3 * (size_t)vi[k]]; nn[k] = {q[0], q[1], q[2]}; }
        V3 in = b.x * nn[0] + b.y * nn[1] + b.z * nn[2];
        if (len_sq(in) != 0) n = normalized(in);
    }
    bool has_uv = !m.uv.empty();
    V3 dpdu, dpdv;
    if (has_uv) {
        float u[3], v[3];
        for (int k = 0; k < 3; k++) { u[k] = m.uv[2 * (size_t)vi[k]]; v[k] = m.uv[2 * (size_t)vi[k] + 1]; }
        float a00 = u[1] - u[0], a01 = v[1] - v[0], a10 = u[2] - u[0], a11 = v[2] - v[0];
        V3 bp0 = p[1] - p[0], bp1 = p[2] - p[0];
        float det = a00 * a11 - a01 * a10;                                           // math.cpp:8
        if (det != 0.f) {
            float inv = 1.f / det;
            dpdu = (a11 * bp0 - a01 * bp1) * inv;
            dpdv = (-a10 * bp0 + a00 * bp1) * inv;
        }
        if (std::fabs(det) < 1e-9f) coordinate_system(ng, &dpdu, &dpdv);
    }
    sc.pos = xform_point(o.o2w, position);
    n = xform_vector(o.nrm, n); n = vdiv(n, length(n));
    ng = xform_vector(o.nrm, ng); ng = vdiv(ng, length(ng));
    dpdu = xform_vector(o.o2w, dpdu);
    if (dot(n, ng) == 0.f) ng = n;
    if (dot(ng, sc.wo) < 0) ng = -ng;
    if (dot(n, ng) < 0) { n = -n; sc.flipped = true; }
    offset_origins(sc.pos, ng, &sc.org_pos, &sc.org_neg);
    adjust_normal(sc.wo, ng, &n);
    V3 dpdu_s, dpdv_s;
    if (!has_uv) coordinate_system(n, &dpdu_s, &dpdv_s);
    else if (std::fabs(dot(n, ng) - 1.f) < 1e-6f) dpdu_s = dpdu;
    else { dpdu_s = project_dir(dpdu, n); dpdu_s = dpdu_s * length(dpdu); }
    sc.ng = ng; sc.n = n; sc.dpdu_s = dpdu_s;
    sc.material = o.material; sc.light = o.light;
    sc.B = normalized(cross(n, dpdu_s));
    sc.T = cross(sc.B, n);
}

// ---------------------------------------------------------------------------------------------
// BSDFs (src/ref/bsdf.cpp:15-200, scattering.cpp, sampling.cpp:43-53,311-359)
// ---------------------------------------------------------------------------------------------
static inline V3 to_world(const Shading& sc, V3 l) {
    return {sc.T.x * l.x + sc.B.x * l.y + sc.n.x * l.z,
            sc.T.y * l.x + sc.B.y * l.y + sc.n.y * l.z,
            sc.T.z * l.x + sc.B.z * l.y + sc.n.z * l.z};
}
static inline V3 to_local(const Shading& sc, V3 w) { return {dot(w, sc.T), dot(w, sc.B), dot(w, sc.n)}; }

static V3 sample_cosine(V2 u) {                                                      // sampling.cpp:43
    float z = std::sqrt(1.f - u.x);
    float r = std::sqrt(u.x);
    float phi = 2.f * Pi * u.y;
    return {r * std::cos(phi), r * std::sin(phi), z};
}

static RGB conductor_fresnel(float cos_i, float eta_i, RGB eta_t, RGB k_t) {        // scattering.cpp:39
    cos_i = std::fabs(std::clamp(cos_i, -1.f, 1.f));
    float c2 = cos_i * cos_i;
    float s2 = 1.f - c2;
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
static float dielectric_fresnel(float cos_i, float eta) {                           // scattering.cpp:17
    cos_i = std::min(std::fabs(cos_i), 1.f);
    float sin_i = std::sqrt(1.f - cos_i * cos_i);
    float sin_t = (1.f / eta) * sin_i;
    if (sin_t >= 1.f) return 1.f;
    float cos_t = std::sqrt(1.f - sin_t * sin_t);
    float Rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float Rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (Rp * Rp + Rs * Rs);
}
static float D_aniso(V3 wh, float ax, float ay) {                                   // scattering.cpp:180
    float c = std::clamp(wh.z, -1.f, 1.f);
    float c2 = c * c, s2 = 1.f - c2;
    float s2inv = 1.f / s2;
    float cp = (s2 == 0.f) ? 1.f : (wh.x * wh.x) * s2inv;
    float sp = (s2 == 0.f) ? 0.f : (wh.y * wh.y) * s2inv;
    float k = c2 + s2 * (cp / (ax * ax) + sp / (ay * ay));
    return 1.f / (Pi * ax * ay * k * k);
}
static float lambda_aniso(V3 v, float ax, float ay) {                               // scattering.cpp:206
    float c = std::clamp(v.z, -1.f, 1.f);
    float c2 = c * c, s2 = 1.f - c2;
    float tan2 = s2 / c2;
    float s2inv = 1.f / s2;
    float cp = (s2 == 0.f) ? 1.f : (v.x * v.x) * s2inv;
    float sp = (s2 == 0.f) ? 0.f : (v.y * v.y) * s2inv;
    float alpha2 = cp * ax * ax + sp * ay * ay;
    return 0.5f * (-1.f + std::sqrt(1.f + alpha2 * tan2));
}
static float G_aniso(V3 wi, V3 wo, float ax, float ay) { return 1.f / (1.f + lambda_aniso(wi, ax, ay) + lambda_aniso(wo, ax, ay)); }
static float G1_aniso(V3 v, float ax, float ay) { return 1.f / (1.f + lambda_aniso(v, ax, ay)); }

static V3 vndf_sample(V2 u, V3 wo, float ax, float ay) {                            // sampling.cpp:311
    V3 N = normalized(V3(ax * wo.x, ay * wo.y, wo.z));
    float lsq = N.x * N.x + N.y * N.y;
    V3 T1 = lsq > 0.f ? vdiv(V3(-N.y, N.x, 0), std::sqrt(lsq)) : V3(1, 0, 0);
    V3 T2 = cross(N, T1);
    float r = std::sqrt(u.x);
    float phi = 2.f * Pi * u.y;
    float t1 = r * std::cos(phi);
    float t2 = r * std::sin(phi);
    float s = 0.5f * (1.f + N.z);
    t2 = (1.f - s) * std::sqrt(1.f - t1 * t1) + s * t2;
    V3 nh = t1 * T1 + t2 * T2 + std::sqrt(std::max(0.f, 1.f - t1 * t1 - t2 * t2)) * N;
    return normalized(V3(ax * nh.x, ay * nh.y, std::max(1e-6f, nh.z)));
}
static float vndf_wi_pdf(V3 wo, V3 wh, float ax, float ay) {                        // scattering.cpp:118, sampling.cpp:349
    float G1 = G1_aniso(wo, ax, ay);
    float D = D_aniso(wh, ax, ay);
    float wh_pdf = G1 * D * std::max(0.f, dot(wo, wh)) / wo.z;
    return wh_pdf / (4 * dot(wh, wo));
}

// Isotropic GGX against the shading normal: GGX_Distribution::D, lambda, G, G1 (scattering.cpp:153-238)
// and the visible-normal reflection pdf (sampling.cpp:337-347, scattering.cpp:107-116).
static float D_iso(V3 wh, V3 n, float alpha) {
    float c = dot(wh, n);
    if (c <= 0.f) return 0.f;
    float c2 = c * c;
    float a2 = alpha * alpha;
    float k = c2 * (a2 - 1) + 1;
    return a2 / (Pi * k * k);
}
static float lambda_iso(V3 v, V3 n, float alpha) {
    float c = dot(v, n);
    float c2 = c * c;
    float t2 = std::max(0.f, (1.f - c2) / c2);
    return 0.5f * (-1.f + std::sqrt(1.f + alpha * alpha * t2));
}
static float G_iso(V3 wi, V3 wo, V3 n, float a) { return 1.f / (1.f + lambda_iso(wi, n, a) + lambda_iso(wo, n, a)); }
static float G1_iso(V3 v, V3 n, float a) { return 1.f / (1.f + lambda_iso(v, n, a)); }
static float vndf_wi_pdf_iso(V3 wo, V3 wh, V3 n, float alpha) {
    float wh_pdf = G1_iso(wo, n, alpha) * D_iso(wh, n, alpha) * std::max(0.f, dot(wo, wh)) / dot(wo, n);
    return wh_pdf / (4 * dot(wh, wo));
}
static RGB schlick_fresnel(RGB r0, float cos_i) {                                    // scattering.cpp:10
    float k = 1.f - std::fabs(cos_i);
    float k5 = (k * k) * (k * k) * k;
    return r0 + (splat(1.f) - r0) * k5;
}

// Plastic_BRDF::evaluate (bsdf.cpp:214-233) / Pbrt3_Plastic_BRDF::evaluate (bsdf_pbrt.cpp:264-291)
static RGB plastic_eval(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    V3 wh = normalized(wo + wi);
    float cos_i = dot(wi, wh);
    RGB F;
    if (m.pbrt3) {
        float rel = 1.f / 1.5f;
        V3 original = sc.flipped ? -sc.n : sc.n;
        if (dot(original, wi) < 0.f) rel = 1.5f / 1.f;
        F = splat(dielectric_fresnel(cos_i, rel));
    } else {
        F = schlick_fresnel(splat(0.04f), cos_i);
    }
    float G = G_iso(wi, wo, sc.n, m.alpha_x);
    float D = D_iso(wh, sc.n, m.alpha_x);
    float won = dot(wo, sc.n), win = dot(wi, sc.n);
    RGB base = F * ((G * D) / (4.f * won * win));
    RGB spec = m.r0 * base;
    RGB diffuse = m.refl * Pi_Inv;
    return diffuse + spec;
}
static float plastic_pdf(const Material& m, const Shading& sc, V3 wo, V3 wi) {       // bsdf.cpp:253-262
    float diffuse_pdf = dot(sc.n, wi) / Pi;
    V3 wh = normalized(wo + wi);
    float spec_pdf = vndf_wi_pdf_iso(wo, wh, sc.n, m.alpha_x);
    return 0.5f * (diffuse_pdf + spec_pdf);
}

// Ashikhmin_Shirley_Phong_BRDF::evaluate / pdf (bsdf.cpp:440-495); Material: refl = diffuse
// reflectance, k = Fresnel r0 per channel.
static RGB coated_eval(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    float cos_i = dot(wil, wh);
    RGB F = schlick_fresnel(m.k, cos_i);
    float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    RGB spec = F * (D / (4.f * cos_i * std::max(wol.z, wil.z)));
    auto pow5 = [](float v) { return (v * v) * (v * v) * v; };
    RGB diffuse = (m.refl * (splat(1.f) - m.k)) *
                  (28.f / (23.f * Pi) * (1.f - pow5(1.f - 0.5f * wil.z)) * (1.f - pow5(1.f - 0.5f * wol.z)));
    return diffuse + spec;
}
static float coated_pdf(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    float diffuse_pdf = dot(sc.n, wi) / Pi;
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    float spec_pdf = vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
    return 0.5f * (diffuse_pdf + spec_pdf);
}

// Pbrt3_Uber_BRDF (bsdf_pbrt.cpp:14-91): Kd/pi and an anisotropic GGX lobe with the Fresnel of a
// dielectric (ior inverted when leaving), both scaled by the opacity; half of the samples cosine, half VNDF.
static float uber_ior(const Material& m, const Shading& sc) { return sc.flipped ? 1.f / m.ior : m.ior; }
static RGB uber_eval(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    RGB diffuse = (Pi_Inv * m.refl) * m.opacity;
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    float cos_i = dot(wil, wh);
    float F = dielectric_fresnel(cos_i, uber_ior(m, sc));
    float G = G_aniso(wil, wol, m.alpha_x, m.alpha_y);
    float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    float f = (F * G * D) / (4.f * wol.z * wil.z);                                   // scattering.cpp:84
    RGB specular = (m.ks * m.opacity) * f;
    return diffuse + specular;
}
static float uber_pdf(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    float diffuse_pdf = dot(sc.n, wi) / Pi;
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    float spec_pdf = vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
    return 0.5f * (diffuse_pdf + spec_pdf);
}

// Rough_Glass_BSDF (bsdf.cpp:268-420) with scattering.cpp:65-150 and math.h:115-127.  Material:
// refl = Kr, trans = Kt, ior, alpha_x = GGX alpha; enter/exit from the original normal orientation.
static void rough_glass_etas(const Material& m, const Shading& sc, float* eta_o, float* eta_i) {
    bool enter = !sc.flipped;
    *eta_o = enter ? 1.f : m.ior;
    *eta_i = enter ? m.ior : 1.f;
}
static float max_component(RGB c) { return std::max(c.r, std::max(c.g, c.b)); }
static V3 refraction_half_direction(float eta_o, V3 wo, float eta_i, V3 wi, V3 n) {
    V3 wh = -normalized(eta_o * wo + eta_i * wi);
    if (dot(wh, n) < 0.f) wh = -wh;
    return wh;
}
static bool refract(V3 w, V3 n, float eta, V3* out) {
    float ci = dot(w, n);
    float st2 = eta * eta * std::max(0.f, 1.f - ci * ci);
    if (st2 >= 1.f) return false;
    float ct = std::sqrt(1.f - st2);
    *out = (-eta) * w + (eta * ci - ct) * n;
    return true;
}
static float glass_reflection_probability(const Material& m, float fresnel) {
    float max_r = max_component(m.refl), max_t = max_component(m.trans);
    float ratio = max_r / (max_r + max_t);
    float r = fresnel * ratio;
    float t = (1.f - fresnel) * (1 - ratio);
    return ((r + t) == 0) ? 0.f : r / (r + t);
}
static RGB rough_glass_eval(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    float eta_o, eta_i;
    rough_glass_etas(m, sc, &eta_o, &eta_i);
    V3 n = sc.n;
    float a = m.alpha_x;
    if (dot(wo, n) * dot(wi, n) > 0.f) {
        V3 wh = normalized(wo + wi);
        float F = dielectric_fresnel(dot(wi, wh), eta_i / eta_o);
        float G = G_iso(wi, wo, n, a), D = D_iso(wh, n, a);
        float base = (F * G * D) / (4.f * dot(wo, n) * dot(wi, n));
        return m.refl * base;
    }
    V3 wh = refraction_half_direction(eta_o, wo, eta_i, wi, n);
    float wow = dot(wo, wh), wiw = dot(wi, wh);
    if (wow * wiw > 0.f) return RGB{};
    float F = dielectric_fresnel(wiw, eta_o / eta_i);
    if (F == 1.f) return RGB{};
    float G = G_iso(wi, wo, n, a), D = D_iso(wh, n, a);
    float k = std::fabs((wow * wiw) / (dot(wo, n) * dot(wi, n)));
    float k2 = eta_o * wow + eta_i * wiw;
    float base = (eta_o * eta_o * k * G * D * (1.f - F)) / (k2 * k2);
    return m.trans * base;
}
static float rough_glass_pdf(const Material& m, const Shading& sc, V3 wo, V3 wi) {
    float eta_o, eta_i;
    rough_glass_etas(m, sc, &eta_o, &eta_i);
    V3 n = sc.n;
    float a = m.alpha_x;
    if (dot(wo, n) * dot(wi, n) > 0.f) {
        V3 wh = normalized(wo + wi);
        float rpdf = vndf_wi_pdf_iso(wo, wh, n, a);
        return rpdf * glass_reflection_probability(m, dielectric_fresnel(dot(wi, wh), eta_i / eta_o));
    }
    V3 wh = refraction_half_direction(eta_o, wo, eta_i, wi, n);
    if (dot(wo, wh) * dot(wi, wh) > 0.f) return 0.f;
    float rp = glass_reflection_probability(m, dielectric_fresnel(dot(wo, wh), eta_i / eta_o));
    float wh_pdf = G1_iso(wo, n, a) * D_iso(wh, n, a) * std::max(0.f, dot(wo, wh)) / dot(wo, n);
    float denom = eta_o * dot(wo, wh) + eta_i * dot(wi, wh);
    float dwh_dwi = eta_i * eta_i * std::fabs(dot(wi, wh)) / (denom * denom);
    return wh_pdf * dwh_dwi * (1.f - rp);
}

// material_has_bsdf (material.h): the delta layers (perfect reflector, smooth glass) have none
static bool has_bsdf(const Material& m) {
    return !(m.type == YAR_MATERIAL_MIRROR || (m.type == YAR_MATERIAL_GLASS && m.roughness <= 0.f));
}
static bool has_transmission(const Scene& s, const Material& m) {                 // BSDF::transmission_scattering
    if (m.type == YAR_MATERIAL_MIX) {
        const Material& b1 = s.materials[m.sub[0]];
        const Material& b2 = s.materials[m.sub[1]];
        return (has_bsdf(b1) && has_transmission(s, b1)) || (has_bsdf(b2) && has_transmission(s, b2));
    }
    return m.type == YAR_MATERIAL_GLASS;                                             // Rough_Glass_BSDF
}
static bool sub_scatters(const Scene& s, const Material& m, float n_dot_wi) {
    return n_dot_wi > 0.f || (n_dot_wi < 0.f && has_transmission(s, m));
}
static RGB bsdf_eval_m(const Scene& s, const Material& m, const Shading& sc, V3 wo, V3 wi);
static float bsdf_pdf_m(const Scene& s, const Material& m, const Shading& sc, V3 wo, V3 wi);
static RGB bsdf_sample_m(const Scene& s, const Material& m, const Shading& sc, V2 u, float u_scat, V3 wo, V3* wi, float* pdf);
// Mix_BSDF: mix_coeff = White - amount weights the second BSDF, p_coeff = 1 - luminance(mix_coeff) is
// the probability of sampling it; each BSDF counts only where it can scatter towards wi
static RGB mix_coeff(const Material& m) { return splat(1.f) - m.amount; }
// create_mix_bsdf (bsdf.cpp:599-629): with a delta sub-material the BSDF is the other one scaled by
// its weight, luminance(amount) for the first, 1 - luminance(amount) for the second (unclamped)
static bool mix_scaled(const Scene& s, const Material& m, const Material** b, float* scale) {
    const Material& b1 = s.materials[m.sub[0]];
    const Material& b2 = s.materials[m.sub[1]];
    if (has_bsdf(b1) && has_bsdf(b2)) return false;
    const float mix_amount = luminance(m.amount);
    if (!has_bsdf(b1)) { *b = &b2; *scale = 1.f - mix_amount; }
    else { *b = &b1; *scale = mix_amount; }
    return true;
}
static RGB mix_eval(const Scene& s, const Material& m, const Shading& sc, V3 wo, V3 wi) {
    const Material* bs; float scale;
    if (mix_scaled(s, m, &bs, &scale)) return scale * bsdf_eval_m(s, *bs, sc, wo, wi);   // Scaled_BSDF
    const Material& b1 = s.materials[m.sub[0]];
    const Material& b2 = s.materials[m.sub[1]];
    RGB f1, f2;
    float ndw = dot(sc.n, wi);
    if (sub_scatters(s, b1, ndw)) f1 = bsdf_eval_m(s, b1, sc, wo, wi);
    if (sub_scatters(s, b2, ndw)) f2 = bsdf_eval_m(s, b2, sc, wo, wi);
    const RGB mc = mix_coeff(m);
    return (splat(1.f) - mc) * f1 + mc * f2;
}
static float mix_pdf(const Scene& s, const Material& m, const Shading& sc, V3 wo, V3 wi) {
    const Material* bs; float scale;
    if (mix_scaled(s, m, &bs, &scale)) return bsdf_pdf_m(s, *bs, sc, wo, wi);
    const Material& b1 = s.materials[m.sub[0]];
    const Material& b2 = s.materials[m.sub[1]];
    const float p = 1.f - luminance(mix_coeff(m));
    float ndw = dot(sc.n, wi);
    float pdf1 = 0.f, pdf2 = 0.f;
    if (sub_scatters(s, b1, ndw)) pdf1 = bsdf_pdf_m(s, b1, sc, wo, wi);
    if (sub_scatters(s, b2, ndw)) pdf2 = bsdf_pdf_m(s, b2, sc, wo, wi);
    return (1.f - p) * pdf1 + p * pdf2;
}
static RGB mix_sample(const Scene& s, const Material& m, const Shading& sc, V2 u, float u_scat, V3 wo, V3* wi, float* pdf) {
    const Material* bs; float scale;
    if (mix_scaled(s, m, &bs, &scale)) return scale * bsdf_sample_m(s, *bs, sc, u, u_scat, wo, wi, pdf);
    const Material& b1 = s.materials[m.sub[0]];
    const Material& b2 = s.materials[m.sub[1]];
    const RGB mc = mix_coeff(m);
    const float p = 1.f - luminance(mc);
    RGB f1, f2;
    float pdf1 = 0.f, pdf2 = 0.f;
    if (u_scat < p) {
        u_scat = u_scat / p;
        f2 = bsdf_sample_m(s, b2, sc, u, u_scat, wo, wi, &pdf2);
        if (sub_scatters(s, b1, dot(sc.n, *wi))) {
            pdf1 = bsdf_pdf_m(s, b1, sc, wo, *wi);
            f1 = bsdf_eval_m(s, b1, sc, wo, *wi);
        }
    } else {
        u_scat = (u_scat - p) / (1.f - p);
        f1 = bsdf_sample_m(s, b1, sc, u, u_scat, wo, wi, &pdf1);
        if (sub_scatters(s, b2, dot(sc.n, *wi))) {
            pdf2 = bsdf_pdf_m(s, b2, sc, wo, *wi);
            f2 = bsdf_eval_m(s, b2, sc, wo, *wi);
        }
    }
    *pdf = (1.f - p) * pdf1 + p * pdf2;
    return (splat(1.f) - mc) * f1 + mc * f2;
}

static RGB bsdf_eval(const Scene& s, const Shading& sc, V3 wo, V3 wi) {
    return bsdf_eval_m(s, s.materials[sc.material], sc, wo, wi);
}
static float bsdf_pdf(const Scene& s, const Shading& sc, V3 wo, V3 wi) {
    return bsdf_pdf_m(s, s.materials[sc.material], sc, wo, wi);
}
static RGB bsdf_sample(const Scene& s, const Shading& sc, V2 u, float u_scat, V3 wo, V3* wi, float* pdf) {
    return bsdf_sample_m(s, s.materials[sc.material], sc, u, u_scat, wo, wi, pdf);
}

static RGB bsdf_eval_m(const Scene& s, const Material& m, const Shading& sc, V3 wo, V3 wi) {
    if (m.type == YAR_MATERIAL_MIX) return mix_eval(s, m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_GLASS) return rough_glass_eval(m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_DIFFUSE) return Pi_Inv * m.refl;                      // bsdf.cpp:68
    if (m.type == YAR_MATERIAL_PLASTIC) return plastic_eval(m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_COATED_DIFFUSE) return coated_eval(m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_PBRT3_UBER) return uber_eval(m, sc, wo, wi);
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);                               // bsdf.cpp:160
    V3 wh = normalized(wol + wil);
    float ci = dot(wil, wh);
    RGB F = conductor_fresnel(ci, m.eta_i, m.eta, m.k);
    float G = G_aniso(wil, wol, m.alpha_x, m.alpha_y);
    float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    return F * ((G * D) / (4.f * wol.z * wil.z));
}
static float bsdf_pdf_m(const Scene& s, const Material& m, const Shading& sc, V3 wo, V3 wi) {
    if (m.type == YAR_MATERIAL_MIX) return mix_pdf(s, m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_DIFFUSE) return dot(sc.n, wi) / Pi;                   // bsdf.cpp:80
    if (m.type == YAR_MATERIAL_PLASTIC) return plastic_pdf(m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_COATED_DIFFUSE) return coated_pdf(m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_GLASS) return rough_glass_pdf(m, sc, wo, wi);
    if (m.type == YAR_MATERIAL_PBRT3_UBER) return uber_pdf(m, sc, wo, wi);
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);                               // bsdf.cpp:190
    V3 wh = normalized(wol + wil);
    return vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
}
static RGB bsdf_sample_m(const Scene& s, const Material& m, const Shading& sc, V2 u, float u_scat, V3 wo, V3* wi, float* pdf) {
    if (m.type == YAR_MATERIAL_MIX) return mix_sample(s, m, sc, u, u_scat, wo, wi, pdf);
    if (m.type == YAR_MATERIAL_DIFFUSE) {                                           // bsdf.cpp:71
        *wi = to_world(sc, sample_cosine(u));
        *pdf = bsdf_pdf_m(s, m, sc, wo, *wi);
        return bsdf_eval_m(s, m, sc, wo, *wi);
    }
    if (m.type == YAR_MATERIAL_PLASTIC) {                                           // bsdf.cpp:235-251
        if (u_scat < 0.5f) {
            *wi = to_world(sc, sample_cosine(u));
        } else {
            V3 wh = to_world(sc, vndf_sample(u, to_local(sc, wo), m.alpha_x, m.alpha_x));   // bsdf.cpp:41-48
            *wi = reflect(wo, wh);
        }
        if (dot(sc.n, *wi) <= 0.f) return RGB{};
        *pdf = plastic_pdf(m, sc, wo, *wi);
        return plastic_eval(m, sc, wo, *wi);
    }
    if (m.type == YAR_MATERIAL_GLASS) {                                             // bsdf.cpp:339-376
        float eta_o, eta_i;
        rough_glass_etas(m, sc, &eta_o, &eta_i);
        V3 wh = to_world(sc, vndf_sample(u, to_local(sc, wo), m.alpha_x, m.alpha_x));
        V3 rwi = reflect(wo, wh);
        if (dot(rwi, sc.n) <= 0.f) return RGB{};
        float rp = glass_reflection_probability(m, dielectric_fresnel(dot(wo, wh), eta_i / eta_o));
        if (u_scat < rp) {
            *wi = rwi;
        } else {
            if (!refract(wo, wh, eta_o / eta_i, wi)) return RGB{};
            if (dot(*wi, sc.n) >= 0.f) return RGB{};
        }
        *pdf = rough_glass_pdf(m, sc, wo, *wi);
        if (*pdf == 0.f) return RGB{};
        return rough_glass_eval(m, sc, wo, *wi);
    }
    if (m.type == YAR_MATERIAL_COATED_DIFFUSE) {                                    // bsdf.cpp:462-479
        if (u_scat < 0.5f) {
            *wi = to_world(sc, sample_cosine(u));
        } else {
            V3 wol = to_local(sc, wo);
            V3 wh = vndf_sample(u, wol, m.alpha_x, m.alpha_y);
            *wi = to_world(sc, reflect(wol, wh));
        }
        if (dot(sc.n, *wi) <= 0.f) return RGB{};
        *pdf = coated_pdf(m, sc, wo, *wi);
        return coated_eval(m, sc, wo, *wi);
    }
    if (m.type == YAR_MATERIAL_PBRT3_UBER) {                                        // bsdf_pbrt.cpp:56-74
        if (u_scat < 0.5f) {
            *wi = to_world(sc, sample_cosine(u));
        } else {
            V3 wol = to_local(sc, wo);
            V3 wh = vndf_sample(u, wol, m.alpha_x, m.alpha_y);
            *wi = to_world(sc, reflect(wol, wh));
        }
        if (dot(sc.n, *wi) <= 0.f) return RGB{};
        *pdf = uber_pdf(m, sc, wo, *wi);
        return uber_eval(m, sc, wo, *wi);
    }
    V3 wol = to_local(sc, wo);                                                       // bsdf.cpp:176
    V3 wh = vndf_sample(u, wol, m.alpha_x, m.alpha_y);
    V3 wil = reflect(wol, wh);
    if (wil.z <= 0.f) return RGB{};
    *wi = to_world(sc, wil);
    *pdf = vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
    return bsdf_eval_m(s, m, sc, wo, *wi);
}

// ---------------------------------------------------------------------------------------------
// Direct lighting (src/ref/direct_lighting.cpp:18-194, 386-413, 522-578)
// ---------------------------------------------------------------------------------------------
static inline float power_heuristic(float a, float b) {
    if (a == Inf) return 1.f;
    return a * a / (a * a + b * b);
}

struct PathCounters { int64_t closest = 0, any = 0; Counters trav; };

// get_emitted_radiance (direct_lighting.cpp:388-410): rectangles and spheres emit on both sides,
// triangle-mesh lights only where the original shading normal faces the viewer.  Area lights are
// numbered rectangles, then spheres, then meshes.
static RGB area_light_emission(const Scene& s, const Shading& sc) {
    const int n_rects = (int)s.rects.size(), n_spheres = (int)s.spheres.size();
    if (sc.light < n_rects) return s.rects[sc.light].L;
    if (sc.light < n_rects + n_spheres) return s.spheres[sc.light - n_rects].L;
    if (sc.flipped) return RGB{};
    return s.mesh_lights[sc.light - n_rects - n_spheres].L;
}

static RGB light_point(const Scene& s, const Shading& sc, const PointLight& L, PathCounters& pc, Counters* tc) {
    V3 position = sc.origin_point(L.p);
    V3 lv = L.p - position;
    float dist = length(lv);
    V3 ldir = vdiv(lv, dist);
    float ndl = dot(sc.n, ldir);
    if (ndl <= 0.f) return RGB{};
    pc.any++;
    if (intersect_any(s, Ray{position, ldir}, dist * (1.f - 1e-5f), tc)) return RGB{};
    RGB f = bsdf_eval(s, sc, sc.wo, ldir);
    return (L.I * f) * (ndl / (dist * dist));
}

static RGB light_rect(const Scene& s, const Shading& sc, int li, V2 ul, V2 ub, float us, PathCounters& pc, Counters* tc) {
    const RectLight& L = s.rects[li];
    V3 ln = column(L.l2w, 2);
    RGB out;
    {
        V3 local{L.size.x * (ul.x - 0.5f), L.size.y * (ul.y - 0.5f), 0.0f};
        V3 lp = xform_point(L.l2w, local);
        V3 position = sc.origin_point(lp);
        V3 lv = lp - position;
        float d = length(lv);
        V3 wi = vdiv(lv, d);
        float lndw = dot(ln, -wi);
        if (lndw > 1e-4f) {
            float ndw = dot(sc.n, wi);
            if (sub_scatters(s, s.materials[sc.material], ndw)) {   // scattering_possible (direct_lighting.cpp:133-135)
                RGB f = bsdf_eval(s, sc, sc.wo, wi);
                if (!is_black(f)) {
                    pc.any++;
                    if (!intersect_any(s, Ray{position, wi}, d * (1.f - 1e-5f), tc)) {
                        float lpdf = (d * d) / (L.size.x * L.size.y * lndw);
                        float bpdf = bsdf_pdf(s, sc, sc.wo, wi);
                        float w = power_heuristic(lpdf, bpdf);
                        out = out + (L.L * f) * (w * std::fabs(ndw) / lpdf);
                    }
                }
            }
        }
    }
    {
        V3 wi; float bpdf;
        RGB f = bsdf_sample(s, sc, ub, us, sc.wo, &wi, &bpdf);
        if (!is_black(f)) {
            float lndw = dot(ln, -wi);
            if (lndw > 1e-4f) {
                V3 position = sc.origin_dir(wi);
                Hit h{Inf, -1, 0, {}};
                pc.closest++;
                bool found = intersect(s, Ray{position, wi}, h, tc);
                if (found && s.objects[h.obj].light == li) {
                    const Mesh& m = s.meshes[s.objects[h.obj].mesh];
                    V3 p[3]; m.pos(h.tri, p);
                    V3 ph = h.b.x * p[0] + h.b.y * p[1] + h.b.z * p[2];   // mesh space, as in :181
                    float d = length(ph - position);
                    float lpdf = (d * d) / (L.size.x * L.size.y * lndw);
                    float w = power_heuristic(bpdf, lpdf);
                    float ndw = dot(sc.n, wi);
                    out = out + (L.L * f) * (w * std::fabs(ndw) / bpdf);
                }
            }
        }
    }
    (void)us;
    return out;
}

static bool scattering_possible(const Scene& s, const Shading& sc, float n_dot_wi) {   // direct_lighting.cpp:133-135
    return sub_scatters(s, s.materials[sc.material], n_dot_wi);
}

static RGB light_spot(const Scene& s, const Shading& sc, const SpotLight& L, PathCounters& pc, Counters* tc) {
    V3 position = sc.origin_point(L.p);                                      // direct_lighting.cpp:49-89
    V3 lv = L.p - position;
    float dist = length(lv);
    V3 wi = vdiv(lv, dist);
    float cone_cos = std::cos(L.cone_angle);
    float wi_cos = dot(-wi, L.dir);
    if (wi_cos < cone_cos) return RGB{};
    float penumbra = 1.f;
    float penumbra_cos = std::cos(std::max(0.f, L.cone_angle - L.penumbra_angle));
    if (wi_cos < penumbra_cos) {
        float k = (wi_cos - cone_cos) / (penumbra_cos - cone_cos);
        penumbra = (k * k) * (k * k);
    }
    float ndw = dot(sc.n, wi);
    if (!scattering_possible(s, sc, ndw)) return RGB{};
    RGB f = bsdf_eval(s, sc, sc.wo, wi);
    if (is_black(f)) return RGB{};
    pc.any++;
    if (intersect_any(s, Ray{position, wi}, dist * (1.f - 1e-5f), tc)) return RGB{};
    return (L.I * f) * (penumbra * std::fabs(ndw) / (dist * dist));
}

static RGB light_directional(const Scene& s, const Shading& sc, const DirectionalLight& L, PathCounters& pc, Counters* tc) {
    float ndl = dot(sc.n, L.dir);                                            // direct_lighting.cpp:91-107
    if (ndl <= 0.f) return RGB{};
    V3 position = sc.origin_dir(L.dir);
    pc.any++;
    if (intersect_any(s, Ray{position, L.dir}, Inf, tc)) return RGB{};
    RGB f = bsdf_eval(s, sc, sc.wo, L.dir);
    return (L.E * f) * ndl;
}

// Distribution_1D::sample (sampling.cpp:98-147, 168-172): the sample in [0,1) and the remapped u
static float sample_cdf(const std::vector<float>& cdf, float u, float* remapped) {
    const int n = (int)cdf.size();
    const float interval_length = 1.f / n;
    const int k = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    const float cdf_a = (k == 0) ? 0.f : cdf[k - 1];
    const float cdf_b = cdf[k];
    const float fraction = (u - cdf_a) / (cdf_b - cdf_a);
    const float x = (float(k) + fraction) * interval_length;
    *remapped = std::min(fraction, One_Minus_Epsilon);
    return std::min(x, One_Minus_Epsilon);
}

// Emitter normal of a mesh light at (triangle, barycentrics): the interpolated vertex normal when the
// mesh has normals (not normalised), else the geometric normal (light_sampling.cpp:137-142, 168-173)
static V3 mesh_light_normal(const Mesh& m, uint32_t tri, V3 b) {
    if (!m.n.empty()) {
        V3 nv[3];
        for (int k = 0; k < 3; k++) { const float* q = &m.n[3 * (size_t)m.idx[3 * (size_t)tri + k]]; nv[k] = {q[0], q[1], q[2]}; }
        return b.x * nv[0] + b.y * nv[1] + b.z * nv[2];
    }
    V3 p[3]; m.pos(tri, p);
    V3 ng = normalized(cross(p[1] - p[0], p[2] - p[0]));
    return m.reverse ? -ng : ng;
}

static RGB light_mesh(const Scene& s, const Shading& sc, int li, V2 ul, V2 ub, float us, PathCounters& pc, Counters* tc) {
    const MeshLight& ML = s.mesh_lights[li];
    const Mesh& m = s.meshes[ML.mesh];
    const int ntri = (int)(m.idx.size() / 3);
    RGB out;
    {   // Diffuse_Triangle_Mesh_Light_Sampler::sample (light_sampling.cpp:126-161)
        float u0;
        float smp = sample_cdf(ML.cdf, ul.x, &u0);
        uint32_t tri = std::min(uint32_t(smp * ntri), (uint32_t)ntri - 1);
        float a = std::sqrt(u0);                                             // sampling.cpp:69-77
        V3 b;
        b.x = 1 - a;
        b.y = ul.y * a;
        b.z = 1.f - b.x - b.y;
        V3 p[3]; m.pos(tri, p);
        V3 lp = b.x * p[0] + b.y * p[1] + b.z * p[2];
        V3 ln = mesh_light_normal(m, tri, b);
        V3 lv0 = lp - sc.pos;
        float dsq = dot(lv0, lv0);
        V3 w0 = vdiv(lv0, std::sqrt(dsq));
        float lndw0 = dot(ln, -w0);
        float lpdf = lndw0 <= 0.f ? 0.f : dsq / (lndw0 * ML.area);
        if (lpdf != 0.f) {                                                   // direct_lighting.cpp:272-297
            V3 position = sc.origin_point(lp);
            V3 lv = lp - position;
            float d = length(lv);
            V3 wi = vdiv(lv, d);
            float ndw = dot(sc.n, wi);
            if (scattering_possible(s, sc, ndw)) {
                RGB f = bsdf_eval(s, sc, sc.wo, wi);
                if (!is_black(f)) {
                    pc.any++;
                    if (!intersect_any(s, Ray{position, wi}, d * (1.f - 1e-5f), tc)) {
                        float bpdf = bsdf_pdf(s, sc, sc.wo, wi);
                        float w = power_heuristic(lpdf, bpdf);
                        out = out + (ML.L * f) * (w * std::fabs(ndw) / lpdf);
                    }
                }
            }
        }
    }
    {   // BSDF sampling half (direct_lighting.cpp:299-321, light_sampling.cpp:163-184)
        V3 wi; float bpdf;
        RGB f = bsdf_sample(s, sc, ub, us, sc.wo, &wi, &bpdf);
        if (!is_black(f)) {
            V3 position = sc.origin_dir(wi);
            Hit h{Inf, -1, 0, {}};
            pc.closest++;
            bool found = intersect(s, Ray{position, wi}, h, tc);
            if (found && s.objects[h.obj].light == (int)(s.rects.size() + s.spheres.size()) + li) {
                V3 ln = mesh_light_normal(m, h.tri, h.b);
                float lndw = dot(ln, -wi);
                float lpdf = lndw <= 0.f ? 0.f : (h.t * h.t) / (lndw * ML.area);
                float w = power_heuristic(bpdf, lpdf);
                out = out + (ML.L * f) * (w * std::fabs(dot(sc.n, wi)) / bpdf);
            }
        }
    }
    return out;
}

// Diffuse_Sphere_Light_Sampler (light_sampling.cpp:76-121): uniform directions in the cone the
// sphere subtends from the shading point, mapped to the near intersection with the sphere.
struct SphereSampler {
    V3 ax[3];
    float d_center, cos_theta_max, cone_pdf;
    SphereSampler(const SphereLight& L, V3 shading_pos) {
        ax[2] = shading_pos - L.p;
        d_center = length(ax[2]);
        ax[2] = vdiv(ax[2], d_center);
        coordinate_system(ax[2], &ax[0], &ax[1]);
        const float sin_theta_max = L.radius / d_center;
        cos_theta_max = std::sqrt(std::max(0.f, 1.f - sin_theta_max * sin_theta_max));
        cone_pdf = 1.f / (2.f * Pi * (1.f - cos_theta_max));
    }
    V3 sample(const SphereLight& L, V2 u) const {
        const float r2 = L.radius * L.radius, d2 = d_center * d_center;
        const float cos_theta = (1.f - u.x) + u.x * cos_theta_max;
        const float phi = 2.f * Pi * u.y;
        const float sin_theta2 = std::max(0.f, 1.f - cos_theta * cos_theta);
        const float d_sample = d_center * cos_theta - std::sqrt(std::max(0.f, r2 - d2 * sin_theta2));
        const float cos_alpha = (d2 + r2 - d_sample * d_sample) / (2.f * d_center * L.radius);   // law of cosines
        const float sin_alpha = std::sqrt(std::max(0.f, 1.f - cos_alpha * cos_alpha));
        const V3 dir = ((sin_alpha * std::cos(phi)) * ax[0] + (sin_alpha * std::sin(phi)) * ax[1]) + cos_alpha * ax[2];
        return L.radius * dir + L.p;
    }
    bool in_cone(V3 wi) const { return dot(-ax[2], wi) >= cos_theta_max; }
};

static RGB light_sphere(const Scene& s, const Shading& sc, int li, V2 ul, V2 ub, float us, PathCounters& pc, Counters* tc) {
    const SphereLight& L = s.spheres[li];                                    // direct_lighting.cpp:196-260
    const SphereSampler smp(L, sc.pos);
    RGB out;
    {
        const V3 lp = smp.sample(L, ul);
        const V3 position = sc.origin_point(lp);
        const V3 lv = lp - position;
        const float d = length(lv);
        const V3 wi = vdiv(lv, d);
        const float ndw = dot(sc.n, wi);
        if (scattering_possible(s, sc, ndw)) {
            RGB f = bsdf_eval(s, sc, sc.wo, wi);
            if (!is_black(f)) {
                pc.any++;
                if (!intersect_any(s, Ray{position, wi}, d * (1.f - 1e-5f), tc)) {
                    const float lpdf = smp.cone_pdf;
                    const float bpdf = bsdf_pdf(s, sc, sc.wo, wi);
                    const float w = power_heuristic(lpdf, bpdf);
                    out = out + (L.L * f) * (w * std::fabs(ndw) / lpdf);
                }
            }
        }
    }
    {
        V3 wi; float bpdf;
        RGB f = bsdf_sample(s, sc, ub, us, sc.wo, &wi, &bpdf);
        if (!is_black(f) && smp.in_cone(wi)) {
            const V3 position = sc.origin_dir(wi);
            Hit h{Inf, -1, 0, {}};
            pc.closest++;
            const bool found = intersect(s, Ray{position, wi}, h, tc);
            if (found && s.objects[h.obj].light == (int)s.rects.size() + li) {
                const float lpdf = smp.cone_pdf;
                const float w = power_heuristic(bpdf, lpdf);
                out = out + (L.L * f) * (w * std::fabs(dot(sc.n, wi)) / bpdf);
            }
        }
    }
    return out;
}

// Environment_Light_Sampler over a constant (one-texel) map (light_sampling.cpp:11-73,
// sampling.cpp:185-300): the one-cell Distribution_2D samples uv = (u1, u0) with pdf_uv 1.
static void env_uv(V3 d, float* u, float* v) {                                       // get_uv_from_direction
    float phi = std::atan2(d.y, d.x);
    phi = phi < 0 ? phi + Pi2 : phi;
    float theta = std::acos(std::clamp(d.z, -1.f, 1.f));
    *u = std::clamp(phi * Pi2_Inv, 0.f, One_Minus_Epsilon);
    *v = std::min(theta * Pi_Inv, One_Minus_Epsilon);
}
// get_filtered_radiance_for_direction: the bilinear weights of the white texel, summed in float
static RGB env_filtered(const EnvLight& E, V3 world_dir) {
    float u, v;
    env_uv(xform_vector(E.w2l, world_dir), &u, &v);
    float x = u * 1.f - 0.5f, y = v * 1.f - 0.5f;                                   // image_texture.cpp:441-480
    float wx = x - std::floor(x), wy = y - std::floor(y);
    float w = (((1 - wy) * (1 - wx) + (1 - wy) * wx) + wy * (1 - wx)) + wy * wx;
    return splat(w) * E.L;
}
static float env_pdf(const EnvLight& E, V3 world_dir) {
    V3 d = xform_vector(E.w2l, world_dir);
    float sin_theta = std::sqrt(std::max(0.f, 1.f - d.z * d.z));
    return 1.f / (2 * Pi * Pi * sin_theta);
}

static RGB light_env(const Scene& s, const Shading& sc, V2 ul, V2 ub, float us, PathCounters& pc, Counters* tc) {
    const EnvLight& E = s.env;                                               // direct_lighting.cpp:324-384
    RGB out;
    {
        const float phi = ul.y * Pi2, theta = ul.x * Pi;
        const float sin_theta = std::sin(theta);
        const float lpdf = 1.f / (2 * Pi * Pi * sin_theta);
        V3 wi{sin_theta * std::cos(phi), sin_theta * std::sin(phi), std::cos(theta)};
        wi = xform_vector(E.l2w, wi);
        const float ndw = dot(sc.n, wi);
        if (scattering_possible(s, sc, ndw)) {
            RGB f = bsdf_eval(s, sc, sc.wo, wi);
            if (!is_black(f)) {
                const V3 position = sc.origin_dir(wi);
                pc.any++;
                if (!intersect_any(s, Ray{position, wi}, Inf, tc)) {
                    const float bpdf = bsdf_pdf(s, sc, sc.wo, wi);
                    const float w = power_heuristic(lpdf, bpdf);
                    out = out + (E.L * f) * (w * std::fabs(ndw) / lpdf);
                }
            }
        }
    }
    {
        V3 wi; float bpdf;
        RGB f = bsdf_sample(s, sc, ub, us, sc.wo, &wi, &bpdf);
        if (!is_black(f) && !is_black(E.L)) {
            const V3 position = sc.origin_dir(wi);
            pc.any++;
            if (!intersect_any(s, Ray{position, wi}, Inf, tc)) {
                const float lpdf = env_pdf(E, wi);
                const float w = power_heuristic(bpdf, lpdf);
                out = out + (E.L * f) * (w * std::fabs(dot(sc.n, wi)) / bpdf);
            }
        }
    }
    return out;
}

// estimate_direct_lighting_from_single_sample (direct_lighting.cpp:522-578): lights in the order
// point, spot, directional, rectangular, sphere, triangle mesh, environment
static RGB direct_single(const Scene& s, const Shading& sc, float u_sel, V2 ul, V2 ub, float us, PathCounters& pc, Counters* tc) {
    int li = int(u_sel * s.total_lights);
    const float n = (float)s.total_lights;
    if (li < (int)s.points.size()) return n * light_point(s, sc, s.points[li], pc, tc);
    li -= (int)s.points.size();
    if (li < (int)s.spots.size()) return n * light_spot(s, sc, s.spots[li], pc, tc);
    li -= (int)s.spots.size();
    if (li < (int)s.dirs.size()) return n * light_directional(s, sc, s.dirs[li], pc, tc);
    li -= (int)s.dirs.size();
    if (li < (int)s.rects.size()) return n * light_rect(s, sc, li, ul, ub, us, pc, tc);
    li -= (int)s.rects.size();
    if (li < (int)s.spheres.size()) return n * light_sphere(s, sc, li, ul, ub, us, pc, tc);
    li -= (int)s.spheres.size();
    if (li < (int)s.mesh_lights.size()) return n * light_mesh(s, sc, li, ul, ub, us, pc, tc);
    return n * light_env(s, sc, ul, ub, us, pc, tc);
}

// Smooth glass: get_glass_info (delta_scattering.cpp:53-100), enter/exit judged by the original
// shading-normal orientation (no nested-dielectric tracking for pbrt scenes).  Fresnel-chosen
// reflection (always under total internal reflection) or refraction (refract, math.h:115) with the
// transmitted radiance scaled by eta^2, eta = eta_incident / eta_transmitted.  Returns true on reflection.
static bool glass_scatter(const Shading& sc, const Material& m, float u_scat, RGB* att, V3* dir) {
    bool enter = !sc.flipped;
    float etaT_over_etaI = enter ? m.ior : 1.f / m.ior;
    float cos_i = dot(sc.n, sc.wo);
    float F = dielectric_fresnel(cos_i, etaT_over_etaI);
    if (u_scat < F) {
        *att = m.refl;
        *dir = reflect(sc.wo, sc.n);
        return true;
    }
    float eta = 1.f / etaT_over_etaI;
    *att = m.trans;
    *att = *att * (eta * eta);
    float ci = dot(sc.wo, sc.n);
    float st2 = eta * eta * std::max(0.f, 1.f - ci * ci);
    float ct = std::sqrt(1.f - st2);
    *dir = (-eta) * sc.wo + (eta * ci - ct) * sc.n;
    return false;
}

// ---------------------------------------------------------------------------------------------
// Path tracer (src/ref/path_tracing.cpp:11-133)
// ---------------------------------------------------------------------------------------------
static RGB trace_path(const Scene& s, Ray ray, PixelSampler& ps, PathCounters& pc, Counters* tc) {
    const yar_render_config& cfg = s.config;
    RGB coeff{1, 1, 1}, L;
    int bounce = 0;
    bool delta_traced = false;
    Hit hit{Inf, -1, 0, {}};
    bool hit_found = false;
    Shading sc;
    while (true) {
        if (!delta_traced) {
            hit = Hit{Inf, -1, 0, {}};
            pc.closest++;
            hit_found = intersect(s, ray, hit, tc);
            if (hit_found) init_shading(s, ray, hit, sc);
            if (bounce == 0) {
                if (hit_found && sc.light >= 0) L = L + coeff * area_light_emission(s, sc);
                else if (!hit_found && s.env.present) L = L + coeff * env_filtered(s.env, ray.d);
                if (cfg.max_light_bounces == 0) break;
            }
            if (!hit_found || sc.light >= 0) break;
        }
        delta_traced = false;
        float u_scat = ps.next1d();
        float u_sel = ps.next1d();
        float u_scat_next = ps.next1d();
        V2 u_light = ps.next2d();
        V2 u_bsdf = ps.next2d();
        V2 u_bsdf_next = ps.next2d();
        const Material& m = s.materials[sc.material];
        bool delta = (m.type == YAR_MATERIAL_GLASS && m.roughness <= 0.f) || m.type == YAR_MATERIAL_MIRROR;
        // initialize_scattering (shading_context.cpp:174-186): the uber material picks one of its
        // components with u_scat (get_pbrt_uber_info, delta_scattering.cpp:102-152) and hands the
        // rescaled u_scat on; the BSDF continuation is divided by the probability of the BSDF layers
        float bsdf_layer_probability = 1.f;
        RGB uber_att; V3 uber_dir;
        const Material* mix_delta = nullptr;                                         // the delta sub-material picked
        if (m.type == YAR_MATERIAL_MIX) {                                            // get_mix_info, delta_scattering.cpp:154-189
            const float mix_amount = std::min(luminance(m.amount), 1.f);
            const bool first = u_scat < mix_amount;
            if (first) u_scat /= mix_amount;
            else u_scat = (u_scat - mix_amount) / (1.f - mix_amount);
            const Material& b1 = s.materials[m.sub[0]];
            const Material& b2 = s.materials[m.sub[1]];
            if (!has_bsdf(b1) || !has_bsdf(b2)) {
                const Material& picked = first ? b1 : b2;
                if (!has_bsdf(picked)) { mix_delta = &picked; delta = true; }
                bsdf_layer_probability = 1.f - (has_bsdf(b1) ? 1.f - mix_amount : mix_amount);
            }
        }
        if (m.type == YAR_MATERIAL_PBRT3_UBER) {
            const float fp_index = u_scat * float(m.ncomp);
            const int ci = int(fp_index);
            const int comp = m.comps[ci];
            u_scat = std::clamp(fp_index - float(ci), 0.f, One_Minus_Epsilon);
            if (comp == 2) {                                                         // delta reflection
                const float eta = sc.flipped ? 1.f / m.ior : m.ior;
                const float fresnel = dielectric_fresnel(dot(sc.n, sc.wo), eta);
                uber_att = (float(m.ncomp) * fresnel) * m.kr;
                uber_dir = reflect(sc.wo, sc.n);
                delta = true;
            } else if (comp == 4) {                                                  // opacity: pass through
                uber_att = float(m.ncomp) * (splat(1.f) - m.opacity);
                uber_dir = -sc.wo;
                delta = true;
            }
            float delta_terms = 0.f;
            for (int i = 0; i < m.ncomp; i++) if (m.comps[i] >= 2) delta_terms += 1.f;
            bsdf_layer_probability = 1.f - delta_terms / float(m.ncomp);
        }
        const bool layer_bsdf = has_bsdf(m) && bsdf_layer_probability != 0.f;
        if (!delta) {
            RGB dl = direct_single(s, sc, u_sel, u_light, u_bsdf, u_scat, pc, tc);
            L = L + coeff * dl;
            bounce++;
            if (bounce == cfg.max_light_bounces) break;
            V3 wi; float pdf;
            RGB f = bsdf_sample(s, sc, u_bsdf_next, u_scat_next, sc.wo, &wi, &pdf);
            if (is_black(f)) break;
            coeff = coeff * (f * (std::fabs(dot(sc.n, wi)) / (pdf * bsdf_layer_probability)));
            ray.o = sc.origin_dir(wi);
            ray.d = wi;
        } else {
            if (layer_bsdf) {                                                        // path_tracing.cpp:76-79
                RGB dl = direct_single(s, sc, u_sel, u_light, u_bsdf, u_scat, pc, tc);
                L = L + coeff * dl;
            }
            RGB att; V3 dir;
            const Material& dm = mix_delta ? *mix_delta : m;
            if (dm.type == YAR_MATERIAL_MIRROR) {                                    // delta_scattering.cpp:27-33
                att = dm.refl;
                dir = reflect(sc.wo, sc.n);
            } else if (dm.type == YAR_MATERIAL_PBRT3_UBER) {
                att = uber_att;
                dir = uber_dir;
            } else {
                glass_scatter(sc, dm, u_scat, &att, &dir);
            }
            Ray dray{sc.origin_dir(dir), dir};
            hit = Hit{Inf, -1, 0, {}};
            pc.closest++;
            bool dh = intersect(s, dray, hit, tc);
            if (dh) init_shading(s, dray, hit, sc);
            RGB emitted;
            if (dh && sc.light >= 0) emitted = area_light_emission(s, sc);
            else if (!dh && s.env.present) emitted = env_filtered(s.env, dray.d);
            coeff = coeff * att;
            L = L + coeff * emitted;
            bounce++;
            if (bounce == cfg.max_light_bounces) break;
            if (!dh || sc.light >= 0) break;
            ray = dray;
            delta_traced = true;
        }
        if (bounce >= cfg.russian_roulette_bounce_count_threshold) {
            float u_term = ps.next1d();
            float mc = std::max(coeff.r, std::max(coeff.g, coeff.b));
            if (mc < cfg.russian_roulette_threshold) {
                float p = std::max(0.05f, 1.f - mc);
                if (u_term < p) break;
                float inv = 1.f / (1.f - p);
                coeff = coeff * inv;
            }
        }
    }
    return L;
}

// ---------------------------------------------------------------------------------------------
// Camera (src/ref/camera.cpp:5-32) and film (src/ref/film.cpp)
// ---------------------------------------------------------------------------------------------
struct Camera {
    M34 c2w; float W, H, horz, vert; bool zup;
    void init(const yar_view& v) {
        c2w = load_m34(v.camera_to_world);
        W = (float)v.film_width; H = (float)v.film_height; zup = v.z_is_up != 0;
        constexpr float deg2rad = Pi / 180.f;
        float tan_half = std::tan((v.fov_y_degrees / 2.f) * deg2rad);
        vert = tan_half;
        horz = (W / H) * tan_half;
    }
    Ray gen(float fx, float fy) const {
        float u = 2.f * (fx / W) - 1.f;
        float v = 2.f * (fy / H) - 1.f;
        float right = u * horz, up = -v * vert;
        V3 d{right, zup ? 1.f : up, zup ? up : -1.f};
        d = vdiv(d, length(d));
        return Ray{xform_point(c2w, V3{0, 0, 0}), xform_vector(c2w, d)};
    }
};

struct Region { int x0, y0, x1, y1; int w() const { return x1 - x0; } int h() const { return y1 - y0; } };

struct TileResult {
    Region pix;
    std::vector<yar_film_pixel> px;
    double var = 0;
};

struct RenderOut { int64_t closest = 0, any = 0; Counters trav; };

// Pixels [x0, x1) x [y0, y1) a sample at film position (fx, fy) reaches (Film_Tile::add_sample,
// film.cpp:27-50): pixel centres within the filter radius, before clipping to the tile.
static void film_footprint(float fx, float fy, float rad, int* x0, int* y0, int* x1, int* y1) {
    *x0 = (int)std::ceil(fx - rad - 0.5f);
    *y0 = (int)std::ceil(fy - rad - 0.5f);
    *x1 = (int)std::floor(fx + rad - 0.5f) + 1;
    *y1 = (int)std::floor(fy + rad - 0.5f) + 1;
}

// render_tile (reference_renderer.cpp:331-429) for one 64x64 sample tile
static void render_tile(const Scene& s, const yar_view& view, const Camera& cam, const Region& region,
                        const Region& sb, TileResult& tr, PathCounters& pc, Counters* tc, float* per_sample) {
    const yar_render_config& cfg = s.config;
    int md = std::min(10, cfg.max_light_bounces);
    Pcg rng;
    PixelSampler ps;
    ps.init(view.x_pixel_sample_count, view.y_pixel_sample_count, 4 * md, 3 * md, &rng);
    int spp = ps.spp;
    // pixel bounds of the tile (film.cpp:80-85, radius 0.5)
    const float rad = cfg.pixel_filter_radius;
    tr.pix.x0 = std::max((int)std::ceil(sb.x0 - rad - 0.5f), region.x0);
    tr.pix.y0 = std::max((int)std::ceil(sb.y0 - rad - 0.5f), region.y0);
    tr.pix.x1 = std::min((int)std::floor(sb.x1 + rad - 0.5f) + 1, region.x1);
    tr.pix.y1 = std::min((int)std::floor(sb.y1 + rad - 0.5f) + 1, region.y1);
    tr.px.assign((size_t)tr.pix.w() * tr.pix.h(), yar_film_pixel{{0, 0, 0}, 0});
    tr.var = 0;
    for (int y = sb.y0; y < sb.y1; y++) {
        for (int x = sb.x0; x < sb.x1; x++) {
            uint32_t stream = ((uint32_t)x & 0xffffu) | ((uint32_t)y << 16);
            stream += view.rng_seed_offset;
            pcg_seed(rng, 0, stream);
            ps.next_pixel();
            double lsum = 0.0, lsq = 0.0;
            do {
                V2 j = ps.image[ps.cur];
                float fx = (float)x + j.x, fy = (float)y + j.y;
                Ray ray = cam.gen(fx, fy);
                RGB rad_ = trace_path(s, ray, ps, pc, tc);
                float limit = cfg.max_rgb_component_value_of_film_sample;
                float mc = std::max(rad_.r, std::max(rad_.g, rad_.b));
                if (mc > limit) rad_ = rad_ * (limit / mc);
                if (cfg.film_radiance_scale != 1.f) rad_ = rad_ * cfg.film_radiance_scale;
                if (per_sample) {
                    size_t o = (((size_t)(y - region.y0) * region.w() + (x - region.x0)) * spp + ps.cur) * 3;
                    per_sample[o] = rad_.r; per_sample[o + 1] = rad_.g; per_sample[o + 2] = rad_.b;
                }
                // Film_Tile::add_sample (film.cpp:27-50), box filter weight 1
                int px0, py0, px1, py1;
                film_footprint(fx, fy, rad, &px0, &py0, &px1, &py1);
                px0 = std::max(px0, tr.pix.x0); py0 = std::max(py0, tr.pix.y0);
                px1 = std::min(px1, tr.pix.x1); py1 = std::min(py1, tr.pix.y1);
                for (int yy = py0; yy < py1; yy++)
                    for (int xx = px0; xx < px1; xx++) {
                        yar_film_pixel& fp = tr.px[(size_t)(yy - tr.pix.y0) * tr.pix.w() + (xx - tr.pix.x0)];
                        fp.color_sum[0] += 1.f * rad_.r; fp.color_sum[1] += 1.f * rad_.g; fp.color_sum[2] += 1.f * rad_.b;
                        fp.weight_sum += 1.f;
                    }
                float lum = luminance(rad_);
                lsum += lum;
                lsq += lum * lum;
            } while (ps.next_sample());
            if (spp > 1) {
                int n = spp;
                double pv = (lsq - lsum * lsum / n) / (n * (n - 1));
                pv = std::max(0.0, pv);
                tr.var += pv;
            }
        }
    }
}

static bool load_scene(const yar_scene_desc* d, Scene& s, int threads) {
    s.meshes.resize(d->mesh_count);
    for (int i = 0; i < d->mesh_count; i++) {
        const yar_mesh_desc& md = d->meshes[i];
        Mesh& m = s.meshes[i];
        m.v.assign(md.vertices, md.vertices + 3 * (size_t)md.vertex_count);
        if (md.normals) m.n.assign(md.normals, md.normals + 3 * (size_t)md.vertex_count);
        if (md.uvs) m.uv.assign(md.uvs, md.uvs + 2 * (size_t)md.vertex_count);
        m.idx.assign(md.indices, md.indices + 3 * (size_t)md.triangle_count);
        m.reverse = md.reverse_geometric_normal_orientation != 0;
    }
    // per-mesh trees on worker threads (scene_context.cpp:33-61)
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < d->mesh_count; i = next++) {
            Mesh& m = s.meshes[i];
            TriSource src{m.v.data(), m.idx.data()};
            uint32_t nt = (uint32_t)(m.idx.size() / 3);
            std::vector<Box> boxes(nt);
            for (uint32_t t = 0; t < nt; t++) {                                        // triangle_mesh.h:100
                V3 p[3]; m.pos(t, p);
                Box b; b.lo = p[0]; b.hi = p[0]; box_add(b, p[1]); box_add(b, p[2]);
                boxes[t] = b;
            }
            m.tree = build_tree(boxes, &src);
        }
    };
    {
        std::vector<std::thread> pool;
        int nth = std::max(1, std::min(threads, d->mesh_count));
        for (int t = 1; t < nth; t++) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
    }
    s.objects.resize(d->object_count);
    std::vector<Box> wb(d->object_count);
    for (int i = 0; i < d->object_count; i++) {
        const yar_object_desc& od = d->objects[i];
        Object& o = s.objects[i];
        o.mesh = od.mesh; o.material = od.material; o.light = od.area_light;
        o.o2w = load_m34(od.object_to_world); o.w2o = load_m34(od.world_to_object);
        std::memset(o.nrm.a, 0, sizeof(o.nrm.a));                                      // scene_loader.cpp:56-62
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) o.nrm.a[r][c] = o.w2o.a[c][r];
        const Box& lb = s.meshes[o.mesh].tree.bounds;                                 // matrix.cpp:391
        Box b;
        for (int k = 0; k < 8; k++) {
            V3 c{(k & 4) ? lb.hi.x : lb.lo.x, (k & 2) ? lb.hi.y : lb.lo.y, (k & 1) ? lb.hi.z : lb.lo.z};
            box_add(b, xform_point(o.o2w, c));
        }
        wb[i] = b;
    }
    s.scene_tree = build_tree(wb, nullptr);
    s.materials.resize(d->material_count);
    for (int i = 0; i < d->material_count; i++) {
        const yar_material_desc& md = d->materials[i];
        Material& m = s.materials[i];
        m.type = md.type;
        m.refl = {md.reflectance[0], md.reflectance[1], md.reflectance[2]};
        m.trans = {md.transmittance[0], md.transmittance[1], md.transmittance[2]};
        m.eta = {md.eta[0], md.eta[1], md.eta[2]};
        m.k = {md.k[0], md.k[1], md.k[2]};
        if (md.type == YAR_MATERIAL_COATED_DIFFUSE) m.k = {md.specular[0], md.specular[1], md.specular[2]};   // r0
        m.eta_i = md.eta_i;
        m.ior = md.index_of_refraction;
        m.roughness = md.u_roughness;
        m.r0 = md.specular_r0;
        m.pbrt3 = d->config.pbrt3_scene != 0;
        auto remap = [&](float r) {                                                   // scattering.cpp:240-269
            if (md.roughness_is_alpha) return r;
            if (d->config.pbrt3_scene) {
                r = std::max(r, 1e-3f);
                float x = std::log(r);
                return 1.621420000f + 0.819955000f * x + 0.173400000f * x * x + 0.017120100f * x * x * x +
                       0.000640711f * x * x * x * x;
            }
            return r * r;
        };
        m.alpha_x = remap(md.u_roughness);
        m.alpha_y = remap(md.v_roughness);
        if (md.type == YAR_MATERIAL_MIX) {
            m.sub[0] = md.mix_materials[0];
            m.sub[1] = md.mix_materials[1];
            m.amount = {md.mix_amount[0], md.mix_amount[1], md.mix_amount[2]};
        }
        if (md.type == YAR_MATERIAL_PBRT3_UBER) {                                    // pbrt_scene.cpp:585-608
            m.ks = {md.specular[0], md.specular[1], md.specular[2]};
            m.kr = {md.delta_reflectance[0], md.delta_reflectance[1], md.delta_reflectance[2]};
            m.opacity = {md.opacity[0], md.opacity[1], md.opacity[2]};
            if (!is_black(m.refl)) m.comps[m.ncomp++] = 0;
            if (!is_black(m.ks)) m.comps[m.ncomp++] = 1;
            if (!is_black(m.kr)) m.comps[m.ncomp++] = 2;
            if (!(m.opacity.r == 1.f && m.opacity.g == 1.f && m.opacity.b == 1.f)) m.comps[m.ncomp++] = 4;
        }
    }
    for (int i = 0; i < d->rect_light_count; i++) {
        const yar_rect_light_desc& ld = d->rect_lights[i];
        RectLight L; L.l2w = load_m34(ld.light_to_world);
        L.L = {ld.emitted_radiance[0], ld.emitted_radiance[1], ld.emitted_radiance[2]};
        L.size.x = ld.size[0]; L.size.y = ld.size[1];
        s.rects.push_back(L);
    }
    for (int i = 0; i < d->point_light_count; i++) {
        const yar_point_light_desc& ld = d->point_lights[i];
        s.points.push_back(PointLight{{ld.position[0], ld.position[1], ld.position[2]},
                                      {ld.intensity[0], ld.intensity[1], ld.intensity[2]}});
    }
    for (int i = 0; i < d->spot_light_count; i++) {
        const yar_spot_light_desc& ld = d->spot_lights[i];
        s.spots.push_back(SpotLight{{ld.position[0], ld.position[1], ld.position[2]},
                                    {ld.direction[0], ld.direction[1], ld.direction[2]}, ld.cone_angle, ld.penumbra_angle,
                                    {ld.intensity[0], ld.intensity[1], ld.intensity[2]}});
    }
    for (int i = 0; i < d->directional_light_count; i++) {
        const yar_directional_light_desc& ld = d->directional_lights[i];
        s.dirs.push_back(DirectionalLight{{ld.direction[0], ld.direction[1], ld.direction[2]},
                                          {ld.irradiance[0], ld.irradiance[1], ld.irradiance[2]}});
    }
    for (int i = 0; i < d->mesh_light_count; i++) {                          // reference_renderer.cpp:137-152
        const yar_mesh_light_desc& ld = d->mesh_lights[i];
        MeshLight ml;
        ml.object = ld.object;
        ml.mesh = s.objects[ld.object].mesh;
        ml.L = {ld.emitted_radiance[0], ld.emitted_radiance[1], ld.emitted_radiance[2]};
        const Mesh& m = s.meshes[ml.mesh];
        const int n = (int)(m.idx.size() / 3);
        std::vector<float> areas(n);
        for (int t = 0; t < n; t++) {                                         // triangle_mesh.h:103-118
            V3 p[3]; m.pos(t, p);
            areas[t] = length(cross(p[1] - p[0], p[2] - p[0])) * 0.5f;
        }
        ml.area = 0.f;
        for (int t = 0; t < n; t++) ml.area += areas[t];
        float total = 0.f;
        for (int t = 0; t < n; t++) total += areas[t];
        ml.cdf.resize(n);
        float sum = 0.f;
        for (int t = 0; t < n; t++) { sum += areas[t]; ml.cdf[t] = sum / total; }
        s.mesh_lights.push_back(std::move(ml));
    }
    for (int i = 0; i < d->sphere_light_count; i++) {
        const yar_sphere_light_desc& ld = d->sphere_lights[i];
        s.spheres.push_back(SphereLight{{ld.position[0], ld.position[1], ld.position[2]}, ld.radius,
                                        {ld.emitted_radiance[0], ld.emitted_radiance[1], ld.emitted_radiance[2]}});
    }
    if (d->environment_light_count > 0) {
        const yar_environment_light_desc& ld = d->environment_light[0];
        s.env.present = true;
        s.env.l2w = load_m34(ld.light_to_world);
        s.env.w2l = load_m34(ld.world_to_light);
        s.env.L = {ld.radiance[0], ld.radiance[1], ld.radiance[2]};
    }
    s.total_lights = (int)(s.points.size() + s.spots.size() + s.dirs.size() + s.rects.size() + s.spheres.size() +
                           s.mesh_lights.size()) + (s.env.present ? 1 : 0);
    s.camera = d->camera;
    s.config = d->config;
    return true;
}

}  // namespace oref

using namespace oref;

extern "C" {

void oref_pcg32(uint64_t initstate, uint64_t initseq, int32_t n, uint32_t* out) {
    Pcg r; pcg_seed(r, initstate, initseq);
    for (int i = 0; i < n; i++) out[i] = pcg_next(r);
}

// Sample table of one pixel in GPU layout [dim][sample]: img (2*spp), s1 (d1*spp), s2 (2*d2*spp);
// also returns the RNG state after next_pixel (the fallback stream).
void oref_pixel_samples(int32_t x, int32_t y, uint32_t seed_offset, int32_t nx, int32_t ny, int32_t d1, int32_t d2,
                        float* img, float* s1, float* s2, uint64_t* state_after) {
    Pcg rng;
    uint32_t stream = ((uint32_t)x & 0xffffu) | ((uint32_t)y << 16);
    stream += seed_offset;
    pcg_seed(rng, 0, stream);
    PixelSampler ps; ps.init(nx, ny, d1, d2, &rng);
    ps.next_pixel();
    int spp = nx * ny;
    for (int k = 0; k < spp; k++) { img[2 * k] = ps.image[k].x; img[2 * k + 1] = ps.image[k].y; }
    for (int i = 0; i < d1; i++) for (int k = 0; k < spp; k++) s1[(size_t)i * spp + k] = ps.s1[(size_t)k * d1 + i];
    for (int i = 0; i < d2; i++) for (int k = 0; k < spp; k++) {
        s2[((size_t)i * spp + k) * 2] = ps.s2[(size_t)k * d2 + i].x;
        s2[((size_t)i * spp + k) * 2 + 1] = ps.s2[(size_t)k * d2 + i].y;
    }
    if (state_after) *state_after = rng.state;
}

float oref_watertight(const float* ray6, const float* tri9, float* b3) {
    Ray r{{ray6[0], ray6[1], ray6[2]}, {ray6[3], ray6[4], ray6[5]}};
    V3 b;
    float t = intersect_watertight(r, {tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]}, {tri9[6], tri9[7], tri9[8]}, &b);
    if (b3) { b3[0] = b.x; b3[1] = b.y; b3[2] = b.z; }
    return t;
}

// kd-tree over a triangle mesh; caller frees *out_words with oref_free.
int32_t oref_kdtree_build_mesh(const float* v, const int32_t* idx, int32_t ntri, uint32_t** out_words,
                               uint32_t* out_node_count, float* out_bounds, int32_t* out_depth) {
    TriSource src{v, idx};
    std::vector<Box> boxes(ntri);
    for (int t = 0; t < ntri; t++) {
        V3 p[3]; src.get(t, p);
        Box b; b.lo = p[0]; b.hi = p[0]; box_add(b, p[1]); box_add(b, p[2]);
        boxes[t] = b;
    }
    Tree tr = build_tree(boxes, &src);
    *out_node_count = (uint32_t)(tr.words.size() / 2);
    *out_words = (uint32_t*)std::malloc(tr.words.size() * 4 + 4);
    std::memcpy(*out_words, tr.words.data(), tr.words.size() * 4);
    if (out_bounds) { for (int a = 0; a < 3; a++) { out_bounds[a] = tr.bounds.lo[a]; out_bounds[3 + a] = tr.bounds.hi[a]; } }
    if (out_depth) *out_depth = tr.depth;
    return 0;
}
int32_t oref_kdtree_build_boxes(const float* boxes6, int32_t n, uint32_t** out_words, uint32_t* out_node_count,
                                float* out_bounds, int32_t* out_depth) {
    std::vector<Box> boxes(n);
    for (int i = 0; i < n; i++) { boxes[i].lo = {boxes6[6 * i], boxes6[6 * i + 1], boxes6[6 * i + 2]}; boxes[i].hi = {boxes6[6 * i + 3], boxes6[6 * i + 4], boxes6[6 * i + 5]}; }
    Tree tr = build_tree(boxes, nullptr);
    *out_node_count = (uint32_t)(tr.words.size() / 2);
    *out_words = (uint32_t*)std::malloc(tr.words.size() * 4 + 4);
    std::memcpy(*out_words, tr.words.data(), tr.words.size() * 4);
    if (out_bounds) { for (int a = 0; a < 3; a++) { out_bounds[a] = tr.bounds.lo[a]; out_bounds[3 + a] = tr.bounds.hi[a]; } }
    if (out_depth) *out_depth = tr.depth;
    return 0;
}
void oref_free(void* p) { std::free(p); }

void* oref_scene_create(const yar_scene_desc* desc, int32_t threads) {
    Scene* s = new Scene;
    load_scene(desc, *s, threads);
    return s;
}
void oref_scene_free(void* s) { delete (Scene*)s; }

int32_t oref_scene_tree(void* sp, int32_t mesh, uint32_t** out_words, uint32_t* out_node_count, float* out_bounds) {
    Scene& s = *(Scene*)sp;
    const Tree& t = mesh < 0 ? s.scene_tree : s.meshes[mesh].tree;
    *out_node_count = (uint32_t)(t.words.size() / 2);
    *out_words = (uint32_t*)std::malloc(t.words.size() * 4 + 4);
    std::memcpy(*out_words, t.words.data(), t.words.size() * 4);
    for (int a = 0; a < 3; a++) { out_bounds[a] = t.bounds.lo[a]; out_bounds[3 + a] = t.bounds.hi[a]; }
    return t.depth;
}

// counters: 4 int64 per ray (nodes, leaf refs, triangle tests, objects) or NULL.
int32_t oref_trace(void* sp, const yar_ray* rays, int64_t n, int32_t any, int32_t brute, yar_hit* out,
                   int64_t* counters, int32_t threads) {
    const Scene& s = *(const Scene*)sp;
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        const int64_t chunk = 1024;
        for (int64_t base = next.fetch_add(chunk); base < n; base = next.fetch_add(chunk)) {
            int64_t end = std::min(n, base + chunk);
            for (int64_t i = base; i < end; i++) {
                const yar_ray& r = rays[i];
                Ray ray{{r.origin[0], r.origin[1], r.origin[2]}, {r.direction[0], r.direction[1], r.direction[2]}};
                Counters c;
                Counters* cp = counters ? &c : nullptr;
                yar_hit& h = out[i];
                if (any) {
                    bool occ = intersect_any(s, ray, r.tmax, cp);
                    h.t = occ ? 1.f : 0.f; h.object = -1; h.triangle = -1; h.barycentrics[0] = h.barycentrics[1] = h.barycentrics[2] = 0;
                } else {
                    Hit hit{r.tmax, -1, 0, {}};
                    bool found = brute ? intersect_brute(s, ray, hit) : intersect(s, ray, hit, cp);
                    h.t = found ? hit.t : Inf;
                    h.object = found ? hit.obj : -1;
                    h.triangle = found ? (int32_t)hit.tri : -1;
                    h.barycentrics[0] = found ? hit.b.x : 0; h.barycentrics[1] = found ? hit.b.y : 0; h.barycentrics[2] = found ? hit.b.z : 0;
                }
                if (counters) { counters[4 * i] = c.nodes; counters[4 * i + 1] = c.leaf_refs; counters[4 * i + 2] = c.tri_tests; counters[4 * i + 3] = c.objects; }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::max(1, threads); t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return 0;
}

// Renders a view's render region (reference_renderer.cpp:479-584).  film_out: region Film_Pixels
// merged in tile order; per_sample (optional): region pixels * spp * 3 radiance values;
// stats_out (optional, 6 int64): closest, any, nodes, leaf refs, tri tests, objects.
int32_t oref_render(void* sp, const yar_view* view, yar_film_pixel* film_out, float* image_out, double* variance_out,
                    float* per_sample, int64_t* stats_out, int32_t threads) {
    const Scene& s = *(const Scene*)sp;
    Camera cam; cam.init(*view);
    Region region{view->region_x0, view->region_y0, view->region_x1, view->region_y1};
    const float rad = s.config.pixel_filter_radius;
    Region sreg{(int)std::floor(region.x0 + 0.5f - rad), (int)std::floor(region.y0 + 0.5f - rad),
                (int)std::ceil(region.x1 - 1 + 0.5f + rad), (int)std::ceil(region.y1 - 1 + 0.5f + rad)};
    int tgx = (sreg.w() + 63) / 64, tgy = (sreg.h() + 63) / 64;
    int ntiles = tgx * tgy;
    std::vector<TileResult> tiles(ntiles);
    std::vector<Region> sbs(ntiles);
    for (int t = 0; t < ntiles; t++) {
        int tx = t % tgx, ty = t / tgx;
        Region& sb = sbs[t];
        sb.x0 = sreg.x0 + tx * 64; sb.y0 = sreg.y0 + ty * 64;
        sb.x1 = std::min(sb.x0 + 64, sreg.x1); sb.y1 = std::min(sb.y0 + 64, sreg.y1);
    }
    std::atomic<int> next{0};
    std::vector<PathCounters> pcs(std::max(1, threads));
    std::vector<Counters> tcs(std::max(1, threads));
    bool want_trav = stats_out != nullptr;
    auto work = [&](int tid) {
        for (int t = next++; t < ntiles; t = next++)
            render_tile(s, *view, cam, region, sbs[t], tiles[t], pcs[tid], want_trav ? &tcs[tid] : nullptr, per_sample);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::max(1, threads); t++) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    std::vector<yar_film_pixel> film((size_t)region.w() * region.h(), yar_film_pixel{{0, 0, 0}, 0});
    for (int t = 0; t < ntiles; t++) {                                                   // merge_tile, film.cpp:88
        const TileResult& tr = tiles[t];
        for (int y = tr.pix.y0; y < tr.pix.y1; y++)
            for (int x = tr.pix.x0; x < tr.pix.x1; x++) {
                yar_film_pixel& f = film[(size_t)(y - region.y0) * region.w() + (x - region.x0)];
                const yar_film_pixel& p = tr.px[(size_t)(y - tr.pix.y0) * tr.pix.w() + (x - tr.pix.x0)];
                f.color_sum[0] += p.color_sum[0]; f.color_sum[1] += p.color_sum[1]; f.color_sum[2] += p.color_sum[2];
                f.weight_sum += p.weight_sum;
            }
    }
    if (film_out) std::memcpy(film_out, film.data(), film.size() * sizeof(yar_film_pixel));
    if (image_out) {                                                                      // get_image, film.cpp:102
        for (size_t i = 0; i < film.size(); i++) {
            const yar_film_pixel& f = film[i];
            RGB c;
            if (f.weight_sum != 0.f) c = cdivs(RGB{f.color_sum[0], f.color_sum[1], f.color_sum[2]}, f.weight_sum);
            image_out[3 * i] = std::max(0.f, c.r); image_out[3 * i + 1] = std::max(0.f, c.g); image_out[3 * i + 2] = std::max(0.f, c.b);
        }
    }
    int spp = view->x_pixel_sample_count * view->y_pixel_sample_count;
    if (variance_out) {
        *variance_out = 0;
        if (spp > 1) {
            double acc = 0; int64_t cnt = 0;
            for (int t = 0; t < ntiles; t++) { acc += tiles[t].var; cnt += (int64_t)sbs[t].w() * sbs[t].h(); }
            *variance_out = acc / cnt;
        }
    }
    if (stats_out) {
        for (int i = 0; i < 6; i++) stats_out[i] = 0;
        for (size_t t = 0; t < pcs.size(); t++) {
            stats_out[0] += pcs[t].closest; stats_out[1] += pcs[t].any;
            stats_out[2] += tcs[t].nodes; stats_out[3] += tcs[t].leaf_refs; stats_out[4] += tcs[t].tri_tests; stats_out[5] += tcs[t].objects;
        }
    }
    return 0;
}

// BSDF probes in a local frame (T = x, B = y, n = z) for the tests' mathematical pins of the
// oracle's BSDF code (white furnace, Fresnel closed forms, pdf normalisation, sample/pdf
// consistency, reciprocity).  type: YAR_MATERIAL_DIFFUSE or YAR_MATERIAL_METAL.
static Scene probe_scene(int type, const float* refl, const float* eta, const float* k, float eta_i, float ax, float ay) {
    Scene s;
    Material m{};
    m.type = type;
    m.refl = RGB{refl[0], refl[1], refl[2]};
    m.eta = RGB{eta[0], eta[1], eta[2]};
    m.k = RGB{k[0], k[1], k[2]};
    m.eta_i = eta_i; m.alpha_x = ax; m.alpha_y = ay;
    s.materials.push_back(m);
    return s;
}
static Shading probe_frame() {
    Shading sc;
    sc.material = 0;
    sc.T = V3(1, 0, 0); sc.B = V3(0, 1, 0); sc.n = V3(0, 0, 1); sc.ng = V3(0, 0, 1);
    return sc;
}

void oref_bsdf_eval_local(int32_t type, const float* refl, const float* eta, const float* k, float eta_i, float ax,
                          float ay, const float* wo, const float* wi, float* f_out, float* pdf_out) {
    Scene s = probe_scene(type, refl, eta, k, eta_i, ax, ay);
    Shading sc = probe_frame();
    V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    RGB f = bsdf_eval(s, sc, o, i);
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = bsdf_pdf(s, sc, o, i);
}

int32_t oref_bsdf_sample_local(int32_t type, const float* refl, const float* eta, const float* k, float eta_i, float ax,
                               float ay, const float* wo, const float* u, float* wi_out, float* f_out, float* pdf_out) {
    Scene s = probe_scene(type, refl, eta, k, eta_i, ax, ay);
    Shading sc = probe_frame();
    V3 wi;
    float pdf = 0.f;
    RGB f = bsdf_sample(s, sc, V2{u[0], u[1]}, 0.f, V3(wo[0], wo[1], wo[2]), &wi, &pdf);
    wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z;
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = pdf;
    return (f.r != 0.f || f.g != 0.f || f.b != 0.f) ? 1 : 0;
}

float oref_dielectric_fresnel(float cos_i, float eta) { return dielectric_fresnel(cos_i, eta); }

// Plastic BRDF in the local frame (n = z): evaluate + pdf at (wo, wi), or sample with (u0, u1, u_scat).
// pbrt3 selects Pbrt3_Plastic_BRDF; flipped = the original shading normal faced away from wo.
void oref_plastic_eval(int32_t pbrt3, const float* kd, float r0, float alpha, int32_t flipped, const float* wo,
                       const float* wi, float* f_out, float* pdf_out) {
    Material m{};
    m.type = YAR_MATERIAL_PLASTIC; m.refl = RGB{kd[0], kd[1], kd[2]}; m.r0 = r0; m.alpha_x = m.alpha_y = alpha;
    m.pbrt3 = pbrt3 != 0;
    Shading sc = probe_frame();
    sc.flipped = flipped != 0;
    const V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    const RGB f = plastic_eval(m, sc, o, i);
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = plastic_pdf(m, sc, o, i);
}
int32_t oref_plastic_sample(int32_t pbrt3, const float* kd, float r0, float alpha, int32_t flipped, const float* wo,
                            const float* u3, float* wi_out, float* f_out, float* pdf_out) {
    Scene s;
    Material m{};
    m.type = YAR_MATERIAL_PLASTIC; m.refl = RGB{kd[0], kd[1], kd[2]}; m.r0 = r0; m.alpha_x = m.alpha_y = alpha;
    m.pbrt3 = pbrt3 != 0;
    s.materials.push_back(m);
    Shading sc = probe_frame();
    sc.flipped = flipped != 0;
    V3 wi;
    float pdf = 0.f;
    const RGB f = bsdf_sample(s, sc, V2{u3[0], u3[1]}, u3[2], V3(wo[0], wo[1], wo[2]), &wi, &pdf);
    wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z;
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = pdf;
    return (f.r != 0.f || f.g != 0.f || f.b != 0.f) ? 1 : 0;
}

// Coated diffuse BRDF in the local frame (n = z, T = x, B = y): evaluate + pdf, or sample.
void oref_coated_eval(const float* kd, const float* r0, float ax, float ay, const float* wo, const float* wi,
                      float* f_out, float* pdf_out) {
    Material m{};
    m.type = YAR_MATERIAL_COATED_DIFFUSE; m.refl = RGB{kd[0], kd[1], kd[2]}; m.k = RGB{r0[0], r0[1], r0[2]};
    m.alpha_x = ax; m.alpha_y = ay;
    Shading sc = probe_frame();
    const V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    const RGB f = coated_eval(m, sc, o, i);
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = coated_pdf(m, sc, o, i);
}
int32_t oref_coated_sample(const float* kd, const float* r0, float ax, float ay, const float* wo, const float* u3,
                           float* wi_out, float* f_out, float* pdf_out) {
    Scene s;
    Material m{};
    m.type = YAR_MATERIAL_COATED_DIFFUSE; m.refl = RGB{kd[0], kd[1], kd[2]}; m.k = RGB{r0[0], r0[1], r0[2]};
    m.alpha_x = ax; m.alpha_y = ay;
    s.materials.push_back(m);
    Shading sc = probe_frame();
    V3 wi;
    float pdf = 0.f;
    const RGB f = bsdf_sample(s, sc, V2{u3[0], u3[1]}, u3[2], V3(wo[0], wo[1], wo[2]), &wi, &pdf);
    wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z;
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = pdf;
    return (f.r != 0.f || f.g != 0.f || f.b != 0.f) ? 1 : 0;
}

// pbrt3 uber BRDF in the local frame (n = z, T = x, B = y): evaluate + pdf, or sample; flipped =
// the original normal faced away (ior inverted).
static Material uber_probe(const float* kd, const float* ks, const float* opacity, float ior, float ax, float ay) {
    Material m{};
    m.type = YAR_MATERIAL_PBRT3_UBER; m.ior = ior; m.alpha_x = ax; m.alpha_y = ay;
    m.refl = RGB{kd[0], kd[1], kd[2]}; m.ks = RGB{ks[0], ks[1], ks[2]}; m.opacity = RGB{opacity[0], opacity[1], opacity[2]};
    return m;
}
void oref_uber_eval(const float* kd, const float* ks, const float* opacity, float ior, float ax, float ay, int32_t flipped,
                    const float* wo, const float* wi, float* f_out, float* pdf_out) {
    const Material m = uber_probe(kd, ks, opacity, ior, ax, ay);
    Shading sc = probe_frame();
    sc.flipped = flipped != 0;
    const V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    const RGB f = uber_eval(m, sc, o, i);
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = uber_pdf(m, sc, o, i);
}
int32_t oref_uber_sample(const float* kd, const float* ks, const float* opacity, float ior, float ax, float ay,
                         int32_t flipped, const float* wo, const float* u3, float* wi_out, float* f_out, float* pdf_out) {
    Scene s;
    s.materials.push_back(uber_probe(kd, ks, opacity, ior, ax, ay));
    Shading sc = probe_frame();
    sc.flipped = flipped != 0;
    V3 wi;
    float pdf = 0.f;
    const RGB f = bsdf_sample(s, sc, V2{u3[0], u3[1]}, u3[2], V3(wo[0], wo[1], wo[2]), &wi, &pdf);
    wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z;
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = pdf;
    return (f.r != 0.f || f.g != 0.f || f.b != 0.f) ? 1 : 0;
}

// Rough glass BSDF in the local frame (n = z): evaluate + pdf, or sample; flipped = leaving the dielectric.
static Material rough_glass_probe(float ior, float alpha, const float* kr, const float* kt) {
    Material m{};
    m.type = YAR_MATERIAL_GLASS; m.ior = ior; m.alpha_x = m.alpha_y = alpha; m.roughness = 1.f;
    m.refl = RGB{kr[0], kr[1], kr[2]}; m.trans = RGB{kt[0], kt[1], kt[2]};
    return m;
}
void oref_rough_glass_eval(float ior, float alpha, const float* kr, const float* kt, int32_t flipped, const float* wo,
                           const float* wi, float* f_out, float* pdf_out) {
    Material m = rough_glass_probe(ior, alpha, kr, kt);
    Shading sc = probe_frame();
    sc.flipped = flipped != 0;
    const V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    const RGB f = rough_glass_eval(m, sc, o, i);
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = rough_glass_pdf(m, sc, o, i);
}
int32_t oref_rough_glass_sample(float ior, float alpha, const float* kr, const float* kt, int32_t flipped,
                                const float* wo, const float* u3, float* wi_out, float* f_out, float* pdf_out) {
    Scene s;
    s.materials.push_back(rough_glass_probe(ior, alpha, kr, kt));
    Shading sc = probe_frame();
    sc.flipped = flipped != 0;
    V3 wi;
    float pdf = 0.f;
    const RGB f = bsdf_sample(s, sc, V2{u3[0], u3[1]}, u3[2], V3(wo[0], wo[1], wo[2]), &wi, &pdf);
    wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z;
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = pdf;
    return (f.r != 0.f || f.g != 0.f || f.b != 0.f) ? 1 : 0;
}

// Smooth-glass event at a surface with shading normal n (facing wo), original orientation flipped or not.
int32_t oref_glass_scatter(const float* n, const float* wo, int32_t flipped, float ior, const float* kr, const float* kt,
                           float u_scat, float* dir_out, float* att_out) {
    Shading sc;
    sc.n = V3(n[0], n[1], n[2]);
    sc.wo = V3(wo[0], wo[1], wo[2]);
    sc.flipped = flipped != 0;
    Material m{};
    m.type = YAR_MATERIAL_GLASS;
    m.ior = ior;
    m.refl = RGB{kr[0], kr[1], kr[2]};
    m.trans = RGB{kt[0], kt[1], kt[2]};
    RGB att; V3 dir;
    const bool refl = glass_scatter(sc, m, u_scat, &att, &dir);
    dir_out[0] = dir.x; dir_out[1] = dir.y; dir_out[2] = dir.z;
    att_out[0] = att.r; att_out[1] = att.g; att_out[2] = att.b;
    return refl ? 1 : 0;
}

// adjust_shading_normal (shading_context.cpp:17-63): returns 1 when the normal was changed.
int32_t oref_adjust_normal(const float* wo, const float* ng, const float* n, float* out) {
    V3 nn(n[0], n[1], n[2]);
    const bool changed = adjust_normal(V3(wo[0], wo[1], wo[2]), V3(ng[0], ng[1], ng[2]), &nn);
    out[0] = nn.x; out[1] = nn.y; out[2] = nn.z;
    return changed ? 1 : 0;
}

// Camera::generate_ray (camera.cpp:16-32) of a view at film position (fx, fy).
void oref_camera_ray(const yar_view* view, float fx, float fy, float* o_out, float* d_out) {
    Camera cam;
    cam.init(*view);
    const Ray r = cam.gen(fx, fy);
    o_out[0] = r.o.x; o_out[1] = r.o.y; o_out[2] = r.o.z;
    d_out[0] = r.d.x; d_out[1] = r.d.y; d_out[2] = r.d.z;
}

// Pixels [x0, x1) x [y0, y1) a film sample at pixel (x, y) + jitter reaches (Film_Tile::add_sample,
// film.cpp:27-50, box filter of the given radius), before clipping to the tile.
void oref_film_footprint(int32_t x, int32_t y, float jx, float jy, float radius, int32_t* out4) {
    film_footprint((float)x + jx, (float)y + jy, radius, &out4[0], &out4[1], &out4[2], &out4[3]);
}

void oref_conductor_fresnel(float cos_i, float eta_i, const float* eta, const float* k, float* out) {
    RGB r = conductor_fresnel(cos_i, eta_i, RGB{eta[0], eta[1], eta[2]}, RGB{k[0], k[1], k[2]});
    out[0] = r.r; out[1] = r.g; out[2] = r.b;
}

}  // extern "C"

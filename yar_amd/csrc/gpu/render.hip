// render.hip - wavefront path-tracer kernels for gfx950 (CDNA4).
//
// One path per work item; the bounce loop is split into queue-driven kernels:
//   sample_table -> raygen -> [reset -> extend -> shade -> shadow, mis -> resolve]* -> film
// Queue kernels read their work counts from device memory and traversal kernels fetch work in
// wave-sized batches from a device cursor, so an iteration needs no host round trip and the whole
// bounce loop can be captured in a hipGraph.  Each kernel restates one part of the reference
// renderer's per-sample loop (kennyalive/YAR src/ref/reference_renderer.cpp:331-429 render_tile,
// path_tracing.cpp:11-133 trace_path, direct_lighting.cpp:18-578, pixel_sampling.cpp:54-173,
// film.cpp:27-50) with the reference's floating-point operation order, so a path's radiance
// matches the CPU result whenever the device libm agrees with the host's (sin/cos in direction
// sampling are the only libm calls on the path).
#include "kernels.hpp"
#include "shading.hpp"

namespace yg {

// ---------------------------------------------------------------------------------------------
// wave-level helpers
// ---------------------------------------------------------------------------------------------
// Queue appends aggregated over the block: one atomic per block and queue.  Atomics on one address
// serialize in L2 (about 10 ns each), so with one atomic per wave the appends of a 4M-entry launch
// alone would take most of a millisecond.  Every thread of the block must make the call, the same
// number of times (pred may be false).  Entries keep wave order inside the block and lane order
// inside a wave.  Blocks of at most 1024 threads.
template <int NQ>
YG_INLINE void block_append(uint32_t* const (&ctr)[NQ], const bool (&pred)[NQ], uint32_t (&idx)[NQ]) {
    __shared__ uint32_t s_cnt[NQ][16];
    __shared__ uint32_t s_base[NQ];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t pre[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        const uint64_t m = __ballot(pred[k]);
        pre[k] = (uint32_t)__popcll(m & below);
        if (lane == 0) s_cnt[k][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        if (threadIdx.x == (uint32_t)k) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < nw; w++) {
                const uint32_t c = s_cnt[k][w];
                s_cnt[k][w] = tot;
                tot += c;
            }
            s_base[k] = tot ? atomicAdd(ctr[k], tot) : 0u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; k++) idx[k] = s_base[k] + s_cnt[k][wave] + pre[k];
    __syncthreads();
}

// Next `size` work items (a multiple of 64) for this wave from a device cursor (persistent
// traversal kernels).
YG_INLINE uint32_t wave_fetch(uint32_t* cursor, uint32_t size) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(cursor, size);
    return __shfl(base, 0);
}

// Traversal counters per kernel kind: totals[4*kind + 0..3] = nodes, leaf refs, triangle tests,
// objects; kind 0 extend, 1 shadow, 2 mis, 3 trace_rays.
template <bool COUNT>
YG_INLINE void flush_counters(unsigned long long* totals, int kind, const TravCounters& c) {
    if (!COUNT) return;
    unsigned long long v[4] = {c.nodes, c.leaf_refs, c.tri_tests, c.objects};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (__lane_id() == 0 && x) atomicAdd(&totals[4 * kind + k], x);
    }
}

YG_INLINE uint32_t stream_id(int2 px, uint32_t seed_offset) {     // reference_renderer.cpp:348-349
    uint32_t stream = ((uint32_t)px.x & 0xffffu) | ((uint32_t)px.y << 16);
    return stream + seed_offset;
}

// Conservative ray / light-rectangle test: false only when the ray clearly misses the rectangle
// (by more than 2% of its size plus 1e-4 per unit distance plus 1e-5 of the coordinate magnitude),
// where the closest-hit query of the MIS half could not end on the light's triangles anyway; such
// rays need no traversal.  Only used for lights whose emitter triangles were verified at scene
// creation to lie inside their rectangle (size.z == 0).
YG_INLINE bool ray_may_hit_rect(const DevRect& R, V3 o, V3 d) {
    if (R.size.z != 0.f) return true;
    const V3 c = v3(R.l0.w, R.l1.w, R.l2.w);
    const V3 ex = v3(R.l0.x, R.l1.x, R.l2.x), ey = v3(R.l0.y, R.l1.y, R.l2.y), ln = v3(R.l0.z, R.l1.z, R.l2.z);
    const float denom = dot(d, ln);
    if (!(denom < 0.f)) return true;
    const float t = dot(c - o, ln) / denom;
    if (t != t) return true;
    if (!(t > 1e-4f)) return t > -1e-4f;          // plane behind the origin: miss; at the origin: may hit
    const V3 q = (o + t * d) - c;
    const float mag = fmaxf(fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fmaxf(fabsf(o.z), fabsf(c.x))), fmaxf(fabsf(c.y), fabsf(c.z)));
    const float tol = 1e-4f * (1.f + t) + 1e-5f * mag;
    const float mu = 0.51f * R.size.x + tol, mv = 0.51f * R.size.y + tol;
    return fabsf(dot(q, ex)) <= mu && fabsf(dot(q, ey)) <= mv;
}

// Distribution_1D::sample (sampling.cpp:98-147, 168-172) over a CDF of n entries ending in 1:
// returns the sample in [0,1) and the remapped u of the chosen interval.
YG_INLINE float sample_cdf(const float* cdf, int n, float u, float& remapped) {
    int lo = 0, hi = n;                                       // lower_bound: first cdf[k] >= u
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] < u) lo = mid + 1; else hi = mid;
    }
    const int k = min(lo, n - 1);
    const float interval_length = 1.f / (float)n;
    const float a = k == 0 ? 0.f : cdf[k - 1], b = cdf[k];
    const float fraction = (u - a) / (b - a);
    const float x = ((float)k + fraction) * interval_length;
    remapped = fminf(fraction, kOneMinusEps);
    return fminf(x, kOneMinusEps);
}

// Point and emitter normal of a triangle-mesh light at (global triangle, barycentrics), in the mesh's
// coordinates: the interpolated vertex normal (not normalised) when the mesh has normals, else the
// geometric normal (Triangle_Mesh::get_position / get_normal / get_geometric_normal,
// light_sampling.cpp:133-142, 168-173).
YG_INLINE void mesh_light_point(const DevScene& s, uint32_t gt, V3 b, V3& p, V3& n) {
    const V3 p0 = from4(s.tri_verts[3 * (size_t)gt]), p1 = from4(s.tri_verts[3 * (size_t)gt + 1]),
             p2 = from4(s.tri_verts[3 * (size_t)gt + 2]);
    p = b.x * p0 + b.y * p1 + b.z * p2;
    const int4 vi = s.tri_vidx[gt];
    const int mflags = vi.w;
    if (mflags & 1) {
        n = b.x * from4(s.vtx_n[vi.x]) + b.y * from4(s.vtx_n[vi.y]) + b.z * from4(s.vtx_n[vi.z]);
    } else {
        n = normalized(cross(p1 - p0, p2 - p0));
        if (mflags & 4) n = -n;
    }
}

// Diffuse_Sphere_Light_Sampler (light_sampling.cpp:76-121): the cone of directions the sphere
// subtends from the shading point, its frame and the constant pdf of uniform sampling over it.
struct SphereCone { V3 a0, a1, a2; float d_center, cos_max, pdf; };
YG_INLINE SphereCone sphere_cone(const DevSphere& L, V3 pos) {
    SphereCone c;
    const V3 v = pos - from4(L.p);
    c.d_center = length(v);
    c.a2 = vdiv(v, c.d_center);
    coordinate_system(c.a2, c.a0);
    c.a1 = cross(c.a2, c.a0);
    const float sin_max = L.p.w / c.d_center;
    c.cos_max = sqrtf(fmaxf(0.f, 1.f - sin_max * sin_max));
    c.pdf = 1.f / (2.f * kPi * (1.f - c.cos_max));
    return c;
}
// A uniform direction of the cone taken to the near side of the sphere, as a point on the sphere.
YG_INLINE V3 sphere_cone_point(const DevSphere& L, const SphereCone& c, float u0, float u1) {
    const float r = L.p.w, r2 = r * r, d2 = c.d_center * c.d_center;
    const float cos_theta = (1.f - u0) + u0 * c.cos_max;
    const float phi = 2.f * kPi * u1;
    const float sin_theta2 = fmaxf(0.f, 1.f - cos_theta * cos_theta);
    const float d_sample = c.d_center * cos_theta - sqrtf(fmaxf(0.f, r2 - d2 * sin_theta2));
    const float cos_alpha = (d2 + r2 - d_sample * d_sample) / (2.f * c.d_center * r);
    const float sin_alpha = sqrtf(fmaxf(0.f, 1.f - cos_alpha * cos_alpha));
    const V3 dir = ((sin_alpha * cosf(phi)) * c.a0 + (sin_alpha * sinf(phi)) * c.a1) + cos_alpha * c.a2;
    return r * dir + from4(L.p);
}

// get_emitted_radiance (direct_lighting.cpp:388-410): rectangles and spheres emit on both sides,
// triangle-mesh lights only where the original shading normal faces the viewer (not flipped by
// init_shading).
YG_INLINE RGB mesh_light_emission(const DevScene& s, int j, int obj, uint32_t tri, V3 b, V3 ray_d) {
    Shading sc;
    init_shading(s, ray_d, obj, tri, b, sc);
    return sc.flipped ? rgb(0.f, 0.f, 0.f) : rgb4(s.mesh_lights[j].L);
}
YG_INLINE RGB area_emission(const DevScene& s, int light, int obj, uint32_t tri, V3 b, V3 ray_d) {
    if (light < s.n_rects) return rgb4(s.rects[light].L);
    if (light < s.n_rects + s.n_spheres) return rgb4(s.spheres[light - s.n_rects].L);
    return mesh_light_emission(s, light - s.n_rects - s.n_spheres, obj, tri, b, ray_d);
}

// Environment_Light_Sampler over a constant (one-texel) map (light_sampling.cpp:11-73): direction
// to (u, v) in the light frame, the radiance a missed ray collects (bilinear weights of the white
// texel summed in float, image_texture.cpp:441-480, times the radiance) and the solid-angle pdf of
// the one-cell Distribution_2D (pdf_uv = 1).
constexpr float kPi2 = 6.2831853f;                                  // src/lib/math.h:10
constexpr float kPi2Inv = 1.f / kPi2;
// Out of line: inlined at the two miss sites its atan2/acos expansion costs k_shade an occupancy step.
__device__ __noinline__ RGB env_filtered(const DevEnv* __restrict__ env, V3 world_dir) {
    const DevEnv& E = *env;
    const V3 d = xform_vector(M34{E.w2l0, E.w2l1, E.w2l2}, world_dir);
    float phi = atan2f(d.y, d.x);
    phi = phi < 0 ? phi + kPi2 : phi;
    const float theta = acosf(fminf(fmaxf(d.z, -1.f), 1.f));
    const float u = fminf(fmaxf(phi * kPi2Inv, 0.f), kOneMinusEps);
    const float v = fminf(theta * kPiInv, kOneMinusEps);
    const float x = u * 1.f - 0.5f, y = v * 1.f - 0.5f;
    const float wx = x - floorf(x), wy = y - floorf(y);
    const float w = (((1 - wy) * (1 - wx) + (1 - wy) * wx) + wy * (1 - wx)) + wy * wx;
    return rgb(w, w, w) * rgb4(E.L);
}
YG_INLINE float env_pdf(const DevScene& s, V3 world_dir) {
    const DevEnv& E = *s.env;
    const V3 d = xform_vector(M34{E.w2l0, E.w2l1, E.w2l2}, world_dir);
    const float sin_theta = sqrtf(fmaxf(0.f, 1.f - d.z * d.z));
    return 1.f / (2 * kPi * kPi * sin_theta);
}

// Guided form: a pcg32 stream per (pixel, sample), disjoint from the pixel streams (whose upper
// 32 sequence bits are zero), for strategy choices, mixture sampling and fallback dimensions.
YG_INLINE uint64_t guide_stream(const Chunk& c, uint2 pk) {
    const DevView& v = c.views[c.pix_view[pk.x]];
    return ((uint64_t)(pk.y + 1) << 32) | stream_id(c.pixels[pk.x], v.seed_offset);
}
YG_INLINE uint64_t guide_inc(const Chunk& c, uint2 pk) { return (guide_stream(c, pk) << 1u) | 1u; }

YG_INLINE float luminance(RGB c) { return 0.2126729f * c.r + 0.7151522f * c.g + 0.0721750f * c.b; }   // color.cpp:21

YG_INLINE void scene_frame(const DevScene& s, V3& lo, V3& inv_ext) {
    lo = from4(s.scene_lo);
    const V3 e = from4(s.scene_hi) - lo;
    inv_ext = v3(e.x > 0.f ? 1.f / e.x : 0.f, e.y > 0.f ? 1.f / e.y : 0.f, e.z > 0.f ? 1.f / e.z : 0.f);
}

// ---------------------------------------------------------------------------------------------
// Stratified pixel sampler (Stratified_Pixel_Sampler::next_pixel, pixel_sampling.cpp:54-85;
// generate_stratified_sequence_1d/2d, sampling.cpp:79-96; shuffle, math.h:138-144;
// get_next_1d/2d_sample, pixel_sampling.cpp:159-173).
// next_pixel() draws, from the pixel's pcg32 stream in this order: the image plane (spp jittered 2D
// values, row-major strata, x then y), then per 1D dimension spp jittered values and spp-1 shuffle
// draws, then per 2D dimension 2 spp jittered values and spp-1 shuffle draws.  The value at shuffled
// position k of a run is therefore stratum perm[k] jittered by the stream draw at a fixed offset, so
// only the permutations are stored (k_sample_perms); consumers recompute values with the jump-ahead.
// ---------------------------------------------------------------------------------------------
YG_INLINE uint32_t dim1_base(uint32_t spp, uint32_t i) { return 2u * spp + i * (2u * spp - 1u); }
YG_INLINE uint32_t dim2_base(uint32_t spp, uint32_t d1, uint32_t i) {
    return 2u * spp + d1 * (2u * spp - 1u) + i * (3u * spp - 1u);
}
// Draws next_pixel() makes in total (the RNG fallback stream continues after them).
YG_INLINE uint32_t sampler_draws(uint32_t spp, uint32_t d1, uint32_t d2) { return dim2_base(spp, d1, d2); }

// A path's view of its pixel's sampler: seeded stream and first permutation entry.
struct PixelSampler {
    uint64_t s0, inc;
    long long perm0;
    int spp, nx, ny;
};
YG_INLINE PixelSampler pixel_sampler(const Chunk& c, int p) {
    const DevView& v = c.views[c.pix_view[p]];
    PixelSampler ps;
    ps.s0 = pcg_seed_state(stream_id(c.pixels[p], v.seed_offset), ps.inc);
    ps.perm0 = c.pix_first[p] * (long long)(c.d1 + c.d2);
    ps.spp = v.spp; ps.nx = v.nx; ps.ny = v.ny;
    return ps;
}
YG_INLINE uint32_t perm_entry(const Chunk& c, long long e) {
    return c.perm_bytes == 1 ? (uint32_t)c.perm[e] : (uint32_t)reinterpret_cast<const uint16_t*>(c.perm)[e];
}
// generate_stratified_sequence_2d value of stratum j from the draws starting at stream offset n
YG_INLINE void strat2(const PixelSampler& ps, const JumpTables& jt, uint32_t n, uint32_t j, float& vx, float& vy) {
    uint64_t st = pcg_jump(jt, ps.s0, ps.inc, n);
    const float dx = 1.f / float(ps.nx), dy = 1.f / float(ps.ny);
    const uint32_t x = j % (uint32_t)ps.nx, y = j / (uint32_t)ps.nx;
    vx = fminf((float(x) + pcg_float(st, ps.inc)) * dx, kOneMinusEps);
    vy = fminf((float(y) + pcg_float(st, ps.inc)) * dy, kOneMinusEps);
}
// Image-plane sample k (not shuffled).
YG_INLINE void image_sample(const PixelSampler& ps, const JumpTables& jt, int k, float& vx, float& vy) {
    strat2(ps, jt, 2u * (uint32_t)k, (uint32_t)k, vx, vy);
}
// Sample k of 1D dimension i.
YG_INLINE float sample_1d(const Chunk& c, const PixelSampler& ps, int i, int k) {
    const uint32_t spp = (uint32_t)ps.spp;
    const uint32_t j = perm_entry(c, ps.perm0 + (long long)i * spp + k);
    uint64_t st = pcg_jump(c.jt, ps.s0, ps.inc, dim1_base(spp, (uint32_t)i) + j);
    return fminf((float(j) + pcg_float(st, ps.inc)) * (1.f / float(spp)), kOneMinusEps);
}
// Sample k of 2D dimension i.
YG_INLINE void sample_2d(const Chunk& c, const PixelSampler& ps, int i, int k, float& x, float& y) {
    const uint32_t spp = (uint32_t)ps.spp;
    const uint32_t j = perm_entry(c, ps.perm0 + (long long)(c.d1 + i) * spp + k);
    strat2(ps, c.jt, dim2_base(spp, (uint32_t)c.d1, (uint32_t)i) + 2u * j, j, x, y);
}

// Shuffle permutations of large runs (spp > 128): one wave per (pixel, dimension) run.  The lanes
// draw the run's swap indices in parallel (pcg32 jump-ahead to each lane's first draw, then steps
// of 64 draws) into LDS and initialise the index array there; one lane then applies the swaps in
// order (math.h:138-144) and the wave writes the permutation out.  Every array access is in LDS.
template <class T>
__global__ __launch_bounds__(64) void k_sample_perms_large(const Chunk* __restrict__ cp, int n_dims, ulonglong2 step64) {
    extern __shared__ uint16_t big_lds[];
    const Chunk c = *cp;
    const long long total = (long long)c.n_pixels * n_dims;
    const long long g = blockIdx.x;
    if (g >= total) return;
    const int p = (int)(g / n_dims), i = (int)(g % n_dims);
    const DevView& v = c.views[c.pix_view[p]];
    const int spp = v.spp;
    uint16_t* a = big_lds;                       // index array, spp entries
    uint16_t* jv = big_lds + c.max_spp;          // swap index of draw m (array size n = spp - m)
    const int lane = (int)threadIdx.x;
    uint64_t inc;
    const uint64_t s0 = pcg_seed_state(stream_id(c.pixels[p], v.seed_offset), inc);
    const uint32_t first = i < c.d1 ? dim1_base((uint32_t)spp, (uint32_t)i) + (uint32_t)spp
                                    : dim2_base((uint32_t)spp, (uint32_t)c.d1, (uint32_t)(i - c.d1)) + 2u * (uint32_t)spp;
    uint64_t st = pcg_jump(c.jt, s0, inc, first + (uint32_t)lane);
    const uint64_t g64 = step64.y * inc;
    for (int m = lane; m < spp - 1; m += 64) {
        const uint32_t n = (uint32_t)(spp - m);
        jv[m] = (uint16_t)((pcg_output(st) * (uint64_t)n) >> 32);
        st = step64.x * st + g64;
    }
    for (int k = lane; k < spp; k += 64) a[k] = (uint16_t)k;
    __syncthreads();
    if (lane == 0) {
        for (int m = 0; m < spp - 1; m++) {
            const int n = spp - m;
            const uint32_t k = jv[m];
            const uint16_t x = a[n - 1];
            a[n - 1] = a[k];
            a[k] = x;
        }
    }
    __syncthreads();
    T* out = reinterpret_cast<T*>(const_cast<uint8_t*>(c.perm)) + c.pix_first[p] * (long long)n_dims + (long long)i * spp;
    for (int k = lane; k < spp; k += 64) out[k] = (T)a[k];
}

// Shuffle permutations: one work item per (pixel, 1D or 2D dimension) applies the run's Fisher-Yates
// swaps (math.h:138-144, the swap draws follow the run's value draws) to an index array and writes
// the spp resulting stratum indices.  STAGED: the index arrays live in LDS as entries of type T,
// 4 / sizeof(T) consecutive entries of one lane per 32-bit word and words laid out [word][thread], so
// a wave's accesses fall in distinct banks while a 64-entry byte run takes 4 KB per block instead of
// 16 KB (more resident waves to hide the serial draw-and-swap chain); the block's runs, adjacent in
// the table, are written as one contiguous range.  Otherwise the swaps run in place on the output
// (large spp).
template <class T, bool STAGED>
__global__ void k_sample_perms(const Chunk* __restrict__ cp, int n_dims) {
    extern __shared__ uint32_t perm_lds[];
    __shared__ long long run_off[256];
    const Chunk c = *cp;
    T* out = reinterpret_cast<T*>(const_cast<uint8_t*>(c.perm));
    const long long total = (long long)c.n_pixels * n_dims;
    const long long g0 = (long long)blockIdx.x * blockDim.x;
    const long long g = g0 + threadIdx.x;
    const int nt = blockDim.x;
    long long off = 0;
    int spp = 0;
    if (g < total) {
        const int p = (int)(g / n_dims), i = (int)(g % n_dims);
        const DevView& v = c.views[c.pix_view[p]];
        spp = v.spp;
        off = c.pix_first[p] * (long long)n_dims + (long long)i * spp;
        uint64_t inc;
        const uint64_t s0 = pcg_seed_state(stream_id(c.pixels[p], v.seed_offset), inc);
        const uint32_t first = i < c.d1 ? dim1_base((uint32_t)spp, (uint32_t)i) + (uint32_t)spp
                                        : dim2_base((uint32_t)spp, (uint32_t)c.d1, (uint32_t)(i - c.d1)) + 2u * (uint32_t)spp;
        uint64_t st = pcg_jump(c.jt, s0, inc, first);
        if (STAGED) {
            constexpr int E = 4 / (int)sizeof(T);
            T* a = reinterpret_cast<T*>(perm_lds + threadIdx.x);
            const int row = nt * E;                 // entry k at a[(k / E) * row + k % E]
            for (int k = 0; k < spp; k++) a[(k / E) * row + (k % E)] = (T)k;
            for (int n = spp; n > 1; n--) {
                const int k = (int)(((uint64_t)pcg_next(st, inc) * (uint64_t)n) >> 32);
                T* an = a + ((n - 1) / E) * row + ((n - 1) % E);
                T* ak = a + (k / E) * row + (k % E);
                const T x = *an;
                *an = *ak;
                *ak = x;
            }
        } else {
            T* a = out + off;
            for (int k = 0; k < spp; k++) a[k] = (T)k;
            for (int n = spp; n > 1; n--) {
                const int k = (int)(((uint64_t)pcg_next(st, inc) * (uint64_t)n) >> 32);
                const T x = a[n - 1];
                a[n - 1] = a[k];
                a[k] = x;
            }
        }
    }
    if (!STAGED) return;
    if (c.uniform_spp > 0 && (c.uniform_spp & 3) == 0 && sizeof(T) == 1) {
        // every run of the chunk has the same 4-aligned length: each lane copies its own run's LDS
        // words, which already hold four consecutive entries each, to the table
        if (g < total) {
            const uint32_t* a = perm_lds + threadIdx.x;
            uint32_t* o = reinterpret_cast<uint32_t*>(out + off);
            for (int w = 0; w < (spp >> 2); w++) o[w] = a[w * nt];
        }
        return;
    }
    // coalesced write-out of the block's contiguous range [run_off[0], end)
    run_off[threadIdx.x] = g < total ? off : -1;
    __syncthreads();
    const long long base = run_off[0];
    const int last = (int)((total < g0 + nt ? total : g0 + nt) - g0) - 1;
    __shared__ long long end_s;
    if ((int)threadIdx.x == last) end_s = off + spp;
    __syncthreads();
    const int n_out = (int)(end_s - base);
    const int uni = c.uniform_spp;
    for (int j = threadIdx.x; j < n_out; j += nt) {
        int t, k;
        if (uni > 0) {
            t = j / uni;
            k = j - t * uni;
        } else {                                    // the run holding entry j: last t with run_off[t] - base <= j
            int lo = 0, hi = last;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (run_off[mid] - base <= j) lo = mid; else hi = mid - 1;
            }
            t = lo;
            k = (int)(base + j - run_off[t]);
        }
        constexpr int E = 4 / (int)sizeof(T);
        out[base + j] = reinterpret_cast<const T*>(perm_lds)[((k / E) * nt + t) * E + (k % E)];
    }
}

constexpr int kPermBlock = 64;
constexpr int kPermLargeMaxSpp = 8192;     // 2 x 2 B x spp of LDS per run (32 KB)

void launch_sample_perms(const Chunk& hc, const Chunk* dc, hipStream_t s) {
    const int n_dims = hc.d1 + hc.d2;
    if (n_dims == 0) return;
    const long long items = (long long)hc.n_pixels * n_dims;
    const unsigned blocks = (unsigned)((items + kPermBlock - 1) / kPermBlock);
    if (hc.max_spp <= 128) {
        const int per_word = hc.perm_bytes == 1 ? 4 : 2;
        const size_t shm = (size_t)kPermBlock * ((hc.max_spp + per_word - 1) / per_word) * sizeof(uint32_t);
        if (hc.perm_bytes == 1) hipLaunchKernelGGL((k_sample_perms<uint8_t, true>), dim3(blocks), dim3(kPermBlock), shm, s, dc, n_dims);
        else hipLaunchKernelGGL((k_sample_perms<uint16_t, true>), dim3(blocks), dim3(kPermBlock), shm, s, dc, n_dims);
        return;
    }
    if (hc.max_spp > kPermLargeMaxSpp) {          // runs beyond the LDS of one workgroup: in place in HBM
        if (hc.perm_bytes == 1) hipLaunchKernelGGL((k_sample_perms<uint8_t, false>), dim3(blocks), dim3(kPermBlock), 0, s, dc, n_dims);
        else hipLaunchKernelGGL((k_sample_perms<uint16_t, false>), dim3(blocks), dim3(kPermBlock), 0, s, dc, n_dims);
        return;
    }
    // one wave per run; pcg32 multiplier and increment factor of 64 steps (A_64, G_64)
    uint64_t a64 = 1, g64 = 0;
    for (int k = 0; k < 64; k++) { g64 = g64 * kPcgMult + 1; a64 *= kPcgMult; }
    const ulonglong2 step64 = make_ulonglong2(a64, g64);
    const size_t shm = (size_t)2 * hc.max_spp * sizeof(uint16_t);
    const unsigned runs = (unsigned)items;
    if (hc.perm_bytes == 1) hipLaunchKernelGGL((k_sample_perms_large<uint8_t>), dim3(runs), dim3(64), shm, s, dc, n_dims, step64);
    else hipLaunchKernelGGL((k_sample_perms_large<uint16_t>), dim3(runs), dim3(64), shm, s, dc, n_dims, step64);
}

// ---------------------------------------------------------------------------------------------
// Ray generation (render_tile, reference_renderer.cpp:345-365; Camera::generate_ray,
// camera.cpp:16-32).  Parallel mode: all samples of all chunk pixels at once (slot = sample);
// sequential mode: sample `pass` of every pixel (slot = pixel).
// ---------------------------------------------------------------------------------------------
__global__ void k_raygen(const Chunk* __restrict__ cp, Paths P, Queues q) {
    const Chunk c = *cp;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = false;
    int p = 0, k = 0;
    if (s < c.n_paths) {
        if (c.guided) {
            const int S = c.k1 - c.k0;
            p = (int)(s / S);
            k = c.k0 + (int)(s % S);
            active = k < c.views[c.pix_view[p]].spp;
        } else if (c.sequential) {
            p = (int)s;
            k = c.pass;
            active = k < c.views[c.pix_view[p]].spp;
        } else if (c.uniform_spp > 0) {
            p = (int)(s / c.uniform_spp);
            k = (int)(s % c.uniform_spp);
            active = true;
        } else {
            int lo = 0, hi = c.n_pixels - 1;            // pixel owning slot s: pix_first[p] <= s
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (c.pix_first[mid] <= s) lo = mid; else hi = mid - 1;
            }
            p = lo;
            k = (int)(s - c.pix_first[p]);
            active = true;
        }
    }
    if (active) {
        const DevView& v = c.views[c.pix_view[p]];
        const PixelSampler ps = pixel_sampler(c, p);
        const int2 px = c.pixels[p];
        float jx, jy;
        image_sample(ps, c.jt, k, jx, jy);
        const float fx = (float)px.x + jx;
        const float fy = (float)px.y + jy;
        const float u = 2.f * (fx / v.W) - 1.f;
        const float vv = 2.f * (fy / v.H) - 1.f;
        const float right = u * v.horz, up = -vv * v.vert;
        V3 d = v3(right, v.zup ? 1.f : up, v.zup ? up : -1.f);
        d = vdiv(d, length(d));
        const M34 c2w = {v.c0, v.c1, v.c2};
        const V3 o = xform_point(c2w, v3(0.f, 0.f, 0.f));
        const V3 wd = xform_vector(c2w, d);
        P.ray_o[s] = f4(o, 0.f);
        P.ray_d[s] = f4(wd, 0.f);
        P.coeff[s] = make_float4(1.f, 1.f, 1.f, 0.f);
        P.L[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        P.state[s] = 0u;
        P.pk[s] = make_uint2((uint32_t)p, (uint32_t)k);
        if (c.guided) {
            uint64_t ginc;
            P.grng[s] = pcg_seed_state(guide_stream(c, make_uint2((uint32_t)p, (uint32_t)k)), ginc);
        }
        if (c.sequential && c.pass == 0) {
            // fallback stream: the pixel's RNG after next_pixel() drew every table value
            P.rng[s] = pcg_jump(c.jt, ps.s0, ps.inc, sampler_draws((uint32_t)v.spp, (uint32_t)c.d1, (uint32_t)c.d2));
        }
    }
    if (c.uniform_spp > 0 && !c.guided && !c.sequential) {
        // every slot below n_paths is a path: the queue is the identity, no appends
        if (active) q.active0[s] = (uint32_t)s;
        if (s == 0) q.counts[0] = (uint32_t)c.n_paths;
        return;
    }
    uint32_t* const ctr[1] = {&q.counts[0]};
    const bool pred[1] = {active};
    uint32_t idx[1];
    block_append<1>(ctr, pred, idx);
    if (active) q.active0[idx[0]] = (uint32_t)s;
}

// Clears the queue counters and cursors before a pass; the film side-list counter [5] and the
// training-record counter [10] run on.
__global__ void k_zero_counts(uint32_t* counts, int n) {
    if ((int)threadIdx.x < n && threadIdx.x != 5 && threadIdx.x != 10) counts[threadIdx.x] = 0u;
}

void launch_raygen(const Chunk* c, long long n_paths_max, const Paths& p, Queues q, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, q.counts, 16);
    hipLaunchKernelGGL(k_raygen, dim3((unsigned)((n_paths_max + 255) / 256)), dim3(256), 0, s, c, p, q);
}

// ---------------------------------------------------------------------------------------------
// Traversal kernels: extend (closest hit for path rays), shadow (any hit), mis (closest hit for
// BSDF-sampled light rays, direct_lighting.cpp:163-191).  Persistent: each wave fetches 64 rays
// at a time from a device cursor until the queue is drained.
// ---------------------------------------------------------------------------------------------
struct StackCfg { int sd, md, refill; };

YG_INLINE void lds_stacks(uint32_t* lds_base, StackCfg sc, LdsStack& sst, LdsStack& mst) {
    const int T = blockDim.x, t = threadIdx.x;
    uint32_t* snode = lds_base;
    float* stm = reinterpret_cast<float*>(lds_base + sc.sd * T);
    uint32_t* mnode = lds_base + 2 * sc.sd * T;
    float* mtm = reinterpret_cast<float*>(lds_base + 2 * sc.sd * T + sc.md * T);
    sst = LdsStack{snode + t, stm + t, T};
    mst = LdsStack{mnode + t, mtm + t, T};
}

constexpr int kPrivDepth = 32;   // private stack capacity per level (restart beyond)

// Runs `body(sst, sd, mst, md)` with the stacks of the selected storage.
template <bool PRIV, class Body>
YG_INLINE void with_stacks(StackCfg cfg, Body body) {
    if (PRIV) {
        PrivStack<kPrivDepth> sst, mst;
        body(sst, min(cfg.sd, kPrivDepth), mst, min(cfg.md, kPrivDepth));
    } else {
        extern __shared__ uint32_t lds_u[];
        LdsStack sst, mst;
        lds_stacks(lds_u, cfg, sst, mst);
        body(sst, cfg.sd, mst, cfg.md);
    }
}

// Persistent traversal over a device queue.  With `refill` a lane whose query ended takes the next
// queue entry as soon as at least 16 lanes of its wave are idle, so a wave is not held by its
// longest ray (pays on small scenes with short, uneven queries); otherwise waves take batches of 64.
// setup(index, o, d, tmax) returns the payload finish(payload, tracer) receives.
template <bool ANY, bool COUNT, class SStack, class MStack, class Setup, class Finish>
YG_INLINE void trace_queue(const DevScene& sc, uint32_t n, uint32_t* cursor, SStack& sst, int sd, MStack& mst, int md,
                           TravCounters& cnt, bool refill, Setup setup, Finish finish) {
    Tracer<ANY, COUNT> tr;
    if (!refill) {
        // batches of 64: every lane of the wave runs its query to completion before the next batch.
        // The first `size` entries of every wave are assigned statically (no atomic, so a launch
        // over a short queue costs almost nothing); further work comes from the cursor, 4 batches
        // per update while more than 8 fetches per wave remain (fewer same-address atomics), then
        // single batches (balanced tail).
        const uint32_t wpb = (blockDim.x + 63u) >> 6;
        const uint32_t waves = gridDim.x * wpb;
        uint32_t size = n > 8u * 256u * waves ? 256u : 64u;
        const uint32_t static_end = waves * size;
        for (uint32_t base = (blockIdx.x * wpb + (threadIdx.x >> 6)) * size; base < n;
             base = static_end + wave_fetch(cursor, size)) {
            const uint32_t end = min(base + size, n);
            for (uint32_t b = base; b < end; b += 64u) {
                const uint32_t i = b + __lane_id();
                if (i < end) {
                    V3 o, d;
                    float tmax;
                    const uint32_t payload = setup(i, o, d, tmax);
                    if (tr.start(sc, o, d, tmax))
                        while (!tr.step(sc, sst, sd, mst, md, cnt)) {}
                    finish(payload, tr);
                }
            }
            if (n - end < 8u * 256u * waves) size = 64u;
        }
        return;
    }
    bool busy = false, exhausted = false;
    uint32_t payload = 0;
    const uint32_t lane = __lane_id();
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!exhausted && (n_idle >= 16 || n_idle == 64)) {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(cursor, n_idle);
            base = __shfl(base, leader);
            exhausted = base + n_idle >= n;
            if (!busy) {
                const uint32_t idx = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (idx < n) {
                    V3 o, d;
                    float tmax;
                    payload = setup(idx, o, d, tmax);
                    busy = tr.start(sc, o, d, tmax);
                    if (!busy) finish(payload, tr);                  // missed the scene bounds
                }
            }
        }
        if (__ballot(busy) == 0) {
            if (exhausted) break;
            continue;
        }
        if (busy && tr.step(sc, sst, sd, mst, md, cnt)) {
            finish(payload, tr);
            busy = false;
        }
    }
}

// The closest-hit and shadow traversals are latency bound: registers capped for five waves per SIMD
// (96 VGPRs; four waves at the uncapped 105) measured faster on config 2 (k_shadow 28.4 -> 27.0 ms
// per step with no spills, k_extend 74.9 -> 74.3 ms with 28-36 B of spills); k_mis keeps its registers
// (40 B of spills there cost more than the wave gains).  Blocks of at most 256 threads.
#define YG_TRAV_BOUNDS __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))

template <bool COUNT, bool PRIV>
__global__ void YG_TRAV_BOUNDS k_extend(DevScene sc, Paths P, Queues q, int parity, StackCfg cfg) {
    const uint32_t* in = parity ? q.active1 : q.active0;
    const uint32_t n = q.counts[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&q.totals[16], (unsigned long long)n);
        atomicAdd(&q.totals[19], 1ull);
    }
    TravCounters cnt = {0, 0, 0, 0};
    with_stacks<PRIV>(cfg, [&](auto& sst, int sd, auto& mst, int md) {
        trace_queue<false, COUNT>(sc, n, &q.counts[6], sst, sd, mst, md, cnt, cfg.refill != 0,
            [&](uint32_t i, V3& o, V3& d, float& tmax) -> uint32_t {
                const uint32_t slot = in[i];
                o = from4(P.ray_o[slot]);
                d = from4(P.ray_d[slot]);
                tmax = f_inf();
                return slot;
            },
            [&](uint32_t slot, const Tracer<false, COUNT>& tr) {
                const HitRec& h = tr.hit;
                P.hit_tb[slot] = make_float4(h.t, h.b.x, h.b.y, h.b.z);
                P.hit_ot[slot] = make_int2(h.obj, (int)h.tri);
            });
    });
    flush_counters<COUNT>(q.totals, 0, cnt);
}

template <bool COUNT, bool PRIV>
__global__ void YG_TRAV_BOUNDS k_shadow(DevScene sc, Paths P, Queues q, StackCfg cfg) {
    const uint32_t n = q.counts[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&q.totals[17], (unsigned long long)n);
    TravCounters cnt = {0, 0, 0, 0};
    with_stacks<PRIV>(cfg, [&](auto& sst, int sd, auto& mst, int md) {
        trace_queue<true, COUNT>(sc, n, &q.counts[7], sst, sd, mst, md, cnt, cfg.refill != 0,
            [&](uint32_t i, V3& o, V3& d, float& tmax) -> uint32_t {
                const float4 o4 = q.shadow_o[i];
                o = from4(o4);
                d = from4(q.shadow_d[i]);
                tmax = o4.w;
                return i;
            },
            [&](uint32_t i, const Tracer<true, COUNT>& tr) {
                if (!tr.result()) P.dl_light[q.shadow_path[i]] = q.shadow_c[i];
            });
    });
    flush_counters<COUNT>(q.totals, 1, cnt);
}

template <bool COUNT, bool PRIV>
__global__ void __launch_bounds__(256) k_mis(DevScene sc, Paths P, Queues q, StackCfg cfg) {
    const uint32_t n = q.counts[3];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&q.totals[18], (unsigned long long)n);
    TravCounters cnt = {0, 0, 0, 0};
    with_stacks<PRIV>(cfg, [&](auto& sst, int sd, auto& mst, int md) {
        trace_queue<false, COUNT>(sc, n, &q.counts[8], sst, sd, mst, md, cnt, cfg.refill != 0,
            [&](uint32_t i, V3& o, V3& d, float& tmax) -> uint32_t {
                o = from4(q.mis_o[i]);
                d = from4(q.mis_d[i]);
                tmax = f_inf();
                return i;
            },
            [&](uint32_t i, const Tracer<false, COUNT>& tr) {
                const HitRec& h = tr.hit;
                const float4 o4 = q.mis_o[i], d4 = q.mis_d[i];
                const int light = (int)fbits(o4.w);
                if (!tr.result() && light == sc.n_rects + sc.n_spheres + sc.n_mesh_lights) {
                    // environment light: the BSDF sample counts when nothing blocks it (direct_lighting.cpp:355-381)
                    const float4 f4v = q.mis_f[i];
                    const float bpdf = f4v.w;
                    const float w = power_heuristic(bpdf, d4.w);
                    const RGB contrib = (rgb4(sc.env->L) * rgb4(f4v)) * (w * fabsf(q.mis_ndw[i]) / bpdf);
                    P.dl_bsdf[q.mis_path[i]] = f4(contrib, 0.f);
                } else if (tr.result() && sc.obj_info[h.obj].y == light) {
                    float lpdf;
                    RGB Le;
                    if (light < sc.n_rects) {
                        const DevRect L = sc.rects[light];
                        const float4 v0 = sc.tri_verts[3 * (size_t)h.tri], v1 = sc.tri_verts[3 * (size_t)h.tri + 1],
                                     v2 = sc.tri_verts[3 * (size_t)h.tri + 2];
                        // mesh-space hit point, as Triangle_Mesh::get_position (direct_lighting.cpp:181)
                        const V3 p = h.b.x * from4(v0) + h.b.y * from4(v1) + h.b.z * from4(v2);
                        const float dd = length(p - from4(o4));
                        const float lndw = d4.w;
                        lpdf = (dd * dd) / (L.size.x * L.size.y * lndw);
                        Le = rgb4(L.L);
                    } else if (light < sc.n_rects + sc.n_spheres) {
                        lpdf = d4.w;                                // the cone pdf (direct_lighting.cpp:247)
                        Le = rgb4(sc.spheres[light - sc.n_rects].L);
                    } else {
                        // Diffuse_Triangle_Mesh_Light_Sampler::pdf (light_sampling.cpp:163-184); a hit on
                        // the emitter's back has pdf 0 and keeps the full BSDF-sample weight, as there
                        const DevMeshLight ML = sc.mesh_lights[light - sc.n_rects - sc.n_spheres];
                        V3 lp, ln;
                        mesh_light_point(sc, h.tri, h.b, lp, ln);
                        const float lndw = dot(ln, -from4(d4));
                        lpdf = lndw <= 0.f ? 0.f : (h.t * h.t) / (lndw * ML.area);
                        Le = rgb4(ML.L);
                    }
                    const float4 f4v = q.mis_f[i];
                    const float bpdf = f4v.w;
                    const float w = power_heuristic(bpdf, lpdf);
                    const RGB contrib = (Le * rgb4(f4v)) * (w * fabsf(q.mis_ndw[i]) / bpdf);
                    P.dl_bsdf[q.mis_path[i]] = f4(contrib, 0.f);
                }
            });
    });
    flush_counters<COUNT>(q.totals, 2, cnt);
}

// ---------------------------------------------------------------------------------------------
// Shade: one iteration of trace_path's loop body (path_tracing.cpp:22-131) for every active path
// ---------------------------------------------------------------------------------------------
// Guided form: one training record per recorded vertex, weighted by the luminance of the radiance
// its continuation brought in divided by the throughput at the vertex (= f Li cos / p_sample).
YG_INLINE void emit_records(const Paths& P, const Queues& q, uint32_t slot, int nrec, RGB L, uint64_t& rs,
                            uint64_t rinc) {
    for (int r = 0; r < nrec; r++) {
        const float* R = P.rec + ((size_t)slot * kRecMax + r) * kRecFloats;
        const RGB d = L - rgb(R[15], R[16], R[17]);
        const RGB ratio = rgb(R[12] > 0.f ? d.r / R[12] : 0.f, R[13] > 0.f ? d.g / R[13] : 0.f,
                              R[14] > 0.f ? d.b / R[14] : 0.f);
        const float wt = luminance(ratio);
        if (!(wt > 0.f) || !(wt < f_inf())) continue;
        if (pcg_float(rs, rinc) >= q.keep_prob) continue;
        // ring buffer over the whole render: training sees every pass's records, newest kept
        const uint32_t idx = atomicAdd(&q.counts[10], 1u) % q.train_cap;
        GuideRecord g;
        for (int j = 0; j < 9; j++) g.x[j] = R[j];
        g.wt = wt;
        g.w[0] = R[9]; g.w[1] = R[10]; g.w[2] = R[11];
        g.pad[0] = g.pad[1] = g.pad[2] = 0.f;
        q.train[idx] = g;
    }
}

YG_INLINE void finalize(const Chunk& c, const Paths& P, uint2 pk, RGB L) {          // render_tile :386-393
    float mc = fmaxf(L.r, fmaxf(L.g, L.b));
    if (mc > c.max_component) L = L * (c.max_component / mc);
    if (c.film_scale != 1.f) L = L * c.film_scale;
    P.radiance[c.pix_first[pk.x] + pk.y] = f4(L, 0.f);
}

// GUIDED selects the guided form's variant at compile time, so the plain kernel keeps its own
// (smaller) register allocation.
template <bool GUIDED, bool MIX>
__global__ void __launch_bounds__(256) k_shade(DevScene s, const Chunk* __restrict__ cp, Paths P, Queues q, int parity) {
    const Chunk& c = *cp;
    const uint32_t* in = parity ? q.active1 : q.active0;
    uint32_t* out = q.bin_rays ? q.unsorted : (parity ? q.active0 : q.active1);
    uint32_t* n_out = &q.counts[parity ^ 1];
    const uint32_t n = q.counts[parity];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t blk = blockIdx.x * blockDim.x; blk < n; blk += stride) {
        const uint32_t i = blk + threadIdx.x;
        const bool valid = i < n;
        bool want_active = false, want_shadow = false, want_mis = false, want_dl = false, want_guide = false;
        bool mis_culled = false;
        float4 g_x0, g_x1, g_x2, g_x3;
        uint32_t slot = 0;
        V3 sh_o = v3(0, 0, 0), sh_d = v3(0, 0, 0), mi_o = v3(0, 0, 0), mi_d = v3(0, 0, 0);
        float sh_tmax = 0.f, mi_pdf = 0.f, mi_ndw = 0.f, mi_lndw = 0.f;
        RGB sh_c = rgb(0, 0, 0), mi_f = rgb(0, 0, 0);
        int mi_light = 0;
        if (valid) {
            slot = in[i];
            uint32_t st = P.state[slot];
            int bounce = (int)(st & 0xffu), c1 = (int)((st >> 8) & 0xffu), c2 = (int)((st >> 16) & 0xffu);
            const bool delta_traced = (st & kDelta) != 0;
            RGB coeff = rgb4(P.coeff[slot]);
            RGB L = rgb4(P.L[slot]);
            const float4 d4 = P.ray_d[slot];
            const int2 ot = P.hit_ot[slot];
            const float4 tb = P.hit_tb[slot];
            const bool hit = ot.x >= 0;
            const int light = hit ? s.obj_info[ot.x].y : -1;
            const uint2 pk = P.pk[slot];
            const int p = (int)pk.x, k = (int)pk.y;
            const DevView& v = c.views[c.pix_view[p]];
            const PixelSampler ps = pixel_sampler(c, p);
            const bool seq = c.sequential != 0;
            uint64_t rs = 0, rinc = 0;
            if (GUIDED && c.guided) {
                rinc = guide_inc(c, pk);
                rs = P.grng[slot];
            } else if (seq) {
                rinc = ((uint64_t)stream_id(c.pixels[p], v.seed_offset) << 1u) | 1u;
                rs = P.rng[slot];
            }
            int nrec = (int)((st >> kRecShift) & 7u);
            auto next1 = [&]() -> float {                               // pixel_sampling.cpp:159
                if (c1 < c.d1) { const float x = sample_1d(c, ps, c1, k); c1++; return x; }
                return pcg_float(rs, rinc);
            };
            auto next2 = [&](float& x, float& y) {                      // pixel_sampling.cpp:167
                if (c2 < c.d2) {
                    sample_2d(c, ps, c2, k, x, y);
                    c2++;
                    return;
                }
                x = pcg_float(rs, rinc);
                y = pcg_float(rs, rinc);
            };
            bool finished = false, pending_dl = false, continue_body = true;
            if (!delta_traced) {
                if (bounce == 0) {
                    if (hit && light >= 0) L = L + coeff * area_emission(s, light, ot.x, (uint32_t)ot.y, v3(tb.y, tb.z, tb.w), from4(d4));
                    else if (!hit && s.n_env) L = L + coeff * env_filtered(s.env, from4(d4));
                    if (c.max_bounces == 0) { finished = true; continue_body = false; }
                }
                if (continue_body && (!hit || light >= 0)) { finished = true; continue_body = false; }
            } else {
                RGB emitted = rgb(0.f, 0.f, 0.f);
                if (hit && light >= 0) emitted = area_emission(s, light, ot.x, (uint32_t)ot.y, v3(tb.y, tb.z, tb.w), from4(d4));
                else if (!hit && s.n_env) emitted = env_filtered(s.env, from4(d4));
                L = L + coeff * emitted;
                bounce++;
                if (bounce == c.max_bounces || !hit || light >= 0) { finished = true; continue_body = false; }
                else if (bounce >= c.rr_bounce_threshold) {            // path_tracing.cpp:112-130
                    float u_term = next1();
                    float mc = fmaxf(coeff.r, fmaxf(coeff.g, coeff.b));
                    if (mc < c.rr_threshold) {
                        float pt = fmaxf(0.05f, 1.f - mc);
                        if (u_term < pt) { finished = true; continue_body = false; }
                        else coeff = coeff * (1.f / (1.f - pt));
                    }
                }
            }
            uint32_t flags = 0;
            if (continue_body) {
                Shading sc;
                init_shading(s, from4(d4), ot.x, (uint32_t)ot.y, v3(tb.y, tb.z, tb.w), sc);
                const float u_scat = next1();
                const float u_sel = next1();
                const float u_scat_next = next1();                      // u_scattering_type_next_segment
                float ulx, uly, ubx, uby, unx, uny;
                next2(ulx, uly);
                next2(ubx, uby);
                next2(unx, uny);
                const DevMaterial m = s.materials[sc.material];
                // initialize_scattering (shading_context.cpp:174-186): the uber material picks one of
                // its components with u_scat and hands the rescaled u_scat on; the BSDF continuation
                // is divided by the probability of its BSDF layers (path_tracing.cpp:68)
                float u_s = u_scat, bsdf_layer_probability = 1.f;
                int uber_delta = 0;                                     // 1 delta reflection, 2 pass-through
                bool mix_delta = false;
                if (MIX && m.type == kMixType) {
                    // get_mix_info (delta_scattering.cpp:154-189): u_scat rescaled to the picked
                    // sub-material's interval; picking a delta sub-material (mirror, smooth glass) is a
                    // delta event, and the BSDF layer's probability is the other one's weight
                    const float mix_amount = m.trans.z;
                    const bool first = u_scat < mix_amount;
                    u_s = first ? u_scat / mix_amount : (u_scat - mix_amount) / (1.f - mix_amount);
                    if (m.eta.x >= 0.f) {
                        const bool delta_first = m.eta.x == 0.f;
                        mix_delta = first == delta_first;
                        bsdf_layer_probability = 1.f - (delta_first ? mix_amount : 1.f - mix_amount);
                    }
                }
                if (m.type == kUberType) {
                    const uint32_t code = (uint32_t)m.trans.w;
                    const uint32_t ncomp = code & 15u;
                    const float fp_index = u_scat * (float)ncomp;
                    const int ci = (int)fp_index;
                    const uint32_t comp = (code >> (4 + 4 * ci)) & 15u;
                    u_s = fminf(fmaxf(fp_index - (float)ci, 0.f), kOneMinusEps);
                    uber_delta = comp == 2u ? 1 : (comp == 4u ? 2 : 0);
                    bsdf_layer_probability = 1.f - (float)((code >> 20) & 15u) / (float)ncomp;
                }
                if (m.type == 4) {
                    // perfect reflector (delta_scattering.cpp:27-33)
                    const V3 dir = reflect(sc.wo, sc.n);
                    P.ray_o[slot] = f4(origin_dir(sc, dir), 0.f);
                    P.ray_d[slot] = f4(dir, 0.f);
                    coeff = coeff * rgb4(m.refl);
                    flags |= kDelta;
                    want_active = true;
                } else if (m.type == 3) {
                    // smooth glass, get_glass_info (delta_scattering.cpp:53-100); enter/exit from the
                    // original shading-normal orientation (no nested-dielectric tracking in pbrt scenes)
                    const float ior = m.trans.w;
                    const bool enter = !sc.flipped;
                    const float etaT_over_etaI = enter ? ior : 1.f / ior;
                    const float cos_i = dot(sc.n, sc.wo);
                    const float F = dielectric_fresnel(cos_i, etaT_over_etaI);
                    RGB att;
                    V3 dir;
                    if (u_scat < F) {
                        att = rgb4(m.refl);
                        dir = reflect(sc.wo, sc.n);
                    } else {
                        const float eta = 1.f / etaT_over_etaI;
                        att = rgb4(m.trans);
                        att = att * (eta * eta);
                        const float ci = dot(sc.wo, sc.n);              // refract, math.h:115
                        const float st2 = eta * eta * fmaxf(0.f, 1.f - ci * ci);
                        const float ct = sqrtf(1.f - st2);
                        dir = (-eta) * sc.wo + (eta * ci - ct) * sc.n;
                    }
                    P.ray_o[slot] = f4(origin_dir(sc, dir), 0.f);
                    P.ray_d[slot] = f4(dir, 0.f);
                    coeff = coeff * att;
                    flags |= kDelta;
                    want_active = true;
                } else {
                    if (bsdf_layer_probability != 0.f) {
                        // estimate_direct_lighting_from_single_sample (direct_lighting.cpp:522-578)
                        const int li = int(u_sel * s.n_lights);
                        pending_dl = true;
                        int lj = li;
                        if (lj < s.n_points) {
                            const DevPoint pl = s.points[lj];               // direct_lighting.cpp:26-47
                            const V3 lp = from4(pl.p);
                            const V3 position = origin_dir(sc, lp - sc.pos);
                            const V3 lv = lp - position;
                            const float dist = length(lv);
                            const V3 ldir = vdiv(lv, dist);
                            const float ndl = dot(sc.n, ldir);
                            if (ndl > 0.f) {
                                const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, ldir, s.materials);
                                sh_c = (rgb4(pl.I) * f) * (ndl / (dist * dist));
                                sh_o = position; sh_d = ldir; sh_tmax = dist * (1.f - 1e-5f);
                                want_shadow = true;
                            }
                        } else if ((lj -= s.n_points) < s.n_spots) {
                            const DevSpot sl = s.spots[lj];                 // direct_lighting.cpp:49-89
                            const V3 lp = from4(sl.p);
                            const V3 position = origin_dir(sc, lp - sc.pos);
                            const V3 lv = lp - position;
                            const float dist = length(lv);
                            const V3 wi = vdiv(lv, dist);
                            const float cone_cos = sl.dir.w;
                            const float wi_cos = dot(-wi, from4(sl.dir));
                            if (!(wi_cos < cone_cos)) {
                                float penumbra = 1.f;
                                const float penumbra_cos = sl.I.w;
                                if (wi_cos < penumbra_cos) {
                                    const float kp = (wi_cos - cone_cos) / (penumbra_cos - cone_cos);
                                    penumbra = (kp * kp) * (kp * kp);
                                }
                                const float ndw = dot(sc.n, wi);
                                if (scattering_possible(m, ndw)) {
                                    const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, wi, s.materials);
                                    if (!is_black(f)) {
                                        sh_c = (rgb4(sl.I) * f) * (penumbra * fabsf(ndw) / (dist * dist));
                                        sh_o = position; sh_d = wi; sh_tmax = dist * (1.f - 1e-5f);
                                        want_shadow = true;
                                    }
                                }
                            }
                        } else if ((lj -= s.n_spots) < s.n_dirs) {
                            const DevDirLight dl = s.dirs[lj];              // direct_lighting.cpp:91-107
                            const V3 ldir = from4(dl.dir);
                            const float ndl = dot(sc.n, ldir);
                            if (ndl > 0.f) {
                                const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, ldir, s.materials);
                                sh_c = (rgb4(dl.E) * f) * ndl;
                                sh_o = origin_dir(sc, ldir); sh_d = ldir; sh_tmax = f_inf();
                                want_shadow = true;
                            }
                        } else if ((lj -= s.n_dirs) < s.n_rects) {
                            const int ri = lj;                              // direct_lighting.cpp:109-194
                            const DevRect R = s.rects[ri];
                            const M34 l2w = {R.l0, R.l1, R.l2};
                            const V3 ln = v3(R.l0.z, R.l1.z, R.l2.z);
                            {
                                const V3 local = v3(R.size.x * (ulx - 0.5f), R.size.y * (uly - 0.5f), 0.0f);
                                const V3 lp = xform_point(l2w, local);
                                const V3 position = origin_dir(sc, lp - sc.pos);
                                const V3 lv = lp - position;
                                const float dd = length(lv);
                                const V3 wi = vdiv(lv, dd);
                                const float lndw = dot(ln, -wi);
                                if (lndw > 1e-4f) {
                                    const float ndw = dot(sc.n, wi);
                                    if (scattering_possible(m, ndw)) {
                                        const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, wi, s.materials);
                                        if (!is_black(f)) {
                                            const float lpdf = (dd * dd) / (R.size.x * R.size.y * lndw);
                                            const float bpdf = bsdf_pdf<MIX>(m, sc, sc.wo, wi, s.materials);
                                            const float w = power_heuristic(lpdf, bpdf);
                                            sh_c = (rgb4(R.L) * f) * (w * fabsf(ndw) / lpdf);
                                            sh_o = position; sh_d = wi; sh_tmax = dd * (1.f - 1e-5f);
                                            want_shadow = true;
                                        }
                                    }
                                }
                            }
                            {
                                V3 wi = v3(0.f, 0.f, 0.f);
                                float bpdf = 0.f;
                                const RGB f = bsdf_sample<MIX>(m, sc, ubx, uby, u_s, sc.wo, wi, bpdf, s.materials);
                                if (!is_black(f)) {
                                    const float lndw = dot(ln, -wi);
                                    if (lndw > 1e-4f) {
                                        mi_o = origin_dir(sc, wi); mi_d = wi; mi_f = f; mi_pdf = bpdf;
                                        mi_ndw = dot(sc.n, wi); mi_lndw = lndw; mi_light = ri;
                                        // only a ray that can reach the light's rectangle can contribute
                                        want_mis = ray_may_hit_rect(R, mi_o, wi);
                                        mis_culled = !want_mis;
                                    }
                                }
                            }
                        } else if ((lj -= s.n_rects) < s.n_spheres) {
                            const DevSphere SL = s.spheres[lj];             // direct_lighting.cpp:196-260
                            const SphereCone cone = sphere_cone(SL, sc.pos);
                            {
                                const V3 lp = sphere_cone_point(SL, cone, ulx, uly);
                                const V3 position = origin_dir(sc, lp - sc.pos);
                                const V3 lv = lp - position;
                                const float dd = length(lv);
                                const V3 wi = vdiv(lv, dd);
                                const float ndw = dot(sc.n, wi);
                                if (scattering_possible(m, ndw)) {
                                    const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, wi, s.materials);
                                    if (!is_black(f)) {
                                        const float bpdf = bsdf_pdf<MIX>(m, sc, sc.wo, wi, s.materials);
                                        const float w = power_heuristic(cone.pdf, bpdf);
                                        sh_c = (rgb4(SL.L) * f) * (w * fabsf(ndw) / cone.pdf);
                                        sh_o = position; sh_d = wi; sh_tmax = dd * (1.f - 1e-5f);
                                        want_shadow = true;
                                    }
                                }
                            }
                            {
                                V3 wi = v3(0.f, 0.f, 0.f);
                                float bpdf = 0.f;
                                const RGB f = bsdf_sample<MIX>(m, sc, ubx, uby, u_s, sc.wo, wi, bpdf, s.materials);
                                if (!is_black(f) && dot(-cone.a2, wi) >= cone.cos_max) {
                                    mi_o = origin_dir(sc, wi); mi_d = wi; mi_f = f; mi_pdf = bpdf;
                                    mi_ndw = dot(sc.n, wi); mi_lndw = cone.pdf; mi_light = s.n_rects + lj;
                                    want_mis = true;
                                }
                            }
                        } else if ((lj -= s.n_spheres) < s.n_mesh_lights) {
                            const DevMeshLight ML = s.mesh_lights[lj];      // direct_lighting.cpp:262-322
                            {
                                // Diffuse_Triangle_Mesh_Light_Sampler::sample (light_sampling.cpp:126-161)
                                float u0;
                                const float smp = sample_cdf(s.mesh_light_cdf + ML.cdf, ML.tri_count, ulx, u0);
                                const uint32_t lt = min((uint32_t)(smp * (float)ML.tri_count), (uint32_t)ML.tri_count - 1u);
                                const float ra = sqrtf(u0);                  // sampling.cpp:69-77
                                V3 bl;
                                bl.x = 1 - ra;
                                bl.y = uly * ra;
                                bl.z = 1.f - bl.x - bl.y;
                                V3 lp, ln;
                                mesh_light_point(s, (uint32_t)ML.tri_base + lt, bl, lp, ln);
                                const V3 lv0 = lp - sc.pos;
                                const float dsq = dot(lv0, lv0);
                                const V3 w0 = vdiv(lv0, sqrtf(dsq));
                                const float lndw0 = dot(ln, -w0);
                                const float lpdf = lndw0 <= 0.f ? 0.f : dsq / (lndw0 * ML.area);
                                if (lpdf != 0.f) {
                                    const V3 position = origin_dir(sc, lp - sc.pos);
                                    const V3 lv = lp - position;
                                    const float dd = length(lv);
                                    const V3 wi = vdiv(lv, dd);
                                    const float ndw = dot(sc.n, wi);
                                    if (scattering_possible(m, ndw)) {
                                        const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, wi, s.materials);
                                        if (!is_black(f)) {
                                            const float bpdf = bsdf_pdf<MIX>(m, sc, sc.wo, wi, s.materials);
                                            const float w = power_heuristic(lpdf, bpdf);
                                            sh_c = (rgb4(ML.L) * f) * (w * fabsf(ndw) / lpdf);
                                            sh_o = position; sh_d = wi; sh_tmax = dd * (1.f - 1e-5f);
                                            want_shadow = true;
                                        }
                                    }
                                }
                            }
                            {
                                V3 wi = v3(0.f, 0.f, 0.f);
                                float bpdf = 0.f;
                                const RGB f = bsdf_sample<MIX>(m, sc, ubx, uby, u_s, sc.wo, wi, bpdf, s.materials);
                                if (!is_black(f)) {
                                    mi_o = origin_dir(sc, wi); mi_d = wi; mi_f = f; mi_pdf = bpdf;
                                    mi_ndw = dot(sc.n, wi); mi_lndw = 0.f; mi_light = s.n_rects + s.n_spheres + lj;
                                    want_mis = true;
                                }
                            }
                        } else {
                            const RGB Le = rgb4(s.env->L);                  // direct_lighting.cpp:324-384
                            {
                                // the one-cell Distribution_2D samples uv = (u1, u0) with pdf_uv = 1
                                const float phi = uly * kPi2, theta = ulx * kPi;
                                const float sin_theta = sinf(theta);
                                const float lpdf = 1.f / (2 * kPi * kPi * sin_theta);
                                const DevEnv& E = *s.env;
                                const V3 wi = xform_vector(M34{E.l2w0, E.l2w1, E.l2w2},
                                                           v3(sin_theta * cosf(phi), sin_theta * sinf(phi), cosf(theta)));
                                const float ndw = dot(sc.n, wi);
                                if (scattering_possible(m, ndw)) {
                                    const RGB f = bsdf_eval<MIX>(m, sc, sc.wo, wi, s.materials);
                                    if (!is_black(f)) {
                                        const float bpdf = bsdf_pdf<MIX>(m, sc, sc.wo, wi, s.materials);
                                        const float w = power_heuristic(lpdf, bpdf);
                                        sh_c = (Le * f) * (w * fabsf(ndw) / lpdf);
                                        sh_o = origin_dir(sc, wi); sh_d = wi; sh_tmax = f_inf();
                                        want_shadow = true;
                                    }
                                }
                            }
                            {
                                V3 wi = v3(0.f, 0.f, 0.f);
                                float bpdf = 0.f;
                                const RGB f = bsdf_sample<MIX>(m, sc, ubx, uby, u_s, sc.wo, wi, bpdf, s.materials);
                                if (!is_black(f) && !is_black(Le)) {
                                    mi_o = origin_dir(sc, wi); mi_d = wi; mi_f = f; mi_pdf = bpdf;
                                    mi_ndw = dot(sc.n, wi); mi_lndw = env_pdf(s, wi);
                                    mi_light = s.n_rects + s.n_spheres + s.n_mesh_lights;
                                    want_mis = true;
                                }
                            }
                        }
                        P.dl_coeff[slot] = f4(coeff, 0.f);
                        P.dl_light[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                        P.dl_bsdf[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                    }
                    if (uber_delta != 0) {
                        // get_pbrt_uber_info (delta_scattering.cpp:102-152): Fresnel-weighted mirror
                        // reflection of Kr, or the opacity pass-through, each scaled by the component count
                        const float ncomp = (float)((uint32_t)m.trans.w & 15u);
                        RGB att;
                        V3 dir;
                        if (uber_delta == 1) {
                            const float F = dielectric_fresnel(dot(sc.n, sc.wo), sc.flipped ? 1.f / m.refl.w : m.refl.w);
                            att = rgb4(m.eta) * (ncomp * F);
                            dir = reflect(sc.wo, sc.n);
                        } else {
                            att = (rgb(1.f, 1.f, 1.f) - rgb4(m.k)) * ncomp;
                            dir = -sc.wo;
                        }
                        P.ray_o[slot] = f4(origin_dir(sc, dir), 0.f);
                        P.ray_d[slot] = f4(dir, 0.f);
                        coeff = coeff * att;
                        flags |= kDelta;
                        want_active = true;
                    } else if (MIX && mix_delta) {
                        // the mix's delta sub-material with the rescaled u_scat: perfect reflector
                        // (delta_scattering.cpp:27-33) or smooth glass (get_glass_info, :53-100)
                        const DevMaterial dm = s.materials[(int)(m.eta.x == 0.f ? m.trans.x : m.trans.y)];
                        V3 dir = reflect(sc.wo, sc.n);
                        RGB att = rgb4(dm.refl);
                        if (dm.type == 3) {
                            const float ior = dm.trans.w;
                            const float etaT_over_etaI = !sc.flipped ? ior : 1.f / ior;
                            const float cos_i = dot(sc.n, sc.wo);
                            if (!(u_s < dielectric_fresnel(cos_i, etaT_over_etaI))) {
                                const float eta = 1.f / etaT_over_etaI;
                                att = rgb4(dm.trans) * (eta * eta);
                                const float st2 = eta * eta * fmaxf(0.f, 1.f - cos_i * cos_i);
                                const float ct = sqrtf(1.f - st2);
                                dir = (-eta) * sc.wo + (eta * cos_i - ct) * sc.n;
                            }
                        }
                        P.ray_o[slot] = f4(origin_dir(sc, dir), 0.f);
                        P.ray_d[slot] = f4(dir, 0.f);
                        coeff = coeff * att;
                        flags |= kDelta;
                        want_active = true;
                    } else {
                    bounce++;
                    if (bounce == c.max_bounces) {
                        finished = true;
                    } else if (GUIDED && c.guided && m.type == 1) {
                        // guided continuation at a diffuse vertex: record the vertex, then either
                        // sample the BSDF here (untrained pass) or queue a network query for the
                        // guide-sample stage.  Glossy vertices keep BSDF sampling: their product
                        // f.Li.cos is a narrow lobe that the mixture cannot match, and a broad
                        // proposal mixed into a glossy lobe only raises the variance.
                        float x[kGIn];
                        V3 lo, inv_ext;
                        scene_frame(s, lo, inv_ext);
                        guide_inputs(sc.pos, lo, inv_ext, sc.wo, sc.n, x);
                        int ridx = -1;
                        if (nrec < kRecMax) {
                            ridx = nrec++;
                            float* R = P.rec + ((size_t)slot * kRecMax + ridx) * kRecFloats;
                            for (int j = 0; j < kGIn; j++) R[j] = x[j];
                            R[9] = R[10] = R[11] = 0.f;
                            R[12] = coeff.r; R[13] = coeff.g; R[14] = coeff.b;
                            flags |= kRecPending;
                        }
                        const float u_term = bounce >= c.rr_bounce_threshold ? next1() : -1.f;
                        if (c.alpha > 0.f) {
                            want_guide = true;
                            g_x0 = make_float4(x[0], x[1], x[2], x[3]);
                            g_x1 = make_float4(x[4], x[5], x[6], x[7]);
                            g_x2 = make_float4(x[8], (float)ridx, unx, uny);
                            g_x3 = make_float4(u_term, u_scat_next, 0.f, 0.f);
                        } else {
                            V3 wi = v3(0.f, 0.f, 0.f);
                            float pdf = 0.f;
                            const RGB f = bsdf_sample<MIX>(m, sc, unx, uny, u_scat_next, sc.wo, wi, pdf, s.materials);
                            if (is_black(f)) {
                                finished = true;
                            } else {
                                coeff = coeff * (f * (fabsf(dot(sc.n, wi)) / pdf));
                                P.ray_o[slot] = f4(origin_dir(sc, wi), 0.f);
                                P.ray_d[slot] = f4(wi, 0.f);
                                if (ridx >= 0) {
                                    float* R = P.rec + ((size_t)slot * kRecMax + ridx) * kRecFloats;
                                    R[9] = wi.x; R[10] = wi.y; R[11] = wi.z;
                                }
                                if (u_term >= 0.f) {
                                    const float mc = fmaxf(coeff.r, fmaxf(coeff.g, coeff.b));
                                    if (mc < c.rr_threshold) {
                                        const float pt = fmaxf(0.05f, 1.f - mc);
                                        if (u_term < pt) finished = true;
                                        else coeff = coeff * (1.f / (1.f - pt));
                                    }
                                }
                                if (!finished) want_active = true;
                            }
                        }
                    } else {
                        V3 wi = v3(0.f, 0.f, 0.f);
                        float pdf = 0.f;
                        const RGB f = bsdf_sample<MIX>(m, sc, unx, uny, u_scat_next, sc.wo, wi, pdf, s.materials);
                        if (is_black(f)) {
                            finished = true;
                        } else {
                            coeff = coeff * (f * (fabsf(dot(sc.n, wi)) / (pdf * bsdf_layer_probability)));
                            P.ray_o[slot] = f4(origin_dir(sc, wi), 0.f);
                            P.ray_d[slot] = f4(wi, 0.f);
                            if (bounce >= c.rr_bounce_threshold) {
                                float u_term = next1();
                                float mc = fmaxf(coeff.r, fmaxf(coeff.g, coeff.b));
                                if (mc < c.rr_threshold) {
                                    float pt = fmaxf(0.05f, 1.f - mc);
                                    if (u_term < pt) finished = true;
                                    else coeff = coeff * (1.f / (1.f - pt));
                                }
                            }
                            if (!finished) want_active = true;
                        }
                    }
                    }
                }
            }
            if (pending_dl) { flags |= kPendingDL; want_dl = true; if (finished) flags |= kDone; }
            else if (finished) {
                if (GUIDED && c.guided) emit_records(P, q, slot, nrec, L, rs, rinc);
                finalize(c, P, pk, L);
            }
            P.state[slot] = (uint32_t)bounce | ((uint32_t)c1 << 8) | ((uint32_t)c2 << 16) | flags |
                            ((uint32_t)nrec << kRecShift);
            P.coeff[slot] = f4(coeff, 0.f);
            P.L[slot] = f4(L, 0.f);
            if (GUIDED && c.guided) P.grng[slot] = rs;
            else if (seq) P.rng[slot] = rs;
        }
        if (q.count_culls) {
            const uint64_t m = __ballot(mis_culled);
            if (__lane_id() == 0 && m) atomicAdd(&q.totals[20], (unsigned long long)__popcll(m));
        }
        uint32_t* const ctr[4] = {n_out, &q.counts[2], &q.counts[3], &q.counts[4]};
        const bool pred[4] = {want_active, want_shadow, want_mis, want_dl};
        uint32_t qi[4];
        block_append<4>(ctr, pred, qi);
        const uint32_t ia = qi[0], is = qi[1], im = qi[2], id = qi[3];
        if (want_active) out[ia] = slot;
        if (want_shadow) {
            q.shadow_path[is] = slot;
            q.shadow_o[is] = f4(sh_o, sh_tmax);
            q.shadow_d[is] = f4(sh_d, 0.f);
            q.shadow_c[is] = f4(sh_c, 0.f);
        }
        if (want_mis) {
            q.mis_path[im] = slot;
            q.mis_o[im] = f4(mi_o, bitsf((uint32_t)mi_light));
            q.mis_d[im] = f4(mi_d, mi_lndw);
            q.mis_f[im] = f4(mi_f, mi_pdf);
            q.mis_ndw[im] = mi_ndw;
        }
        if (want_dl) q.dl_path[id] = slot;
        if (GUIDED && c.guided) {
            uint32_t* const gctr[1] = {&q.counts[9]};
            const bool gpred[1] = {want_guide};
            uint32_t gidx[1];
            block_append<1>(gctr, gpred, gidx);
            const uint32_t ig = gidx[0];
            if (want_guide) {
                q.guide_path[ig] = slot;
                float4* gi = reinterpret_cast<float4*>(q.guide_in + (size_t)ig * 16);
                gi[0] = g_x0; gi[1] = g_x1; gi[2] = g_x2; gi[3] = g_x3;
            }
        }
    }
}

// Guided continuation (one-sample MIS of the learned mixture and the BSDF, Veach's balance-weight
// one-sample model): with probability alpha the direction comes from the mixture, otherwise from
// the BSDF with the path's own sampler dimensions; the weight is f |n.w| / (alpha q + (1-alpha) p_bsdf)
// either way, so the estimator stays unbiased for any network output.
__global__ void k_guide_sample(DevScene s, const Chunk* __restrict__ cp, Paths P, Queues q, int parity) {
    const Chunk c = *cp;
    uint32_t* out = q.bin_rays ? q.unsorted : (parity ? q.active0 : q.active1);
    uint32_t* n_out = &q.counts[parity ^ 1];
    const uint32_t n = q.counts[9];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t blk = blockIdx.x * blockDim.x; blk < n; blk += stride) {
        const uint32_t i = blk + threadIdx.x;
        bool alive = false;
        uint32_t slot = 0;
        if (i < n) {
            slot = q.guide_path[i];
            const float4* gi = reinterpret_cast<const float4*>(q.guide_in + (size_t)i * 16);
            const float4 g2 = gi[2];
            const int ridx = (int)g2.y;
            const float4 g3 = gi[3];
            const float unx = g2.z, uny = g2.w, u_term = g3.x, u_scat_next = g3.y;
            const uint2 pk = P.pk[slot];
            const uint64_t rinc = guide_inc(c, pk);
            uint64_t rs = P.grng[slot];
            const int2 ot = P.hit_ot[slot];
            const float4 tb = P.hit_tb[slot];
            Shading sc;
            init_shading(s, from4(P.ray_d[slot]), ot.x, (uint32_t)ot.y, v3(tb.y, tb.z, tb.w), sc);
            const DevMaterial m = s.materials[sc.material];
            VmfLobe lobes[kGLobes];
            mixture_from_outputs(q.guide_out + (size_t)i * kGOut, lobes);
            const float a = c.alpha;
            V3 wi = v3(0.f, 0.f, 0.f);
            RGB f = rgb(0.f, 0.f, 0.f);
            float pb = 0.f;
            if (pcg_float(rs, rinc) < a) {
                const float u0 = pcg_float(rs, rinc), u1 = pcg_float(rs, rinc), u2 = pcg_float(rs, rinc);
                wi = mixture_sample(lobes, u0, u1, u2);
                const bool upper = dot(sc.n, wi) > 0.f && (m.type == 1 || to_local(sc, wi).z > 0.f);
                if (upper) {
                    f = bsdf_eval_base(m, sc, sc.wo, wi);
                    pb = bsdf_pdf_base(m, sc, sc.wo, wi);
                }
            } else {
                f = bsdf_sample_base(m, sc, unx, uny, u_scat_next, sc.wo, wi, pb);
                if (!(dot(sc.n, wi) > 0.f)) f = rgb(0.f, 0.f, 0.f);
            }
            const float pm = is_black(f) ? 0.f : a * mixture_pdf(lobes, wi) + (1.f - a) * pb;
            RGB coeff = rgb4(P.coeff[slot]);
            uint32_t st = P.state[slot];
            if (pm > 0.f && pm < f_inf()) {
                coeff = coeff * (f * (fabsf(dot(sc.n, wi)) / pm));
                P.ray_o[slot] = f4(origin_dir(sc, wi), 0.f);
                P.ray_d[slot] = f4(wi, 0.f);
                if (ridx >= 0) {
                    float* R = P.rec + ((size_t)slot * kRecMax + ridx) * kRecFloats;
                    R[9] = wi.x; R[10] = wi.y; R[11] = wi.z;
                }
                alive = true;
                if (u_term >= 0.f) {                                     // path_tracing.cpp:112-130
                    const float mc = fmaxf(coeff.r, fmaxf(coeff.g, coeff.b));
                    if (mc < c.rr_threshold) {
                        const float pt = fmaxf(0.05f, 1.f - mc);
                        if (u_term < pt) alive = false;
                        else coeff = coeff * (1.f / (1.f - pt));
                    }
                }
            }
            if (!alive) st |= kDone;
            P.state[slot] = st;
            P.coeff[slot] = f4(coeff, 0.f);
            P.grng[slot] = rs;
        }
        uint32_t* const ctr[1] = {n_out};
        const bool pred[1] = {alive};
        uint32_t idx[1];
        block_append<1>(ctr, pred, idx);
        if (alive) out[idx[0]] = slot;
    }
}

// Resolve: L += coeff * N * (light half + BSDF half) (path_tracing.cpp:56-57, direct_lighting.cpp:551)
__global__ void k_resolve(DevScene s, const Chunk* __restrict__ cp, Paths P, Queues q) {
    const uint32_t n = q.counts[4];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t slot = q.dl_path[i];
        const RGB dl = (rgb4(P.dl_light[slot]) + rgb4(P.dl_bsdf[slot])) * (float)s.n_lights;
        const RGB L = rgb4(P.L[slot]) + rgb4(P.dl_coeff[slot]) * dl;
        P.L[slot] = f4(L, 0.f);
        const uint32_t st = P.state[slot];
        const int nrec = (int)((st >> kRecShift) & 7u);
        if (st & kRecPending) {
            float* R = P.rec + ((size_t)slot * kRecMax + (nrec - 1)) * kRecFloats;
            R[15] = L.r; R[16] = L.g; R[17] = L.b;
        }
        if (st & kDone) {
            const Chunk& c = *cp;
            const uint2 pk = P.pk[slot];
            if (c.guided) {
                uint64_t rs = P.grng[slot];
                emit_records(P, q, slot, nrec, L, rs, guide_inc(c, pk));
                P.grng[slot] = rs;
            }
            finalize(c, P, pk, L);
        }
        P.state[slot] = st & ~(kPendingDL | kDone | kRecPending);
    }
}

// Resets the output counters of an iteration (a kernel, not a memset, so it can sit in a graph).
__global__ void k_iter_reset(uint32_t* counts, int parity) {
    const int t = threadIdx.x;
    if (t == 0) counts[parity ^ 1] = 0u;
    else if (t >= 2 && t <= 4) counts[t] = 0u;
    else if (t >= 6 && t <= 9) counts[t] = 0u;
}

// Cube-map cell of a direction: dominant axis and sign pick one of 6 faces, the other two
// components over the dominant one pick one of C x C cells.
template <int C>
YG_INLINE uint32_t dir_cell(float4 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t f;
    float u, v, m;
    if (ax >= ay && ax >= az) { f = d.x < 0.f ? 1u : 0u; u = d.y; v = d.z; m = ax; }
    else if (ay >= az) { f = d.y < 0.f ? 3u : 2u; u = d.x; v = d.z; m = ay; }
    else { f = d.z < 0.f ? 5u : 4u; u = d.x; v = d.y; m = az; }
    const float sc = 0.5f * (float)C / m;
    const uint32_t iu = (uint32_t)fminf(fmaxf((u + m) * sc, 0.f), (float)(C - 1));
    const uint32_t iv = (uint32_t)fminf(fmaxf((v + m) * sc, 0.f), (float)(C - 1));
    return (f * C + iu) * C + iv;
}

// Sorts the unsorted next active list window by window (4,096 entries per window) by direction
// cell: a counting sort in LDS, written back to the same window of the active list.  The order
// inside a cell is arbitrary; every path's result depends only on its own slot, so the images do
// not depend on it.
constexpr int kBinWindow = 4096;
template <int C>
__global__ __launch_bounds__(256) void k_bin_local(Paths P, Queues q, int parity) {
    constexpr int NB = 6 * C * C;
    __shared__ uint32_t s_slot[kBinWindow];
    __shared__ uint16_t s_key[kBinWindow];
    __shared__ uint32_t s_cur[NB];
    uint32_t* out = parity ? q.active0 : q.active1;
    const uint32_t n = q.counts[parity ^ 1];
    const int t = threadIdx.x;
    for (uint32_t base = blockIdx.x * (uint32_t)kBinWindow; base < n; base += gridDim.x * (uint32_t)kBinWindow) {
        const int m = (int)min((uint32_t)kBinWindow, n - base);
        for (int j = t; j < NB; j += 256) s_cur[j] = 0u;
        __syncthreads();
        for (int j = t; j < m; j += 256) {
            const uint32_t slot = q.unsorted[base + j];
            const uint32_t key = dir_cell<C>(P.ray_d[slot]);
            s_slot[j] = slot;
            s_key[j] = (uint16_t)key;
            atomicAdd(&s_cur[key], 1u);
        }
        __syncthreads();
        if (t < 64) {
            // exclusive prefix over the NB cells: each lane sums a run of cells, then a wave scan
            constexpr int RUN = (NB + 63) / 64;
            uint32_t v[RUN], sum = 0;
#pragma unroll
            for (int r = 0; r < RUN; r++) {
                const int c = t * RUN + r;
                v[r] = c < NB ? s_cur[c] : 0u;
                sum += v[r];
            }
            uint32_t incl = sum;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(incl, off);
                if (t >= off) incl += y;
            }
            uint32_t acc = incl - sum;
#pragma unroll
            for (int r = 0; r < RUN; r++) {
                const int c = t * RUN + r;
                if (c < NB) s_cur[c] = acc;
                acc += v[r];
            }
        }
        __syncthreads();
        for (int j = t; j < m; j += 256) {
            const uint32_t pos = atomicAdd(&s_cur[s_key[j]], 1u);
            out[base + pos] = s_slot[j];
        }
        __syncthreads();
    }
}

static size_t stack_lds(const LaunchCfg& cfg) {
    if (cfg.priv_stack) return 0;
    return (size_t)2 * (cfg.scene_stack + cfg.mesh_stack) * cfg.block * sizeof(uint32_t);
}

#define YG_LAUNCH_TRAV(KERNEL, GRID, ...)                                                                  \
    do {                                                                                                   \
        StackCfg st{cfg.scene_stack, cfg.mesh_stack, cfg.refill ? 1 : 0};                                  \
        dim3 grid_((unsigned)(GRID)), block_(cfg.block);                                                   \
        size_t shm = stack_lds(cfg);                                                                       \
        if (cfg.counters) {                                                                                \
            if (cfg.priv_stack) hipLaunchKernelGGL((KERNEL<true, true>), grid_, block_, shm, cfg.stream, __VA_ARGS__, st); \
            else hipLaunchKernelGGL((KERNEL<true, false>), grid_, block_, shm, cfg.stream, __VA_ARGS__, st);             \
        } else {                                                                                           \
            if (cfg.priv_stack) hipLaunchKernelGGL((KERNEL<false, true>), grid_, block_, shm, cfg.stream, __VA_ARGS__, st); \
            else hipLaunchKernelGGL((KERNEL<false, false>), grid_, block_, shm, cfg.stream, __VA_ARGS__, st);            \
        }                                                                                                  \
    } while (0)

// Resident blocks per CU of a persistent traversal kernel (0 = k_extend, 1 = k_shadow, 2 = k_mis).
int traversal_blocks_per_cu(const LaunchCfg& cfg, int kernel) {
    int nb = 0;
    const size_t shm = stack_lds(cfg);
    hipError_t e;
#define YG_OCC(K) (cfg.priv_stack ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (K<false, true>), cfg.block, shm) \
                                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (K<false, false>), cfg.block, shm))
    if (kernel == 1) e = YG_OCC(k_shadow);
    else if (kernel == 2) e = YG_OCC(k_mis);
    else e = YG_OCC(k_extend);
#undef YG_OCC
    if (e != hipSuccess || nb <= 0) nb = 1;
    return nb;
}

void launch_stage(int stage, const DevScene& sc, const Chunk* c, const Paths& p, Queues q, int parity, const LaunchCfg& cfg) {
    switch (stage) {
    case kStageReset:
        hipLaunchKernelGGL(k_iter_reset, dim3(1), dim3(64), 0, cfg.stream, q.counts, parity);
        break;
    case kStageBin:
        if (q.bin_rays == 1) hipLaunchKernelGGL(k_bin_local<2>, dim3(cfg.path_grid), dim3(256), 0, cfg.stream, p, q, parity);
        else if (q.bin_rays == 2) hipLaunchKernelGGL(k_bin_local<4>, dim3(cfg.path_grid), dim3(256), 0, cfg.stream, p, q, parity);
        else if (q.bin_rays >= 3) hipLaunchKernelGGL(k_bin_local<8>, dim3(cfg.path_grid), dim3(256), 0, cfg.stream, p, q, parity);
        break;
    case kStageExtend: YG_LAUNCH_TRAV(k_extend, cfg.trav_grid, sc, p, q, parity); break;
    case kStageShade:
        // scenes with mix materials take the variant whose BSDF code evaluates sub-materials
        if (cfg.guided && sc.has_mix) hipLaunchKernelGGL((k_shade<true, true>), dim3(cfg.path_grid), dim3(256), 0, cfg.stream, sc, c, p, q, parity);
        else if (cfg.guided) hipLaunchKernelGGL((k_shade<true, false>), dim3(cfg.path_grid), dim3(256), 0, cfg.stream, sc, c, p, q, parity);
        else if (sc.has_mix) hipLaunchKernelGGL((k_shade<false, true>), dim3(cfg.path_grid), dim3(256), 0, cfg.stream, sc, c, p, q, parity);
        else hipLaunchKernelGGL((k_shade<false, false>), dim3(cfg.path_grid), dim3(256), 0, cfg.stream, sc, c, p, q, parity);
        break;
    case kStageShadow: YG_LAUNCH_TRAV(k_shadow, cfg.shadow_grid, sc, p, q); break;
    case kStageMis: YG_LAUNCH_TRAV(k_mis, cfg.mis_grid, sc, p, q); break;
    case kStageGuideForward:
        if (cfg.guided) guide_forward_dev(cfg.net, q.guide_in, 16, &q.counts[9], q.guide_out, kGOut, cfg.guide_grid, cfg.stream);
        break;
    case kStageGuideSample:
        if (cfg.guided)
            hipLaunchKernelGGL(k_guide_sample, dim3(cfg.path_grid), dim3(256), 0, cfg.stream, sc, c, p, q, parity);
        break;
    case kStageResolve: hipLaunchKernelGGL(k_resolve, dim3(cfg.path_grid), dim3(256), 0, cfg.stream, sc, c, p, q); break;
    default: break;
    }
}

// One bounce-loop iteration: reset -> extend -> shade -> shadow -> mis -> resolve.
void launch_iteration(const DevScene& sc, const Chunk* c, const Paths& p, Queues q, int parity, const LaunchCfg& cfg) {
    for (int stage = 0; stage < kStageCount; stage++) launch_stage(stage, sc, c, p, q, parity, cfg);
}

// ---------------------------------------------------------------------------------------------
// Film (Film_Tile::add_sample, film.cpp:27-50, box filter radius 0.5; per-pixel luminance variance,
// reference_renderer.cpp:395-409).  Because film_pos = pixel + jitter is a float sum, a sample can
// reach the 3x3 neighbourhood of its pixel.  Contributions inside the source's tile are gathered
// on the device in render_tile's visitation order; contributions to a tile's one-pixel margin go
// to a side list that the host adds to the Film_Tile.
// ---------------------------------------------------------------------------------------------
struct Footprint { int x0, x1, y0, y1; };   // [x0, x1) x [y0, y1)

// The image-plane samples of a pixel from sample k0 on, in order (consecutive stream draws).
struct ImageSamples {
    uint64_t st, inc;
    float dx, dy;
    int nx, x, y;
    YG_INLINE ImageSamples(const PixelSampler& ps, const JumpTables& jt, int k0) {
        inc = ps.inc;
        st = pcg_jump(jt, ps.s0, ps.inc, 2u * (uint32_t)k0);
        dx = 1.f / float(ps.nx);
        dy = 1.f / float(ps.ny);
        nx = ps.nx;
        x = k0 % ps.nx;
        y = k0 / ps.nx;
    }
    YG_INLINE void next(float& vx, float& vy) {
        vx = fminf((float(x) + pcg_float(st, inc)) * dx, kOneMinusEps);
        vy = fminf((float(y) + pcg_float(st, inc)) * dy, kOneMinusEps);
        if (++x == nx) { x = 0; y++; }
    }
};

YG_INLINE Footprint footprint(int x, int y, float jx, float jy) {
    const float fx = (float)x + jx, fy = (float)y + jy;   // film_pos, reference_renderer.cpp:363
    Footprint f;
    f.x0 = (int)ceilf(fx - 0.5f - 0.5f);
    f.y0 = (int)ceilf(fy - 0.5f - 0.5f);
    f.x1 = (int)floorf(fx + 0.5f - 0.5f) + 1;
    f.y1 = (int)floorf(fy + 0.5f - 0.5f) + 1;
    return f;
}

struct TileBox { int sx0, sy0, sx1, sy1, px0, py0, px1, py1, first; };

YG_INLINE TileBox tile_of(const DevView& v, const int* tile_first, int x, int y) {   // film.cpp:73-86
    TileBox t;
    const int tx = (x - v.sx0) >> 6, ty = (y - v.sy0) >> 6;
    t.sx0 = v.sx0 + tx * 64; t.sy0 = v.sy0 + ty * 64;
    t.sx1 = min(t.sx0 + 64, v.sx1); t.sy1 = min(t.sy0 + 64, v.sy1);
    t.px0 = max(t.sx0 - 1, v.rx0); t.py0 = max(t.sy0 - 1, v.ry0);
    t.px1 = min(t.sx1 + 1, v.rx1); t.py1 = min(t.sy1 + 1, v.ry1);
    t.first = tile_first[v.tile_base + ty * v.tgx + tx];
    return t;
}

// Pass 1: which neighbours a pixel's samples reach (bit b = (dy+1)*3 + (dx+1), centre excluded),
// and the samples that reach the tile's margin (outside its sample bounds) for the host.
__global__ void k_film_spill(const Chunk* __restrict__ cp, Paths P, FilmParams f, Queues q) {
    const Chunk& c = *cp;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.n_pixels) return;
    const int vi = c.pix_view[p];
    const DevView& v = c.views[vi];
    const int2 px = c.pixels[p];
    const PixelSampler ps = pixel_sampler(c, p);
    const TileBox t = tile_of(v, f.tile_first, px.x, px.y);
    uint32_t bits = 0;
    const int ka = c.guided ? c.k0 : 0, kb = c.guided ? min(c.k1, v.spp) : v.spp;
    ImageSamples is(ps, c.jt, ka);
    for (int k = ka; k < kb; k++) {
        float jx, jy;
        is.next(jx, jy);
        const Footprint fp = footprint(px.x, px.y, jx, jy);
        if (fp.x0 == px.x && fp.x1 == px.x + 1 && fp.y0 == px.y && fp.y1 == px.y + 1) continue;
        bool margin = false;
        for (int yy = fp.y0; yy < fp.y1; yy++)
            for (int xx = fp.x0; xx < fp.x1; xx++) {
                if (xx == px.x && yy == px.y) continue;
                if (xx < t.px0 || xx >= t.px1 || yy < t.py0 || yy >= t.py1) continue;   // outside Film_Tile
                if (xx < t.sx0 || xx >= t.sx1 || yy < t.sy0 || yy >= t.sy1) margin = true;
                else bits |= 1u << ((yy - px.y + 1) * 3 + (xx - px.x + 1));
            }
        if (margin) {
            const uint32_t e = atomicAdd(&q.counts[5], 1u);
            if ((int)e < f.side_capacity) {
                f.side_rad[e] = P.radiance[c.pix_first[p] + k];
                f.side_info[e] = make_int4(px.x, px.y, k, (fp.x0 - px.x + 1) | ((fp.x1 - px.x) << 2) |
                                                              ((fp.y0 - px.y + 1) << 4) | ((fp.y1 - px.y) << 6));
                f.side_view[e] = vi;
            }
        }
    }
    f.spill[p] = (uint8_t)(bits & 0x0fu) | (uint8_t)((bits >> 1) & 0xf0u);   // 9 bits minus the centre
}

YG_INLINE bool spill_bit(uint8_t packed, int dx, int dy) {       // direction from source to target
    int b = (dy + 1) * 3 + (dx + 1);
    if (b == 4) return false;
    return b < 4 ? ((packed >> b) & 1) : ((packed >> (b - 1)) & 1);
}

// Adds the samples of in-tile neighbour (nx, ny) whose footprint covers (x, y), in sample order.
YG_INLINE void gather(const Chunk& c, const Paths& P, const FilmParams& f, const TileBox& t, int ka, int kb, int x,
                      int y, int nx, int ny, float4& acc) {
    if (nx < t.sx0 || nx >= t.sx1 || ny < t.sy0 || ny >= t.sy1) return;
    const long long li = t.first + (long long)(ny - t.sy0) * (t.sx1 - t.sx0) + (nx - t.sx0) - f.base;   // chunk-local
    if (!spill_bit(f.spill[li], x - nx, y - ny)) return;
    ImageSamples is(pixel_sampler(c, (int)li), c.jt, ka);
    for (int k = ka; k < kb; k++) {
        float jx, jy;
        is.next(jx, jy);
        const Footprint fp = footprint(nx, ny, jx, jy);
        if (x >= fp.x0 && x < fp.x1 && y >= fp.y0 && y < fp.y1) {
            const float4 r = P.radiance[c.pix_first[li] + k];
            acc.x += 1.f * r.x; acc.y += 1.f * r.y; acc.z += 1.f * r.z; acc.w += 1.f;
        }
    }
}

// Pass 2: a pixel's Film_Tile value in its own tile, accumulated in render_tile's visitation order
// (row-major sources, samples in order): earlier neighbours, own samples, later neighbours.
__global__ void k_film(const Chunk* __restrict__ cp, Paths P, FilmParams f) {
    const Chunk& c = *cp;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.n_pixels) return;
    const DevView& v = c.views[c.pix_view[p]];
    const int spp = v.spp;
    const int2 px = c.pixels[p];
    const int x = px.x, y = px.y;
    const TileBox t = tile_of(v, f.tile_first, x, y);
    // plain renders gather all samples at once; guided passes add samples [k0, k1) to the film
    // and to running luminance moments, and resolve the variance on the pixel's last pass
    const int ka = c.guided ? c.k0 : 0, kb = c.guided ? min(c.k1, spp) : spp;
    if (ka >= kb) return;
    const bool add = f.accumulate && ka > 0;
    float4 acc = add ? f.out_film[f.base + p] : make_float4(0.f, 0.f, 0.f, 0.f);
    gather(c, P, f, t, ka, kb, x, y, x - 1, y - 1, acc);
    gather(c, P, f, t, ka, kb, x, y, x, y - 1, acc);
    gather(c, P, f, t, ka, kb, x, y, x + 1, y - 1, acc);
    gather(c, P, f, t, ka, kb, x, y, x - 1, y, acc);
    double lsum = add ? f.m1[f.base + p] : 0.0, lsq = add ? f.m2[f.base + p] : 0.0;
    for (int k = ka; k < kb; k++) {
        const float4 r = P.radiance[c.pix_first[p] + k];
        acc.x += 1.f * r.x; acc.y += 1.f * r.y; acc.z += 1.f * r.z; acc.w += 1.f;
        const float lv = 0.2126729f * r.x + 0.7151522f * r.y + 0.0721750f * r.z;   // color.cpp:21
        lsum += (double)lv;
        lsq += (double)(lv * lv);
    }
    gather(c, P, f, t, ka, kb, x, y, x + 1, y, acc);
    gather(c, P, f, t, ka, kb, x, y, x - 1, y + 1, acc);
    gather(c, P, f, t, ka, kb, x, y, x, y + 1, acc);
    gather(c, P, f, t, ka, kb, x, y, x + 1, y + 1, acc);
    f.out_film[f.base + p] = acc;
    if (f.accumulate) {
        f.m1[f.base + p] = lsum;
        f.m2[f.base + p] = lsq;
    }
    if (kb < spp) return;
    double pv = 0.0;
    const int n = spp;
    if (n > 1) {
        pv = (lsq - lsum * lsum / n) / (n * (n - 1));
        pv = pv > 0.0 ? pv : 0.0;
    }
    f.out_var[f.base + p] = pv;
}

void launch_film(const Chunk* c, int n_pixels, const Paths& p, FilmParams f, Queues q, hipStream_t s) {
    hipLaunchKernelGGL(k_film_spill, dim3((n_pixels + 255) / 256), dim3(256), 0, s, c, p, f, q);
    hipLaunchKernelGGL(k_film, dim3((n_pixels + 255) / 256), dim3(256), 0, s, c, p, f);
}

// ---------------------------------------------------------------------------------------------
// Scene-tree ray queries for yar_gpu_trace_rays (32-byte yar_ray in, t/b + object/triangle out)
// ---------------------------------------------------------------------------------------------
template <bool COUNT, bool PRIV>
__global__ void k_trace_rays(DevScene sc, const float4* __restrict__ rays, long long n, int any, float4* out_tb,
                             int2* out_ot, unsigned long long* totals, StackCfg cfg) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    TravCounters cnt = {0, 0, 0, 0};
    if (i < n) {
        const float4 o4 = rays[2 * i], d4 = rays[2 * i + 1];
        HitRec h;
        with_stacks<PRIV>(cfg, [&](auto& sst, int sd, auto& mst, int md) {
            if (any) {
                bool occ = trace<true, COUNT>(sc, from4(o4), from4(d4), o4.w, h, sst, sd, mst, md, cnt);
                out_tb[i] = make_float4(occ ? 1.f : 0.f, 0.f, 0.f, 0.f);
                out_ot[i] = make_int2(-1, -1);
            } else {
                bool found = trace<false, COUNT>(sc, from4(o4), from4(d4), o4.w, h, sst, sd, mst, md, cnt);
                if (found) {
                    const int tri_base = fbits(sc.objs[h.obj].hi_tri.w);
                    out_tb[i] = make_float4(h.t, h.b.x, h.b.y, h.b.z);
                    out_ot[i] = make_int2(h.obj, (int)h.tri - tri_base);
                } else {
                    out_tb[i] = make_float4(f_inf(), 0.f, 0.f, 0.f);
                    out_ot[i] = make_int2(-1, -1);
                }
            }
        });
    }
    flush_counters<COUNT>(totals, 3, cnt);
}

void launch_trace_rays(const DevScene& sc, const float4* rays, int64_t n, int any, float4* out_tb, int2* out_ot,
                       unsigned long long* totals, const LaunchCfg& cfg) {
    if (n <= 0) return;
    const long long count = n;
    YG_LAUNCH_TRAV(k_trace_rays, (count + cfg.block - 1) / cfg.block, sc, rays, count, any, out_tb, out_ot, totals);
}

}  // namespace yg

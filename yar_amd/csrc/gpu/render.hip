// render.hip - wavefront path-tracer kernels for gfx950 (CDNA4).
//
// One path per work item; the bounce loop is split into queue-driven kernels:
//   sample_table -> raygen -> [extend -> shade -> shadow, mis -> resolve]* -> film
// Each kernel restates one part of the reference renderer's per-sample loop (kennyalive/YAR
// src/ref/reference_renderer.cpp:331-429 render_tile, path_tracing.cpp:11-133 trace_path,
// direct_lighting.cpp:18-578, shading_context.cpp:17-240, bsdf.cpp, scattering.cpp,
// sampling.cpp, pixel_sampling.cpp:54-173, film.cpp:27-50) with the reference's floating-point
// operation order, so a path's radiance matches the CPU result whenever the device libm agrees
// with the host's (sin/cos in direction sampling are the only libm calls on the path).
#include "kernels.hpp"

namespace yg {

// ---------------------------------------------------------------------------------------------
// wave-aggregated queue append: one atomic per wave, lane order preserved inside the wave
// ---------------------------------------------------------------------------------------------
YG_INLINE uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    uint32_t lane = __lane_id();
    uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + prefix;
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

// ---------------------------------------------------------------------------------------------
// Sample tables (Stratified_Pixel_Sampler::next_pixel, pixel_sampling.cpp:54-85;
// generate_stratified_sequence_1d/2d, sampling.cpp:79-96; shuffle, math.h:138-144).
// One work item per (pixel, dimension); the dimension's RNG draws start at a fixed offset of the
// pixel's pcg32 stream, reached with the LCG jump-ahead, and the Fisher-Yates shuffle runs in LDS.
// Table layout per pixel: image plane [2*spp], 1D dims [d1][spp], 2D dims [d2][spp][2].
// ---------------------------------------------------------------------------------------------
__global__ void k_sample_table(Chunk c, float* __restrict__ table, int n_dims) {
    extern __shared__ float lds[];
    const int T = blockDim.x, t = threadIdx.x;
    const long long g = (long long)blockIdx.x * T + t;
    if (g >= (long long)c.n_pixels * n_dims) return;
    const int p = (int)(g / n_dims), dim = (int)(g % n_dims);
    const int spp = c.spp, nx = c.nx, ny = c.ny;
    const int2 px = c.pixels[p];
    uint32_t stream = ((uint32_t)px.x & 0xffffu) | ((uint32_t)px.y << 16);
    stream += c.seed_offset;
    uint64_t inc;
    uint64_t st = pcg_seed_state(stream, inc);
    float* out = table + (long long)p * c.stride;
    float* X = lds + t;                 // X[i * T], Y[i * T]: this item's values
    float* Y = lds + (long long)T * spp + t;
    if (dim == 0) {
        const float dx = 1.f / float(nx), dy = 1.f / float(ny);
        int k = 0;
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++, k++) {
                float vx = fminf((float(x) + pcg_float(st, inc)) * dx, kOneMinusEps);
                float vy = fminf((float(y) + pcg_float(st, inc)) * dy, kOneMinusEps);
                out[2 * k] = vx;
                out[2 * k + 1] = vy;
            }
        return;
    }
    if (dim <= c.d1) {
        const int i = dim - 1;
        st = pcg_advance(st, inc, (uint64_t)(2 * spp) + (uint64_t)i * (uint64_t)(2 * spp - 1));
        const float dx = 1.f / float(spp);
        for (int x = 0; x < spp; x++) X[x * T] = fminf((float(x) + pcg_float(st, inc)) * dx, kOneMinusEps);
        for (int n = spp; n > 1; n--) {
            uint32_t k = (uint32_t)(((uint64_t)pcg_next(st, inc) * (uint64_t)n) >> 32);
            float a = X[(n - 1) * T];
            X[(n - 1) * T] = X[k * T];
            X[k * T] = a;
        }
        float* o1 = out + 2 * spp + (long long)i * spp;
        for (int k = 0; k < spp; k++) o1[k] = X[k * T];
        return;
    }
    const int i = dim - 1 - c.d1;
    st = pcg_advance(st, inc, (uint64_t)(2 * spp) + (uint64_t)c.d1 * (uint64_t)(2 * spp - 1) +
                                  (uint64_t)i * (uint64_t)(3 * spp - 1));
    const float dx = 1.f / float(nx), dy = 1.f / float(ny);
    int k = 0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++, k++) {
            X[k * T] = fminf((float(x) + pcg_float(st, inc)) * dx, kOneMinusEps);
            Y[k * T] = fminf((float(y) + pcg_float(st, inc)) * dy, kOneMinusEps);
        }
    for (int n = spp; n > 1; n--) {
        uint32_t kk = (uint32_t)(((uint64_t)pcg_next(st, inc) * (uint64_t)n) >> 32);
        float ax = X[(n - 1) * T], ay = Y[(n - 1) * T];
        X[(n - 1) * T] = X[kk * T]; Y[(n - 1) * T] = Y[kk * T];
        X[kk * T] = ax; Y[kk * T] = ay;
    }
    float* o2 = out + 2 * spp + (long long)c.d1 * spp + 2 * (long long)i * spp;
    for (int kk = 0; kk < spp; kk++) { o2[2 * kk] = X[kk * T]; o2[2 * kk + 1] = Y[kk * T]; }
}

void launch_sample_table(const Chunk& c, float* table, hipStream_t s) {
    const int n_dims = 1 + c.d1 + c.d2;
    int T = 256;
    while (T > 1 && (long long)T * c.spp * 2 * 4 > 32768) T >>= 1;   // <= 32 KB of LDS per block
    long long items = (long long)c.n_pixels * n_dims;
    long long blocks = (items + T - 1) / T;
    size_t shm = (size_t)T * c.spp * 2 * sizeof(float);
    hipLaunchKernelGGL(k_sample_table, dim3((unsigned)blocks), dim3(T), shm, s, c, table, n_dims);
}

// ---------------------------------------------------------------------------------------------
// Ray generation (render_tile, reference_renderer.cpp:345-365; Camera::generate_ray,
// camera.cpp:16-32)
// ---------------------------------------------------------------------------------------------
__global__ void k_raygen(Chunk c, Paths P, Queues q) {
    const int n = c.n_pixels * c.s_par;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p = s / c.s_par, j = s % c.s_par;
    const int k = c.pass * c.s_par + j;
    const float* tab = c.table + (long long)p * c.stride;
    const int2 px = c.pixels[p];
    const float fx = (float)px.x + tab[2 * k];
    const float fy = (float)px.y + tab[2 * k + 1];
    const float u = 2.f * (fx / c.W) - 1.f;
    const float v = 2.f * (fy / c.H) - 1.f;
    const float right = u * c.horz, up = -v * c.vert;
    V3 d = v3(right, c.zup ? 1.f : up, c.zup ? up : -1.f);
    d = vdiv(d, length(d));
    const V3 o = xform_point(c.c2w, v3(0.f, 0.f, 0.f));
    const V3 wd = xform_vector(c.c2w, d);
    P.ray_o[s] = f4(o, 0.f);
    P.ray_d[s] = f4(wd, 0.f);
    P.coeff[s] = make_float4(1.f, 1.f, 1.f, 0.f);
    P.L[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.state[s] = 0u;
    if (c.s_par == 1 && c.pass == 0 && P.rng) {
        // fallback stream: the pixel's RNG after next_pixel() drew every table value
        uint32_t stream = ((uint32_t)px.x & 0xffffu) | ((uint32_t)px.y << 16);
        stream += c.seed_offset;
        uint64_t inc;
        uint64_t st = pcg_seed_state(stream, inc);
        const uint64_t spp = (uint64_t)c.spp;
        uint64_t draws = 2 * spp + (uint64_t)c.d1 * (2 * spp - 1) + (uint64_t)c.d2 * (3 * spp - 1);
        P.rng[s] = pcg_advance(st, inc, draws);
    }
    q.active_in[s] = (uint32_t)s;
}

void launch_raygen(const Chunk& c, const Paths& p, Queues q, hipStream_t s) {
    int n = c.n_pixels * c.s_par;
    hipLaunchKernelGGL(k_raygen, dim3((n + 255) / 256), dim3(256), 0, s, c, p, q);
}

// ---------------------------------------------------------------------------------------------
// Traversal kernels: extend (closest hit for path rays), shadow (any hit), mis (closest hit for
// BSDF-sampled light rays, direct_lighting.cpp:163-191).
// ---------------------------------------------------------------------------------------------
struct StackCfg { int sd, md; };

YG_INLINE void stacks(uint32_t* lds_base, StackCfg sc, LdsStack& sst, LdsStack& mst) {
    const int T = blockDim.x, t = threadIdx.x;
    uint32_t* snode = lds_base;
    float* stm = reinterpret_cast<float*>(lds_base + sc.sd * T);
    uint32_t* mnode = lds_base + 2 * sc.sd * T;
    float* mtm = reinterpret_cast<float*>(lds_base + 2 * sc.sd * T + sc.md * T);
    sst = LdsStack{snode + t, stm + t, T};
    mst = LdsStack{mnode + t, mtm + t, T};
}

template <bool COUNT>
__global__ void k_extend(DevScene sc, Paths P, Queues q, int count, StackCfg cfg) {
    extern __shared__ uint32_t lds_u[];
    LdsStack sst, mst;
    stacks(lds_u, cfg, sst, mst);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    TravCounters cnt = {0, 0, 0, 0};
    if (i < count) {
        const uint32_t slot = q.active_in[i];
        const float4 o4 = P.ray_o[slot], d4 = P.ray_d[slot];
        HitRec h;
        trace<false, COUNT>(sc, from4(o4), from4(d4), f_inf(), h, sst, cfg.sd, mst, cfg.md, cnt);
        P.hit_tb[slot] = make_float4(h.t, h.b.x, h.b.y, h.b.z);
        P.hit_ot[slot] = make_int2(h.obj, (int)h.tri);
    }
    flush_counters<COUNT>(q.totals, 0, cnt);
}

template <bool COUNT>
__global__ void k_shadow(DevScene sc, Paths P, Queues q, int count, StackCfg cfg) {
    extern __shared__ uint32_t lds_u[];
    LdsStack sst, mst;
    stacks(lds_u, cfg, sst, mst);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    TravCounters cnt = {0, 0, 0, 0};
    if (i < count) {
        const float4 o4 = q.shadow_o[i], d4 = q.shadow_d[i];
        HitRec h;
        bool occluded = trace<true, COUNT>(sc, from4(o4), from4(d4), o4.w, h, sst, cfg.sd, mst, cfg.md, cnt);
        if (!occluded) P.dl_light[q.shadow_path[i]] = q.shadow_c[i];
    }
    flush_counters<COUNT>(q.totals, 1, cnt);
}

YG_INLINE float power_heuristic(float a, float b) {             // direct_lighting.cpp:18
    if (a == f_inf()) return 1.f;
    return a * a / (a * a + b * b);
}

template <bool COUNT>
__global__ void k_mis(DevScene sc, Paths P, Queues q, int count, StackCfg cfg) {
    extern __shared__ uint32_t lds_u[];
    LdsStack sst, mst;
    stacks(lds_u, cfg, sst, mst);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    TravCounters cnt = {0, 0, 0, 0};
    if (i < count) {
        const float4 o4 = q.mis_o[i], d4 = q.mis_d[i];
        const V3 o = from4(o4);
        HitRec h;
        bool found = trace<false, COUNT>(sc, o, from4(d4), f_inf(), h, sst, cfg.sd, mst, cfg.md, cnt);
        const int light = (int)fbits(o4.w);
        if (found && sc.obj_info[h.obj].y == light) {
            const DevRect L = sc.rects[light];
            const float4 v0 = sc.tri_verts[3 * (size_t)h.tri], v1 = sc.tri_verts[3 * (size_t)h.tri + 1],
                         v2 = sc.tri_verts[3 * (size_t)h.tri + 2];
            // mesh-space hit point, as Triangle_Mesh::get_position (direct_lighting.cpp:181)
            const V3 p = h.b.x * from4(v0) + h.b.y * from4(v1) + h.b.z * from4(v2);
            const float dd = length(p - o);
            const float lndw = d4.w;
            const float lpdf = (dd * dd) / (L.size.x * L.size.y * lndw);
            const float4 f4v = q.mis_f[i];
            const float bpdf = f4v.w;
            const float w = power_heuristic(bpdf, lpdf);
            const RGB f = rgb4(f4v);
            const RGB contrib = (rgb4(L.L) * f) * (w * fabsf(q.mis_ndw[i]) / bpdf);
            P.dl_bsdf[q.mis_path[i]] = f4(contrib, 0.f);
        }
    }
    flush_counters<COUNT>(q.totals, 2, cnt);
}

static size_t stack_lds(const LaunchCfg& cfg) {
    return (size_t)2 * (cfg.scene_stack + cfg.mesh_stack) * cfg.block * sizeof(uint32_t);
}

void launch_extend(const DevScene& sc, const Paths& p, Queues q, int count, const LaunchCfg& cfg) {
    if (count <= 0) return;
    StackCfg st{cfg.scene_stack, cfg.mesh_stack};
    dim3 grid((count + cfg.block - 1) / cfg.block);
    if (cfg.counters) hipLaunchKernelGGL(k_extend<true>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, p, q, count, st);
    else hipLaunchKernelGGL(k_extend<false>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, p, q, count, st);
}
void launch_shadow(const DevScene& sc, const Paths& p, Queues q, int count, const LaunchCfg& cfg) {
    if (count <= 0) return;
    StackCfg st{cfg.scene_stack, cfg.mesh_stack};
    dim3 grid((count + cfg.block - 1) / cfg.block);
    if (cfg.counters) hipLaunchKernelGGL(k_shadow<true>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, p, q, count, st);
    else hipLaunchKernelGGL(k_shadow<false>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, p, q, count, st);
}
void launch_mis(const DevScene& sc, const Paths& p, Queues q, int count, const LaunchCfg& cfg) {
    if (count <= 0) return;
    StackCfg st{cfg.scene_stack, cfg.mesh_stack};
    dim3 grid((count + cfg.block - 1) / cfg.block);
    if (cfg.counters) hipLaunchKernelGGL(k_mis<true>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, p, q, count, st);
    else hipLaunchKernelGGL(k_mis<false>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, p, q, count, st);
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

YG_INLINE RGB bsdf_eval(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    if (m.type == 1) return kPiInv * rgb4(m.refl);
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    float ci = dot(wil, wh);
    RGB F = conductor_fresnel(ci, m.eta_i, rgb4(m.eta), rgb4(m.k));
    float G = G_aniso(wil, wol, m.alpha_x, m.alpha_y);
    float D = D_aniso(wh, m.alpha_x, m.alpha_y);
    return F * ((G * D) / (4.f * wol.z * wil.z));
}
YG_INLINE float bsdf_pdf(const DevMaterial& m, const Shading& sc, V3 wo, V3 wi) {
    if (m.type == 1) return dot(sc.n, wi) / kPi;
    V3 wol = to_local(sc, wo), wil = to_local(sc, wi);
    V3 wh = normalized(wol + wil);
    return vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
}
YG_INLINE RGB bsdf_sample(const DevMaterial& m, const Shading& sc, float u0, float u1, V3 wo, V3& wi, float& pdf) {
    if (m.type == 1) {
        float z = sqrtf(1.f - u0);                                                     // sampling.cpp:43
        float r = sqrtf(u0);
        float phi = 2.f * kPi * u1;
        V3 l = v3(r * cosf(phi), r * sinf(phi), z);
        wi = to_world(sc, l);
        pdf = bsdf_pdf(m, sc, wo, wi);
        return bsdf_eval(m, sc, wo, wi);
    }
    V3 wol = to_local(sc, wo);
    V3 wh = vndf_sample(u0, u1, wol, m.alpha_x, m.alpha_y);
    V3 wil = reflect(wol, wh);
    if (wil.z <= 0.f) return rgb(0.f, 0.f, 0.f);
    wi = to_world(sc, wil);
    pdf = vndf_wi_pdf(wol, wh, m.alpha_x, m.alpha_y);
    return bsdf_eval(m, sc, wo, wi);
}

// ---------------------------------------------------------------------------------------------
// Shade: one iteration of trace_path's loop body (path_tracing.cpp:22-131) for every active path
// ---------------------------------------------------------------------------------------------
YG_INLINE void finalize(const Chunk& c, const Paths& P, uint32_t slot, RGB L) {   // render_tile :386-393
    float mc = fmaxf(L.r, fmaxf(L.g, L.b));
    if (mc > c.max_component) L = L * (c.max_component / mc);
    if (c.film_scale != 1.f) L = L * c.film_scale;
    const uint32_t p = slot / (uint32_t)c.s_par, j = slot % (uint32_t)c.s_par;
    P.radiance[(long long)p * c.spp + c.pass * c.s_par + j] = f4(L, 0.f);
}

__global__ void k_shade(DevScene s, Chunk c, Paths P, Queues q, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < count;
    bool want_active = false, want_shadow = false, want_mis = false, want_dl = false;
    uint32_t slot = 0;
    V3 sh_o, sh_d, mi_o, mi_d;
    float sh_tmax = 0.f, mi_pdf = 0.f, mi_ndw = 0.f, mi_lndw = 0.f;
    RGB sh_c, mi_f;
    int mi_light = 0;
    if (valid) {
        slot = q.active_in[i];
        uint32_t st = P.state[slot];
        int bounce = (int)(st & 0xffu), c1 = (int)((st >> 8) & 0xffu), c2 = (int)((st >> 16) & 0xffu);
        bool delta_traced = (st & kDelta) != 0;
        RGB coeff = rgb4(P.coeff[slot]);
        RGB L = rgb4(P.L[slot]);
        const float4 d4 = P.ray_d[slot];
        const int2 ot = P.hit_ot[slot];
        const float4 tb = P.hit_tb[slot];
        const bool hit = ot.x >= 0;
        const int light = hit ? s.obj_info[ot.x].y : -1;
        const int p = (int)(slot / (uint32_t)c.s_par), j = (int)(slot % (uint32_t)c.s_par);
        const int k = c.pass * c.s_par + j;
        const float* tab = c.table + (long long)p * c.stride;
        const bool seq = (c.s_par == 1) && P.rng;
        uint64_t rs = 0, rinc = 0;
        if (seq) {
            const int2 px = c.pixels[p];
            uint32_t stream = ((uint32_t)px.x & 0xffffu) | ((uint32_t)px.y << 16);
            stream += c.seed_offset;
            rinc = ((uint64_t)stream << 1u) | 1u;
            rs = P.rng[slot];
        }
        auto next1 = [&]() -> float {                                   // pixel_sampling.cpp:159
            if (c1 < c.d1) { float v = tab[2 * c.spp + c1 * c.spp + k]; c1++; return v; }
            return pcg_float(rs, rinc);
        };
        auto next2 = [&](float& x, float& y) {                          // pixel_sampling.cpp:167
            if (c2 < c.d2) {
                const float* q2 = tab + 2 * c.spp + c.d1 * c.spp + 2 * (c2 * c.spp + k);
                x = q2[0]; y = q2[1]; c2++;
                return;
            }
            x = pcg_float(rs, rinc);
            y = pcg_float(rs, rinc);
        };
        bool finished = false, pending_dl = false;
        bool continue_body = true;
        if (!delta_traced) {
            if (bounce == 0) {
                if (hit && light >= 0) L = L + coeff * rgb4(s.rects[light].L);
                if (c.max_bounces == 0) { finished = true; continue_body = false; }
            }
            if (continue_body && (!hit || light >= 0)) { finished = true; continue_body = false; }
        } else {
            RGB emitted = rgb(0.f, 0.f, 0.f);
            if (hit && light >= 0) emitted = rgb4(s.rects[light].L);
            L = L + coeff * emitted;
            bounce++;
            if (bounce == c.max_bounces || !hit || light >= 0) { finished = true; continue_body = false; }
            else if (bounce >= c.rr_bounce_threshold) {                // path_tracing.cpp:112-130
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
            next1();                                                    // u_scattering_type_next_segment
            float ulx, uly, ubx, uby, unx, uny;
            next2(ulx, uly);
            next2(ubx, uby);
            next2(unx, uny);
            const DevMaterial m = s.materials[sc.material];
            if (m.type == 3) {
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
                    const float ci = dot(sc.wo, sc.n);                  // refract, math.h:115
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
                // estimate_direct_lighting_from_single_sample (direct_lighting.cpp:522-578)
                const int li = int(u_sel * s.n_lights);
                pending_dl = true;
                if (li < s.n_points) {
                    const DevPoint pl = s.points[li];                   // direct_lighting.cpp:26-47
                    const V3 lp = from4(pl.p);
                    const V3 position = origin_dir(sc, lp - sc.pos);
                    const V3 lv = lp - position;
                    const float dist = length(lv);
                    const V3 ldir = vdiv(lv, dist);
                    const float ndl = dot(sc.n, ldir);
                    if (ndl > 0.f) {
                        const RGB f = bsdf_eval(m, sc, sc.wo, ldir);
                        sh_c = (rgb4(pl.I) * f) * (ndl / (dist * dist));
                        sh_o = position; sh_d = ldir; sh_tmax = dist * (1.f - 1e-5f);
                        want_shadow = true;
                    }
                } else {
                    const int ri = li - s.n_points;                     // direct_lighting.cpp:109-194
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
                            if (ndw > 0.f) {
                                const RGB f = bsdf_eval(m, sc, sc.wo, wi);
                                if (!is_black(f)) {
                                    const float lpdf = (dd * dd) / (R.size.x * R.size.y * lndw);
                                    const float bpdf = bsdf_pdf(m, sc, sc.wo, wi);
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
                        const RGB f = bsdf_sample(m, sc, ubx, uby, sc.wo, wi, bpdf);
                        if (!is_black(f)) {
                            const float lndw = dot(ln, -wi);
                            if (lndw > 1e-4f) {
                                mi_o = origin_dir(sc, wi); mi_d = wi; mi_f = f; mi_pdf = bpdf;
                                mi_ndw = dot(sc.n, wi); mi_lndw = lndw; mi_light = ri;
                                want_mis = true;
                            }
                        }
                    }
                }
                P.dl_coeff[slot] = f4(coeff, 0.f);
                P.dl_light[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                P.dl_bsdf[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                bounce++;
                if (bounce == c.max_bounces) {
                    finished = true;
                } else {
                    V3 wi = v3(0.f, 0.f, 0.f);
                    float pdf = 0.f;
                    const RGB f = bsdf_sample(m, sc, unx, uny, sc.wo, wi, pdf);
                    if (is_black(f)) {
                        finished = true;
                    } else {
                        coeff = coeff * (f * (fabsf(dot(sc.n, wi)) / (pdf * 1.f)));
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
        if (pending_dl) { flags |= kPendingDL; want_dl = true; if (finished) flags |= kDone; }
        else if (finished) finalize(c, P, slot, L);
        P.state[slot] = (uint32_t)bounce | ((uint32_t)c1 << 8) | ((uint32_t)c2 << 16) | flags;
        P.coeff[slot] = f4(coeff, 0.f);
        P.L[slot] = f4(L, 0.f);
        if (seq) P.rng[slot] = rs;
    }
    const uint32_t ia = wave_append(&q.counters[0], want_active);
    const uint32_t is = wave_append(&q.counters[1], want_shadow);
    const uint32_t im = wave_append(&q.counters[2], want_mis);
    const uint32_t id = wave_append(&q.counters[3], want_dl);
    if (want_active) q.active_out[ia] = slot;
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
}

void launch_shade(const DevScene& sc, const Chunk& c, const Paths& p, Queues q, int count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_shade, dim3((count + 255) / 256), dim3(256), 0, s, sc, c, p, q, count);
}

// Resolve: L += coeff * N * (light half + BSDF half) (path_tracing.cpp:56-57, direct_lighting.cpp:551)
__global__ void k_resolve(DevScene s, Chunk c, Paths P, Queues q, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t slot = q.dl_path[i];
    const RGB dl = (rgb4(P.dl_light[slot]) + rgb4(P.dl_bsdf[slot])) * (float)s.n_lights;
    const RGB L = rgb4(P.L[slot]) + rgb4(P.dl_coeff[slot]) * dl;
    P.L[slot] = f4(L, 0.f);
    uint32_t st = P.state[slot];
    if (st & kDone) finalize(c, P, slot, L);
    P.state[slot] = st & ~(kPendingDL | kDone);
}

void launch_resolve(const DevScene& sc, const Chunk& c, const Paths& p, Queues q, int count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_resolve, dim3((count + 255) / 256), dim3(256), 0, s, sc, c, p, q, count);
}

// ---------------------------------------------------------------------------------------------
// Film (Film_Tile::add_sample, film.cpp:27-50, box filter radius 0.5; per-pixel luminance variance,
// reference_renderer.cpp:395-409).
// ---------------------------------------------------------------------------------------------
struct Footprint { int x0, x1, y0, y1; };   // [x0, x1) x [y0, y1)

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

YG_INLINE TileBox tile_of(const FilmAcc& f, int x, int y) {          // Film::get_tile_bounds, film.cpp:73-86
    TileBox t;
    const int tx = (x - f.sx0) >> 6, ty = (y - f.sy0) >> 6;
    t.sx0 = f.sx0 + tx * 64; t.sy0 = f.sy0 + ty * 64;
    t.sx1 = min(t.sx0 + 64, f.sx1); t.sy1 = min(t.sy0 + 64, f.sy1);
    t.px0 = max(t.sx0 - 1, f.rx0); t.py0 = max(t.sy0 - 1, f.ry0);
    t.px1 = min(t.sx1 + 1, f.rx1); t.py1 = min(t.sy1 + 1, f.ry1);
    t.first = f.tile_first[ty * f.tgx + tx];
    return t;
}

// Pass 1: which neighbours a pixel's samples reach (bit b = (dy+1)*3 + (dx+1), centre excluded),
// and the samples that reach the tile's margin (outside its sample bounds) for the host.
__global__ void k_film_spill(Chunk c, Paths P, FilmAcc f, Queues q) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.n_pixels) return;
    const int2 px = c.pixels[p];
    const float* tab = c.table + (long long)p * c.stride;
    const TileBox t = tile_of(f, px.x, px.y);
    uint32_t bits = 0;
    for (int k = 0; k < c.spp; k++) {
        const Footprint fp = footprint(px.x, px.y, tab[2 * k], tab[2 * k + 1]);
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
            const uint32_t e = atomicAdd(&q.counters[4], 1u);
            if ((int)e < f.side_capacity) {
                f.side_rad[e] = P.radiance[(long long)p * c.spp + k];
                f.side_info[e] = make_int4(px.x, px.y, k, (fp.x0 - px.x + 1) | ((fp.x1 - px.x) << 2) |
                                                              ((fp.y0 - px.y + 1) << 4) | ((fp.y1 - px.y) << 6));
            }
        }
    }
    f.spill[p] = (uint8_t)(bits & 0xffu) | (uint8_t)((bits >> 1) & 0xf0u);   // pack 9 bits minus centre
}

YG_INLINE bool spill_bit(uint8_t packed, int dx, int dy) {       // direction from source to target
    int b = (dy + 1) * 3 + (dx + 1);
    if (b == 4) return false;
    return b < 4 ? ((packed >> b) & 1) : ((packed >> (b - 1)) & 1);
}

// Adds the samples of in-tile neighbour (nx, ny) whose footprint covers (x, y), in sample order.
YG_INLINE void gather(const Chunk& c, const Paths& P, const FilmAcc& f, const TileBox& t, int x, int y, int nx, int ny,
                      float4& acc) {
    if (nx < t.sx0 || nx >= t.sx1 || ny < t.sy0 || ny >= t.sy1) return;
    const int li = t.first + (ny - t.sy0) * (t.sx1 - t.sx0) + (nx - t.sx0) - f.base;   // chunk-local index
    if (!spill_bit(f.spill[li], x - nx, y - ny)) return;
    const float* tab = c.table + (long long)li * c.stride;
    for (int k = 0; k < c.spp; k++) {
        const Footprint fp = footprint(nx, ny, tab[2 * k], tab[2 * k + 1]);
        if (x >= fp.x0 && x < fp.x1 && y >= fp.y0 && y < fp.y1) {
            const float4 r = P.radiance[(long long)li * c.spp + k];
            acc.x += 1.f * r.x; acc.y += 1.f * r.y; acc.z += 1.f * r.z; acc.w += 1.f;
        }
    }
}

// Pass 2: a pixel's Film_Tile value in its own tile, accumulated in render_tile's visitation order
// (row-major sources, samples in order): earlier neighbours, own samples, later neighbours.
__global__ void k_film(Chunk c, Paths P, FilmAcc f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.n_pixels) return;
    const int2 px = c.pixels[p];
    const int x = px.x, y = px.y;
    const TileBox t = tile_of(f, x, y);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    gather(c, P, f, t, x, y, x - 1, y - 1, acc);
    gather(c, P, f, t, x, y, x, y - 1, acc);
    gather(c, P, f, t, x, y, x + 1, y - 1, acc);
    gather(c, P, f, t, x, y, x - 1, y, acc);
    double lsum = 0.0, lsq = 0.0;
    for (int k = 0; k < c.spp; k++) {
        const float4 r = P.radiance[(long long)p * c.spp + k];
        acc.x += 1.f * r.x; acc.y += 1.f * r.y; acc.z += 1.f * r.z; acc.w += 1.f;
        const float lv = 0.2126729f * r.x + 0.7151522f * r.y + 0.0721750f * r.z;   // color.cpp:21
        lsum += (double)lv;
        lsq += (double)(lv * lv);
    }
    gather(c, P, f, t, x, y, x + 1, y, acc);
    gather(c, P, f, t, x, y, x - 1, y + 1, acc);
    gather(c, P, f, t, x, y, x, y + 1, acc);
    gather(c, P, f, t, x, y, x + 1, y + 1, acc);
    f.out_film[f.base + p] = acc;
    double pv = 0.0;
    const int n = c.spp;
    if (n > 1) {
        pv = (lsq - lsum * lsum / n) / (n * (n - 1));
        pv = pv > 0.0 ? pv : 0.0;
    }
    f.out_var[f.base + p] = pv;
}

void launch_film(const Chunk& c, const Paths& p, FilmAcc f, Queues q, hipStream_t s) {
    hipLaunchKernelGGL(k_film_spill, dim3((c.n_pixels + 255) / 256), dim3(256), 0, s, c, p, f, q);
    hipLaunchKernelGGL(k_film, dim3((c.n_pixels + 255) / 256), dim3(256), 0, s, c, p, f);
}

// ---------------------------------------------------------------------------------------------
// Scene-tree ray queries for yar_gpu_trace_rays (32-byte yar_ray in, t/b + object/triangle out)
// ---------------------------------------------------------------------------------------------
template <bool COUNT>
__global__ void k_trace_rays(DevScene sc, const float4* __restrict__ rays, long long n, int any, float4* out_tb,
                             int2* out_ot, unsigned long long* totals, StackCfg cfg) {
    extern __shared__ uint32_t lds_u[];
    LdsStack sst, mst;
    stacks(lds_u, cfg, sst, mst);
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    TravCounters cnt = {0, 0, 0, 0};
    if (i < n) {
        const float4 o4 = rays[2 * i], d4 = rays[2 * i + 1];
        HitRec h;
        if (any) {
            bool occ = trace<true, COUNT>(sc, from4(o4), from4(d4), o4.w, h, sst, cfg.sd, mst, cfg.md, cnt);
            out_tb[i] = make_float4(occ ? 1.f : 0.f, 0.f, 0.f, 0.f);
            out_ot[i] = make_int2(-1, -1);
        } else {
            bool found = trace<false, COUNT>(sc, from4(o4), from4(d4), o4.w, h, sst, cfg.sd, mst, cfg.md, cnt);
            if (found) {
                const int tri_base = fbits(sc.objs[h.obj].hi_tri.w);
                out_tb[i] = make_float4(h.t, h.b.x, h.b.y, h.b.z);
                out_ot[i] = make_int2(h.obj, (int)h.tri - tri_base);
            } else {
                out_tb[i] = make_float4(f_inf(), 0.f, 0.f, 0.f);
                out_ot[i] = make_int2(-1, -1);
            }
        }
    }
    flush_counters<COUNT>(totals, 3, cnt);
}

void launch_trace_rays(const DevScene& sc, const float4* rays, int64_t n, int any, float4* out_tb, int2* out_ot,
                       unsigned long long* totals, const LaunchCfg& cfg) {
    if (n <= 0) return;
    StackCfg st{cfg.scene_stack, cfg.mesh_stack};
    dim3 grid((unsigned)((n + cfg.block - 1) / cfg.block));
    if (cfg.counters)
        hipLaunchKernelGGL(k_trace_rays<true>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, rays, (long long)n, any,
                           out_tb, out_ot, totals, st);
    else
        hipLaunchKernelGGL(k_trace_rays<false>, grid, dim3(cfg.block), stack_lds(cfg), cfg.stream, sc, rays, (long long)n, any,
                           out_tb, out_ot, totals, st);
}

}  // namespace yg

// context.hip - GPU context of the path-tracing core and its C ABI (include/yar_gpu.h).
//
// yar_gpu_create flattens a yar_scene_desc into HBM (kd-trees built on the host, byte-identical
// to the reference builder), yar_gpu_render / yar_gpu_render_tiles run the wavefront bounce loop
// over chunks of pixels and assemble Film_Tiles exactly as render_scene does
// (kennyalive/YAR src/ref/reference_renderer.cpp:479-584, film.cpp:73-118), and
// yar_gpu_trace_rays exposes the scene-tree queries (kdtree.cpp:219-439) for parity checks.
#include "kernels.hpp"
#include "../host/kdtree_builder.hpp"

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <cstdio>
#include <string>
#include <exception>
#include <thread>
#include <vector>

namespace {

using namespace yg;

struct HipError {
    std::string msg;
    int code = YAR_ERR_HIP;
};

#define HIPX(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)};                       \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void reserve(size_t count) {
        if (count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        HIPX(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        reserve(v.size());
        if (!v.empty()) HIPX(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Host staging kept across jobs: pinned memory, so copies run at link speed, or pageable memory
// when no pinned memory of that size can be had (the copies then stage through the driver).
struct HostBuf {
    char* p = nullptr;
    size_t n = 0;
    bool pinned = false;
    std::vector<char> pageable;
    void reserve(size_t bytes) {
        if (bytes <= n) return;
        release();
        if (hipHostMalloc((void**)&p, bytes, hipHostMallocDefault) == hipSuccess) {
            pinned = true;
        } else {
            (void)hipGetLastError();          // clear the failed call's error state
            p = nullptr;
            pageable.resize(bytes);
            p = pageable.data();
        }
        n = bytes;
    }
    void release() {
        if (pinned && p) (void)hipHostFree(p);
        std::vector<char>().swap(pageable);
        p = nullptr;
        n = 0;
        pinned = false;
    }
    ~HostBuf() { release(); }
};

static int pow2_at_least(int v) { int p = 1; while (p < v) p <<= 1; return p; }

// body(i) for i in [0, n) on up to 16 host threads, contiguous ranges per thread, each thread taking
// at least `grain` items (starting a thread costs about as much as a few 64x64 tiles of work, so
// small jobs stay on the calling thread)
template <class Body>
static void parallel_for(size_t n, Body&& body, size_t grain = 4) {
    const size_t nt = std::min<size_t>({n / std::max<size_t>(grain, 1), (size_t)16,
                                        (size_t)std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) body(i);
        return;
    }
    // an exception in a worker (std::bad_alloc) is carried to the calling thread, not std::terminate
    std::vector<std::exception_ptr> err(nt);
    auto range = [&](size_t t) {
        try {
            const size_t a = n * t / nt, b = n * (t + 1) / nt;
            for (size_t i = a; i < b; i++) body(i);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    try {
        for (size_t t = 1; t < nt; t++) pool.emplace_back(range, t);
    } catch (...) {                          // no more threads: the ranges not started run here
    }
    for (size_t t = 1 + pool.size(); t < nt; t++) range(t);
    range(0);
    for (auto& th : pool) th.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// Per-kernel HIP-event timing (kind: 0 extend, 1 shade, 2 shadow, 3 mis, 4 other, 5 guide network
// and sampling, 6 guide training).
constexpr int kTimerKinds = 7;
struct KernelTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<int, size_t>> marks;   // kind, index of the start event
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIPX(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int kind) {
        if (!on) return;
        marks.push_back({kind, used});
        HIPX(hipEventRecord(next(), s));
    }
    void end() {
        if (!on) return;
        HIPX(hipEventRecord(next(), s));
    }
    void collect(double out[kTimerKinds]) {
        for (int k = 0; k < kTimerKinds; k++) out[k] = 0;
        for (auto& m : marks) {
            float ms = 0.f;
            HIPX(hipEventElapsedTime(&ms, pool[m.second], pool[m.second + 1]));
            out[m.first] += ms * 1e-3;
        }
        marks.clear();
        used = 0;
    }
    ~KernelTimer() { for (auto e : pool) (void)hipEventDestroy(e); }
};

// Scratch of the allocation-free ray queries (yar_gpu_trace_rays_device_noalloc): per-ray hit staging
// (t and barycentrics, object and triangle) for up to kNoAllocMaxRays rays, then the 16 traversal
// counters.  Each device holds its own copy from module load until the process ends, and every call
// on a device shares it.  In host code `scratch` names no device address: kernels receive the address
// hipGetSymbolAddress returned for the current device (scratch_ptr below).
constexpr int64_t kNoAllocMaxRays = (int64_t)1 << 20;
constexpr size_t kNoAllocTbBytes = (size_t)kNoAllocMaxRays * sizeof(float4);
constexpr size_t kNoAllocOtBytes = (size_t)kNoAllocMaxRays * sizeof(int2);
constexpr size_t kNoAllocScratchBytes = kNoAllocTbBytes + kNoAllocOtBytes + 16 * sizeof(unsigned long long);
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up once per device.
unsigned char* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) throw HipError{"device ordinal beyond the scratch table", YAR_ERR_CAPACITY};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError{"hipGetSymbolAddress returned no address for the scratch array"};
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[device];
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

}  // namespace

struct yar_gpu_ctx {
    std::string error;
    int device = 0;
    int cu_count = 1;
    unsigned char* scratch = nullptr;     // this device's address of the allocation-free scratch array
    hipStream_t stream = nullptr;
    yar_gpu_options opt{};
    yar_render_config cfg{};
    yar_gpu_info info{};
    std::vector<yar_view> views;

    // scene buffers
    DevBuf<uint2> scene_nodes, mesh_nodes;
    DevBuf<ObjTrav> objs;
    DevBuf<float4> tri_verts, obj_o2w, obj_nrm, vtx_n;
    DevBuf<int4> obj_info, mesh_info, tri_vidx;
    DevBuf<float2> vtx_uv;
    DevBuf<DevMaterial> materials;
    DevBuf<DevRect> rects;
    DevBuf<DevPoint> points;
    DevBuf<DevSpot> spots;
    DevBuf<DevDirLight> dirs;
    DevBuf<DevMeshLight> mesh_lights;
    DevBuf<float> mesh_light_cdf;
    DevBuf<DevSphere> spheres;
    DevBuf<DevEnv> env;
    DevScene scene{};
    std::vector<int32_t> tri_base;        // per object (for hit conversion on the host)

    // render buffers
    DevBuf<int2> pixels;
    DevBuf<int> pix_view, tile_first, side_view;
    DevBuf<long long> pix_first;
    DevBuf<DevView> dviews;
    DevBuf<Chunk> chunks, chunk_slot;
    DevBuf<uint8_t> perm;                 // sampler shuffle permutations of the current chunk
    DevBuf<ulonglong2> jump_lo, jump_hi;  // pcg32 jump-ahead tables
    DevBuf<float4> ray_o, ray_d, coeff, L, dl_coeff, dl_light, dl_bsdf, hit_tb, radiance;
    DevBuf<int2> hit_ot;
    DevBuf<uint2> pk;
    DevBuf<uint32_t> state;
    DevBuf<uint64_t> rng;
    DevBuf<uint32_t> q_a, q_b, sh_path, mi_path, dl_path, counts, q_unsorted;
    DevBuf<float4> sh_o, sh_d, sh_c, mi_o, mi_d, mi_f;
    DevBuf<float> mi_ndw;
    DevBuf<unsigned long long> totals;
    DevBuf<float4> out_film, side_rad;
    DevBuf<double> out_var;
    DevBuf<int4> side_info;
    DevBuf<uint8_t> spill;
    // guided form
    DevBuf<uint64_t> grng;
    DevBuf<float> rec, guide_in, guide_out;
    DevBuf<uint32_t> guide_path;
    DevBuf<GuideRecord> train;
    DevBuf<double> m1, m2;
    uint32_t* h_poll = nullptr;           // pinned: active counts read back while the loop runs
    size_t h_poll_n = 0;
    HostBuf h_film, h_var;                // per-pixel film sums and variances read back per job
    HostBuf h_stage;                      // pixel list, pixel views and spp prefix sums sent per job
    std::vector<yar_film_pixel> merge_work;  // one view's merged film, kept to avoid fresh pages per render
    // bounce-loop graph (captured per job shape)
    hipGraphExec_t graph = nullptr;
    long long graph_paths = -1;
    int graph_iters = -1;
    Paths graph_P{};
    Queues graph_q{};
    // trace_rays buffers
    DevBuf<float4> tr_rays, tr_tb;
    DevBuf<int2> tr_ot;
    yar_render_stats last_trace{};

    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_paths = -1;
        graph_iters = -1;
    }

    ~yar_gpu_ctx() {
        if (device >= 0) (void)hipSetDevice(device);
        drop_graph();
        DevBuf<uint2>* b2[] = {&scene_nodes, &mesh_nodes, &pk};
        for (auto* b : b2) b->release();
        objs.release();
        DevBuf<float4>* b4[] = {&tri_verts, &obj_o2w, &obj_nrm, &vtx_n, &ray_o, &ray_d, &coeff, &L, &dl_coeff,
                                &dl_light, &dl_bsdf, &hit_tb, &radiance, &sh_o, &sh_d, &sh_c, &mi_o, &mi_d,
                                &mi_f, &out_film, &side_rad, &tr_rays, &tr_tb};
        for (auto* b : b4) b->release();
        DevBuf<int4>* bi4[] = {&obj_info, &mesh_info, &tri_vidx, &side_info};
        for (auto* b : bi4) b->release();
        DevBuf<int>* bi[] = {&pix_view, &tile_first, &side_view};
        for (auto* b : bi) b->release();
        vtx_uv.release(); materials.release(); rects.release(); points.release();
        spots.release(); dirs.release(); mesh_lights.release(); mesh_light_cdf.release(); spheres.release(); env.release();
        pixels.release(); pix_first.release(); dviews.release(); chunks.release(); chunk_slot.release();
        perm.release(); jump_lo.release(); jump_hi.release(); hit_ot.release(); tr_ot.release(); state.release(); rng.release();
        DevBuf<uint32_t>* bu[] = {&q_a, &q_b, &sh_path, &mi_path, &dl_path, &counts, &q_unsorted};
        for (auto* b : bu) b->release();
        mi_ndw.release(); totals.release(); out_var.release(); spill.release();
        grng.release(); rec.release(); guide_in.release(); guide_out.release(); guide_path.release(); train.release();
        m1.release(); m2.release();
        if (h_poll) (void)hipHostFree(h_poll);
        h_film.release();
        h_var.release();
        h_stage.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    LaunchCfg launch_cfg() const {
        LaunchCfg lc{};
        lc.block = opt.block_size > 0 ? opt.block_size : 64;
        lc.scene_stack = info.stack_depth >> 16;
        lc.mesh_stack = info.stack_depth & 0xffff;
        lc.priv_stack = opt.stack_mode == 2;
        lc.refill = opt.lane_refill == 1;
        lc.counters = opt.enable_counters != 0;
        lc.trav_grid = lc.shadow_grid = lc.mis_grid = 1;
        lc.path_grid = 1;
        lc.stream = stream;
        lc.guided = false;
        lc.net = GuideNet{nullptr, nullptr, nullptr};
        lc.guide_grid = 0;
        return lc;
    }
};

namespace {

static int fail(yar_gpu_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->error = msg;
    yar::set_last_error(msg);
    return code;
}

// Every extern "C" entry point runs its body through guarded(): no C++ exception crosses the C ABI;
// each becomes a status code and the context's (or the thread's) last-error text.
template <class Body>
static int guarded(yar_gpu_ctx* ctx, const char* name, Body&& body) {
    try {
        body();
        return YAR_OK;
    } catch (const HipError& e) {
        return fail(ctx, e.code, std::string(name) + ": " + e.msg);
    } catch (const std::bad_alloc&) {
        return fail(ctx, YAR_ERR_OUT_OF_MEMORY, std::string(name) + ": out of host memory");
    } catch (const std::exception& e) {
        return fail(ctx, YAR_ERR_INTERNAL, std::string(name) + ": " + e.what());
    } catch (...) {
        return fail(ctx, YAR_ERR_INTERNAL, std::string(name) + ": unknown exception");
    }
}

static void m34_rows(const float* m, float4& r0, float4& r1, float4& r2) {
    r0 = make_float4(m[0], m[1], m[2], m[3]);
    r1 = make_float4(m[4], m[5], m[6], m[7]);
    r2 = make_float4(m[8], m[9], m[10], m[11]);
}

static float pbrt3_remap(float r, const yar_render_config& cfg, bool is_alpha) {   // scattering.cpp:240-269
    if (is_alpha) return r;
    if (cfg.pbrt3_scene) {
        r = std::max(r, 1e-3f);
        float x = std::log(r);
        return 1.621420000f + 0.819955000f * x + 0.173400000f * x * x + 0.017120100f * x * x * x +
               0.000640711f * x * x * x * x;
    }
    return r * r;
}

static void upload_scene(yar_gpu_ctx* ctx, const yar_scene_desc* d) {
    using namespace yar;
    auto t0 = std::chrono::steady_clock::now();
    // every host core builds mesh trees, as the reference's per-mesh build does (scene_context.cpp:33-61)
    int threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<KdBuildResult> trees = build_mesh_kdtrees(d, threads, ctx->opt.kdtree_builder == 1);
    // scene tree over object world bounds (build_scene_kdtree, kdtree_builder.cpp:443-463)
    std::vector<float> boxes(6 * (size_t)d->object_count);
    for (int i = 0; i < d->object_count; i++) {
        const yar_object_desc& o = d->objects[i];
        const KdBuildResult& t = trees[o.mesh];
        yar::M34 m;
        std::memcpy(m.a, o.object_to_world, sizeof(m.a));
        float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
        for (int k = 0; k < 8; k++) {
            F3 c((k & 4) ? t.bounds.hi.x : t.bounds.lo.x, (k & 2) ? t.bounds.hi.y : t.bounds.lo.y,
                 (k & 1) ? t.bounds.hi.z : t.bounds.lo.z);
            F3 w = yar::transform_point(m, c);
            for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], w[a]); hi[a] = std::max(hi[a], w[a]); }
        }
        for (int a = 0; a < 3; a++) { boxes[6 * (size_t)i + a] = lo[a]; boxes[6 * (size_t)i + 3 + a] = hi[a]; }
    }
    KdBuildResult stree = build_box_kdtree(boxes.data(), d->object_count);
    auto t1 = std::chrono::steady_clock::now();

    // flatten meshes
    std::vector<int64_t> mtri(d->mesh_count + 1, 0), mvtx(d->mesh_count + 1, 0), mnode(d->mesh_count + 1, 0);
    for (int i = 0; i < d->mesh_count; i++) {
        mtri[i + 1] = mtri[i] + d->meshes[i].triangle_count;
        mvtx[i + 1] = mvtx[i] + d->meshes[i].vertex_count;
        mnode[i + 1] = mnode[i] + (int64_t)(trees[i].words.size() / 2);
    }
    if (mnode[d->mesh_count] >= (int64_t)0x40000000 || mtri[d->mesh_count] >= (int64_t)0x7fffffff)
        throw HipError{"scene exceeds 2^30 kd nodes or 2^31 triangles", YAR_ERR_CAPACITY};
    std::vector<float4> tv(3 * (size_t)mtri[d->mesh_count]);
    std::vector<int4> tvi(mtri[d->mesh_count]);
    std::vector<float4> vn(std::max<int64_t>(1, mvtx[d->mesh_count]), make_float4(0, 0, 0, 0));
    std::vector<float2> vuv(std::max<int64_t>(1, mvtx[d->mesh_count]), make_float2(0, 0));
    std::vector<int4> minfo(d->mesh_count);
    std::vector<uint2> mn(std::max<int64_t>(1, mnode[d->mesh_count]));
    int mesh_depth = 0;
    for (int i = 0; i < d->mesh_count; i++) {
        const yar_mesh_desc& m = d->meshes[i];
        const int mflags = (m.normals ? 1 : 0) | (m.uvs ? 2 : 0) | (m.reverse_geometric_normal_orientation ? 4 : 0);
        for (int t = 0; t < m.triangle_count; t++) {
            int64_t g = mtri[i] + t;
            int vi[3];
            for (int k = 0; k < 3; k++) {
                vi[k] = m.indices[3 * (size_t)t + k];
                const float* p = m.vertices + 3 * (size_t)vi[k];
                tv[3 * (size_t)g + k] = make_float4(p[0], p[1], p[2], 0.f);
            }
            tvi[g] = make_int4((int)(mvtx[i] + vi[0]), (int)(mvtx[i] + vi[1]), (int)(mvtx[i] + vi[2]), mflags);
        }
        for (int v = 0; v < m.vertex_count; v++) {
            if (m.normals) vn[mvtx[i] + v] = make_float4(m.normals[3 * v], m.normals[3 * v + 1], m.normals[3 * v + 2], 0.f);
            if (m.uvs) vuv[mvtx[i] + v] = make_float2(m.uvs[2 * v], m.uvs[2 * v + 1]);
        }
        minfo[i] = make_int4(mflags, 0, 0, 0);
        // rebase the mesh tree: above_child and leaf triangle ids become global indices
        const std::vector<uint32_t>& w = trees[i].words;
        size_t nn = w.size() / 2;
        uint32_t nb = (uint32_t)mnode[i], tb = (uint32_t)mtri[i];
        size_t j = 0;
        while (j < nn) {
            uint32_t w0 = w[2 * j], w1 = w[2 * j + 1];
            if ((w0 & 3u) != 3u) {
                mn[nb + j] = make_uint2((w0 & 3u) | (((w0 >> 2) + nb) << 2), w1);
                j++;
            } else {
                uint32_t cnt = w0 >> 2;
                size_t span = 1 + (cnt > 1 ? cnt / 2 : 0);
                std::vector<uint32_t> ws(w.begin() + 2 * j, w.begin() + 2 * (j + span));
                for (uint32_t q = 0; q < cnt; q++) ws[1 + q] += tb;
                for (size_t q = 0; q < span; q++) mn[nb + j + q] = make_uint2(ws[2 * q], ws[2 * q + 1]);
                j += span;
            }
        }
        mesh_depth = std::max(mesh_depth, trees[i].depth);
    }
    std::vector<uint2> sn(stree.words.size() / 2);
    for (size_t j = 0; j < sn.size(); j++) sn[j] = make_uint2(stree.words[2 * j], stree.words[2 * j + 1]);

    // objects
    std::vector<ObjTrav> ot(d->object_count);
    std::vector<float4> o2w(3 * (size_t)d->object_count), nrm(3 * (size_t)d->object_count);
    std::vector<int4> oinfo(d->object_count);
    ctx->tri_base.assign(d->object_count, 0);
    for (int i = 0; i < d->object_count; i++) {
        const yar_object_desc& o = d->objects[i];
        const KdBuildResult& t = trees[o.mesh];
        ObjTrav& r = ot[i];
        m34_rows(o.world_to_object, r.w2o0, r.w2o1, r.w2o2);
        uint32_t root = (uint32_t)mnode[o.mesh], tb = (uint32_t)mtri[o.mesh];
        float rootf, tbf;
        std::memcpy(&rootf, &root, 4);
        std::memcpy(&tbf, &tb, 4);
        r.lo_node = make_float4(t.bounds.lo.x, t.bounds.lo.y, t.bounds.lo.z, rootf);
        r.hi_tri = make_float4(t.bounds.hi.x, t.bounds.hi.y, t.bounds.hi.z, tbf);
        m34_rows(o.object_to_world, o2w[3 * i], o2w[3 * i + 1], o2w[3 * i + 2]);
        // object_to_world_normal_transform = transpose of world_to_object's 3x3 (scene_loader.cpp:56-62)
        const float* w2o = o.world_to_object;
        nrm[3 * i] = make_float4(w2o[0], w2o[4], w2o[8], 0.f);
        nrm[3 * i + 1] = make_float4(w2o[1], w2o[5], w2o[9], 0.f);
        nrm[3 * i + 2] = make_float4(w2o[2], w2o[6], w2o[10], 0.f);
        oinfo[i] = make_int4(o.material, o.area_light, o.mesh, (int)tb);
        ctx->tri_base[i] = (int32_t)tb;
    }
    std::vector<DevMaterial> mats(std::max(1, d->material_count));
    bool has_mix = false;
    for (int i = 0; i < d->material_count; i++) {
        const yar_material_desc& m = d->materials[i];
        DevMaterial& g = mats[i];
        std::memset(&g, 0, sizeof(g));
        g.type = m.type;
        g.eta_i = m.eta_i;
        g.alpha_x = pbrt3_remap(m.u_roughness, d->config, m.roughness_is_alpha != 0);
        g.alpha_y = pbrt3_remap(m.v_roughness, d->config, m.roughness_is_alpha != 0);
        g.refl = make_float4(m.reflectance[0], m.reflectance[1], m.reflectance[2], 0.f);
        g.trans = make_float4(m.transmittance[0], m.transmittance[1], m.transmittance[2], m.index_of_refraction);
        g.eta = make_float4(m.eta[0], m.eta[1], m.eta[2], 0.f);
        g.k = make_float4(m.k[0], m.k[1], m.k[2], 0.f);
        if (m.type == YAR_MATERIAL_PLASTIC)                      // trans.x = r0, trans.y = pbrt3 form
            g.trans = make_float4(m.specular_r0, d->config.pbrt3_scene ? 1.f : 0.f, 0.f, 0.f);
        if (m.type == YAR_MATERIAL_COATED_DIFFUSE)               // k = Fresnel r0 per channel
            g.k = make_float4(m.specular[0], m.specular[1], m.specular[2], 0.f);
        if (m.type == YAR_MATERIAL_GLASS && m.u_roughness > 0.f) g.type = kRoughGlassType;   // Rough_Glass_BSDF
        if (m.type == YAR_MATERIAL_PBRT3_UBER) {
            // refl = Kd | ior, trans = Ks | component code, eta = Kr, k = opacity; the components in
            // the reference's order (pbrt_scene.cpp:585-608): diffuse 0, specular 1, delta reflection 2,
            // opacity 4 (delta transmission 3 is not supported)
            if (m.transmittance[0] != 0.f || m.transmittance[1] != 0.f || m.transmittance[2] != 0.f)
                throw HipError{"uber delta transmission is not supported", YAR_ERR_UNSUPPORTED};
            uint32_t comps[4], n = 0, ndelta = 0;
            auto black = [](const float* c) { return c[0] == 0.f && c[1] == 0.f && c[2] == 0.f; };
            if (!black(m.reflectance)) comps[n++] = 0;
            if (!black(m.specular)) comps[n++] = 1;
            if (!black(m.delta_reflectance)) comps[n++] = 2;
            if (!(m.opacity[0] == 1.f && m.opacity[1] == 1.f && m.opacity[2] == 1.f)) comps[n++] = 4;
            if (n == 0) throw HipError{"uber material with no scattering component", YAR_ERR_INVALID_ARGUMENT};
            uint32_t code = n;
            for (uint32_t j = 0; j < n; j++) {
                code |= comps[j] << (4 + 4 * j);
                if (comps[j] >= 2) ndelta++;
            }
            code |= ndelta << 20;
            // stored as the integer's value (exact below 2^24), not its bit pattern: small integers'
            // bit patterns are FP32 denormals, which instructions that flush denormals would zero
            const float codef = (float)code;
            g.type = kUberType;
            g.refl = make_float4(m.reflectance[0], m.reflectance[1], m.reflectance[2], m.index_of_refraction);
            g.trans = make_float4(m.specular[0], m.specular[1], m.specular[2], codef);
            g.eta = make_float4(m.delta_reflectance[0], m.delta_reflectance[1], m.delta_reflectance[2], 0.f);
            g.k = make_float4(m.opacity[0], m.opacity[1], m.opacity[2], 0.f);
        }
        if (m.type == YAR_MATERIAL_MIX) {
            // refl = White - amount | 1 - luminance(White - amount), trans = sub indices (values) |
            // min(luminance(amount), 1) | 1 when a sub-BSDF transmits (Mix_BSDF, bsdf.cpp:497-506)
            float t = 0.f, delta_sub = -1.f;
            for (int j = 0; j < 2; j++) {
                const int k = m.mix_materials[j];
                if (k < 0 || k >= d->material_count) throw HipError{"mix sub-material index out of range", YAR_ERR_INVALID_ARGUMENT};
                const yar_material_desc& b = d->materials[k];
                const bool bsdf = b.type == YAR_MATERIAL_DIFFUSE || b.type == YAR_MATERIAL_METAL ||
                                  b.type == YAR_MATERIAL_PLASTIC || b.type == YAR_MATERIAL_COATED_DIFFUSE ||
                                  (b.type == YAR_MATERIAL_GLASS && b.u_roughness > 0.f);
                const bool delta = b.type == YAR_MATERIAL_MIRROR || (b.type == YAR_MATERIAL_GLASS && b.u_roughness <= 0.f);
                if (!bsdf && !delta) throw HipError{"mix sub-materials must be BSDF or delta materials (no uber or mix)", YAR_ERR_UNSUPPORTED};
                if (delta) {
                    if (delta_sub >= 0.f) throw HipError{"mix of two delta sub-materials is not supported", YAR_ERR_UNSUPPORTED};
                    delta_sub = (float)j;
                }
                if (bsdf && b.type == YAR_MATERIAL_GLASS) t = 1.f;
            }
            auto lum = [](float r, float g, float bl) { return 0.2126729f * r + 0.7151522f * g + 0.0721750f * bl; };
            const float c0 = 1.f - m.mix_amount[0], c1 = 1.f - m.mix_amount[1], c2 = 1.f - m.mix_amount[2];
            const float f0 = (float)m.mix_materials[0], f1 = (float)m.mix_materials[1];   // exact values
            g.type = kMixType;
            g.refl = make_float4(c0, c1, c2, 1.f - lum(c0, c1, c2));
            g.trans = make_float4(f0, f1, std::min(lum(m.mix_amount[0], m.mix_amount[1], m.mix_amount[2]), 1.f), t);
            // Scaled_BSDF factor of the BSDF sub-material next to a delta one (bsdf.cpp:617-624)
            const float lum_amount = lum(m.mix_amount[0], m.mix_amount[1], m.mix_amount[2]);
            const float scale = delta_sub == 0.f ? 1.f - lum_amount : lum_amount;
            g.eta = make_float4(delta_sub, delta_sub >= 0.f ? scale : 0.f, 0.f, 0.f);
            g.k = make_float4(0.f, 0.f, 0.f, 0.f);
            has_mix = true;
        }
        if (m.type < YAR_MATERIAL_DIFFUSE || m.type > YAR_MATERIAL_MIX) throw HipError{"unsupported material type", YAR_ERR_UNSUPPORTED};
    }
    std::vector<DevRect> rl(std::max(1, d->rect_light_count));
    for (int i = 0; i < d->rect_light_count; i++) {
        const yar_rect_light_desc& l = d->rect_lights[i];
        m34_rows(l.light_to_world, rl[i].l0, rl[i].l1, rl[i].l2);
        rl[i].L = make_float4(l.emitted_radiance[0], l.emitted_radiance[1], l.emitted_radiance[2], 0.f);
        rl[i].size = make_float4(l.size[0], l.size[1], ctx->opt.disable_mis_cull ? 1.f : 0.f, 0.f);
    }
    // The analytic MIS-ray test assumes every emitter triangle lies inside its light's rectangle; lights
    // whose world-space emitter vertices fall outside it (by more than 1e-4 of the size) trace every ray.
    for (int i = 0; i < d->object_count; i++) {
        const yar_object_desc& o = d->objects[i];
        if (o.area_light < 0 || o.area_light >= d->rect_light_count) continue;
        DevRect& R = rl[o.area_light];
        yar::M34 m;
        std::memcpy(m.a, o.object_to_world, sizeof(m.a));
        const F3 c(R.l0.w, R.l1.w, R.l2.w), ex(R.l0.x, R.l1.x, R.l2.x), ey(R.l0.y, R.l1.y, R.l2.y),
            ez(R.l0.z, R.l1.z, R.l2.z);
        const float hx = 0.5f * R.size.x * (1.f + 1e-4f) + 1e-6f, hy = 0.5f * R.size.y * (1.f + 1e-4f) + 1e-6f;
        const float hz = 1e-4f * std::max(R.size.x, R.size.y) + 1e-6f;
        const yar_mesh_desc& mesh = d->meshes[o.mesh];
        for (int64_t k = 0; k < 3 * (int64_t)mesh.triangle_count && R.size.z == 0.f; k++) {
            const float* v = mesh.vertices + 3 * (size_t)mesh.indices[k];
            const F3 q = yar::transform_point(m, F3(v[0], v[1], v[2])) - c;
            if (std::fabs(dot(q, ex)) > hx || std::fabs(dot(q, ey)) > hy || std::fabs(dot(q, ez)) > hz) R.size.z = 1.f;
        }
    }
    std::vector<DevPoint> pl(std::max(1, d->point_light_count));
    for (int i = 0; i < d->point_light_count; i++) {
        const yar_point_light_desc& l = d->point_lights[i];
        pl[i].p = make_float4(l.position[0], l.position[1], l.position[2], 0.f);
        pl[i].I = make_float4(l.intensity[0], l.intensity[1], l.intensity[2], 0.f);
    }

    std::vector<DevSpot> spl(std::max(1, d->spot_light_count));
    for (int i = 0; i < d->spot_light_count; i++) {
        const yar_spot_light_desc& l = d->spot_lights[i];
        const float cone_cos = std::cos(l.cone_angle);
        const float penumbra_cos = std::cos(std::max(0.f, l.cone_angle - l.penumbra_angle));
        spl[i].p = make_float4(l.position[0], l.position[1], l.position[2], 0.f);
        spl[i].dir = make_float4(l.direction[0], l.direction[1], l.direction[2], cone_cos);
        spl[i].I = make_float4(l.intensity[0], l.intensity[1], l.intensity[2], penumbra_cos);
    }
    std::vector<DevDirLight> dl(std::max(1, d->directional_light_count));
    for (int i = 0; i < d->directional_light_count; i++) {
        const yar_directional_light_desc& l = d->directional_lights[i];
        dl[i].dir = make_float4(l.direction[0], l.direction[1], l.direction[2], 0.f);
        dl[i].E = make_float4(l.irradiance[0], l.irradiance[1], l.irradiance[2], 0.f);
    }
    // triangle-mesh lights: triangle areas and their CDF in the reference's float order
    // (Triangle_Mesh::get_triangle_area / get_area, Distribution_1D::initialize; reference_renderer.cpp:137-152)
    std::vector<DevMeshLight> ml(std::max(1, d->mesh_light_count));
    std::vector<float> mcdf;
    for (int i = 0; i < d->mesh_light_count; i++) {
        const yar_mesh_light_desc& l = d->mesh_lights[i];
        const yar_object_desc& o = d->objects[l.object];
        const yar_mesh_desc& mesh = d->meshes[o.mesh];
        const int n = mesh.triangle_count;
        std::vector<float> areas((size_t)n);
        for (int t = 0; t < n; t++) {
            F3 p[3];
            for (int k = 0; k < 3; k++) {
                const float* v = mesh.vertices + 3 * (size_t)mesh.indices[3 * (size_t)t + k];
                p[k] = F3(v[0], v[1], v[2]);
            }
            areas[(size_t)t] = yar::length(yar::cross(p[1] - p[0], p[2] - p[0])) * 0.5f;
        }
        float area = 0.f, total = 0.f, sum = 0.f;
        for (int t = 0; t < n; t++) area += areas[(size_t)t];
        for (int t = 0; t < n; t++) total += areas[(size_t)t];
        ml[i].obj = l.object;
        ml[i].tri_base = (int)mtri[o.mesh];
        ml[i].tri_count = n;
        ml[i].cdf = (int)mcdf.size();
        ml[i].area = area;
        ml[i].L = make_float4(l.emitted_radiance[0], l.emitted_radiance[1], l.emitted_radiance[2], 0.f);
        for (int t = 0; t < n; t++) { sum += areas[(size_t)t]; mcdf.push_back(sum / total); }
    }
    if (mcdf.empty()) mcdf.push_back(1.f);
    std::vector<DevEnv> envl(1);
    std::memset(envl.data(), 0, sizeof(DevEnv));
    if (d->environment_light_count > 0) {
        const yar_environment_light_desc& l = d->environment_light[0];
        m34_rows(l.light_to_world, envl[0].l2w0, envl[0].l2w1, envl[0].l2w2);
        m34_rows(l.world_to_light, envl[0].w2l0, envl[0].w2l1, envl[0].w2l2);
        envl[0].L = make_float4(l.radiance[0], l.radiance[1], l.radiance[2], 0.f);
    }
    std::vector<DevSphere> sph(std::max(1, d->sphere_light_count));
    for (int i = 0; i < d->sphere_light_count; i++) {
        const yar_sphere_light_desc& l = d->sphere_lights[i];
        sph[i].p = make_float4(l.position[0], l.position[1], l.position[2], l.radius);
        sph[i].L = make_float4(l.emitted_radiance[0], l.emitted_radiance[1], l.emitted_radiance[2], 0.f);
    }

    hipStream_t s = ctx->stream;
    ctx->scene_nodes.upload(sn, s);
    ctx->mesh_nodes.upload(mn, s);
    ctx->objs.upload(ot, s);
    ctx->tri_verts.upload(tv, s);
    ctx->tri_vidx.upload(tvi, s);
    ctx->vtx_n.upload(vn, s);
    ctx->vtx_uv.upload(vuv, s);
    ctx->mesh_info.upload(minfo, s);
    ctx->obj_o2w.upload(o2w, s);
    ctx->obj_nrm.upload(nrm, s);
    ctx->obj_info.upload(oinfo, s);
    ctx->materials.upload(mats, s);
    ctx->rects.upload(rl, s);
    ctx->points.upload(pl, s);
    ctx->spots.upload(spl, s);
    ctx->dirs.upload(dl, s);
    ctx->mesh_lights.upload(ml, s);
    ctx->mesh_light_cdf.upload(mcdf, s);
    ctx->spheres.upload(sph, s);
    ctx->env.upload(envl, s);
    HIPX(hipStreamSynchronize(s));
    auto t2 = std::chrono::steady_clock::now();

    DevScene& S = ctx->scene;
    S.scene_nodes = ctx->scene_nodes.p;
    S.mesh_nodes = ctx->mesh_nodes.p;
    S.objs = ctx->objs.p;
    S.tri_verts = ctx->tri_verts.p;
    S.scene_lo = make_float4(stree.bounds.lo.x, stree.bounds.lo.y, stree.bounds.lo.z, 0.f);
    S.scene_hi = make_float4(stree.bounds.hi.x, stree.bounds.hi.y, stree.bounds.hi.z, 0.f);
    S.obj_o2w = ctx->obj_o2w.p;
    S.obj_nrm = ctx->obj_nrm.p;
    S.obj_info = ctx->obj_info.p;
    S.mesh_info = ctx->mesh_info.p;
    S.tri_vidx = ctx->tri_vidx.p;
    S.vtx_n = ctx->vtx_n.p;
    S.vtx_uv = ctx->vtx_uv.p;
    S.materials = ctx->materials.p;
    S.rects = ctx->rects.p;
    S.points = ctx->points.p;
    S.spots = ctx->spots.p;
    S.dirs = ctx->dirs.p;
    S.mesh_lights = ctx->mesh_lights.p;
    S.mesh_light_cdf = ctx->mesh_light_cdf.p;
    S.spheres = ctx->spheres.p;
    S.has_mix = has_mix ? 1 : 0;
    S.env = ctx->env.p;
    S.n_env = d->environment_light_count;
    S.n_rects = d->rect_light_count;
    S.n_points = d->point_light_count;
    S.n_spots = d->spot_light_count;
    S.n_dirs = d->directional_light_count;
    S.n_mesh_lights = d->mesh_light_count;
    S.n_spheres = d->sphere_light_count;
    S.n_lights = d->rect_light_count + d->point_light_count + d->spot_light_count + d->directional_light_count +
                 d->sphere_light_count + d->mesh_light_count + d->environment_light_count;                            // update_total_light_count, light.cpp:35

    yar_gpu_info& inf = ctx->info;
    inf.scene_bytes = (int64_t)(sn.size() * 8 + mn.size() * 8 + ot.size() * sizeof(ObjTrav) + tv.size() * 16 +
                                tvi.size() * 16 + vn.size() * 16 + vuv.size() * 8 + o2w.size() * 32 + oinfo.size() * 16);
    inf.triangle_count = mtri[d->mesh_count];
    inf.object_count = d->object_count;
    inf.scene_tree_depth = stree.depth;
    inf.mesh_tree_depth = mesh_depth;
    int req = ctx->opt.traversal_stack_depth;
    int sd, md;
    // Default: 8-entry short stacks per level (restart beyond).  Deeper LDS stacks cost occupancy
    // (a 32-entry stack pair is 64 KB per 128-thread block) and measured 2x slower on both the
    // 2M-triangle and the 2k-triangle scenes, with bit-identical results.
    if (req <= 0) req = 8;
    sd = std::min(pow2_at_least(req), pow2_at_least(std::max(1, stree.depth)));
    md = std::min(pow2_at_least(req), pow2_at_least(std::max(1, mesh_depth)));
    inf.stack_depth = (sd << 16) | md;
    inf.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    inf.upload_seconds = std::chrono::duration<double>(t2 - t1).count();
}

// pcg32 jump-ahead tables (see pcg_jump): lo[r] = (M^r, G_r) for r < kJumpLo and hi[q] = (M^qL, G_qL)
// with L = kJumpLo, where G_(n+1) = M G_n + 1 and G_0 = 0.
static void upload_jump_tables(yar_gpu_ctx* ctx) {
    std::vector<ulonglong2> lo(kJumpLo), hi(kJumpHi);
    uint64_t A = 1, G = 0;
    for (uint32_t r = 0; r < kJumpLo; r++) {
        lo[r] = make_ulonglong2(A, G);
        G = kPcgMult * G + 1;
        A = kPcgMult * A;
    }
    const uint64_t AL = A, GL = G;
    uint64_t a = 1, g = 0;
    for (uint32_t q = 0; q < kJumpHi; q++) {
        hi[q] = make_ulonglong2(a, g);
        g = AL * g + GL;
        a = AL * a;
    }
    ctx->jump_lo.upload(lo, ctx->stream);
    ctx->jump_hi.upload(hi, ctx->stream);
    HIPX(hipStreamSynchronize(ctx->stream));
}

// Film layout of a view (Film::Film, film.cpp:52-71; box filter radius 0.5 -> sample region ==
// render region) and the bounds of one tile (Film::get_tile_bounds, film.cpp:73-86).
struct TileGeom {
    int x0, y0, x1, y1;      // sample bounds
    int px0, py0, px1, py1;  // pixel bounds (film.cpp:80-85)
};

static void film_layout(const yar_view& v, int* tgx, int* tgy, int* sx0, int* sy0, int* sx1, int* sy1) {
    *sx0 = (int)std::floor(v.region_x0 + 0.5f - 0.5f);
    *sy0 = (int)std::floor(v.region_y0 + 0.5f - 0.5f);
    *sx1 = (int)std::ceil(v.region_x1 - 1 + 0.5f + 0.5f);
    *sy1 = (int)std::ceil(v.region_y1 - 1 + 0.5f + 0.5f);
    *tgx = (*sx1 - *sx0 + 63) / 64;
    *tgy = (*sy1 - *sy0 + 63) / 64;
}

static TileGeom tile_geom(const yar_view& v, int tile) {
    int tgx, tgy, sx0, sy0, sx1, sy1;
    film_layout(v, &tgx, &tgy, &sx0, &sy0, &sx1, &sy1);
    TileGeom g;
    int tx = tile % tgx, ty = tile / tgx;
    g.x0 = sx0 + tx * 64; g.y0 = sy0 + ty * 64;
    g.x1 = std::min(g.x0 + 64, sx1); g.y1 = std::min(g.y0 + 64, sy1);
    g.px0 = std::max((int)std::ceil(g.x0 - 0.5f - 0.5f), v.region_x0);
    g.py0 = std::max((int)std::ceil(g.y0 - 0.5f - 0.5f), v.region_y0);
    g.px1 = std::min((int)std::floor(g.x1 + 0.5f - 0.5f) + 1, v.region_x1);
    g.py1 = std::min((int)std::floor(g.y1 + 0.5f - 0.5f) + 1, v.region_y1);
    return g;
}

// A render job: tiles of one or more views.  The pixel list holds the parts in order, tiles in
// the given order inside a part, pixels row-major inside a tile's sample bounds (render_tile
// visitation order, reference_renderer.cpp:346-347).
struct JobPart {
    int view;
    std::vector<int> tiles;
};

struct PartResult {
    std::vector<std::vector<yar_film_pixel>> tiles;   // per requested tile, pixel-bounds sized
    std::vector<double> tile_var;
};

static DevView make_dev_view(const yar_view& v, int tile_base) {
    DevView d{};
    m34_rows(v.camera_to_world, d.c0, d.c1, d.c2);
    d.W = (float)v.film_width;
    d.H = (float)v.film_height;
    constexpr float deg2rad = yar::kPi / 180.f;                                 // camera.cpp:5-14
    const float tan_half = std::tan((v.fov_y_degrees / 2.f) * deg2rad);
    d.vert = tan_half;
    d.horz = (d.W / d.H) * tan_half;
    d.zup = v.z_is_up;
    d.nx = v.x_pixel_sample_count;
    d.ny = v.y_pixel_sample_count;
    d.spp = d.nx * d.ny;
    d.seed_offset = v.rng_seed_offset;
    d.tile_base = tile_base;
    int tgy;
    film_layout(v, &d.tgx, &tgy, &d.sx0, &d.sy0, &d.sx1, &d.sy1);
    d.rx0 = v.region_x0; d.ry0 = v.region_y0; d.rx1 = v.region_x1; d.ry1 = v.region_y1;
    return d;
}

// Bounce-loop iterations that always suffice: every shade pass raises a path's bounce count by at
// least one except the first pass after a delta (glass) event, and paths end at max_light_bounces.
static int loop_iterations(int max_bounces) { return std::max(1, max_bounces) + 2; }

// Guided render: the learned proposal, its pass schedule and training-record storage.
struct GuidedRun {
    yar_guide* guide = nullptr;
    std::vector<std::pair<int, int>> passes;      // sample ranges [k0, k1)
};

// `first` samples per pixel with BSDF sampling only, then the rest in count - 1 passes of doubling
// size, each guided by the proposal trained on the records of the passes before it.  Every pass
// runs the whole bounce loop, and a small pass leaves the GPU idle in the long tail of each
// iteration, so the default is two passes.
static std::vector<std::pair<int, int>> guided_passes(int max_spp, int first, int count) {
    std::vector<std::pair<int, int>> p;
    const int k1 = std::min(max_spp, std::max(1, first));
    p.push_back({0, k1});
    const int rest = max_spp - k1, n = std::max(1, count - 1);
    // sizes proportional to 1, 2, 4, ... 2^(n-1) over the remaining samples
    const double unit = (double)rest / (double)((1ll << std::min(n, 30)) - 1);
    int k0 = k1;
    for (int i = 0; i < n && k0 < max_spp; i++) {
        int k = i + 1 == n ? max_spp : k1 + (int)std::llround(unit * (double)((1ll << (i + 1)) - 1));
        k = std::max(k, k0 + 1);
        k = std::min(k, max_spp);
        p.push_back({k0, k});
        k0 = k;
    }
    return p;
}

static void render_job(yar_gpu_ctx* ctx, const std::vector<JobPart>& parts, std::vector<PartResult>& out,
                       yar_render_stats& st, GuidedRun* gr = nullptr) {
    auto t_start = std::chrono::steady_clock::now();
    const yar_render_config& rc = ctx->cfg;
    if (rc.pixel_filter_radius != 0.5f) throw HipError{"only the box filter of radius 0.5 is supported", YAR_ERR_UNSUPPORTED};
    const bool guided = gr != nullptr;
    const int mdim = std::min(10, rc.max_light_bounces);
    const int d1 = 4 * std::max(0, mdim), d2 = 3 * std::max(0, mdim);          // reference_renderer.cpp:87-99
    const int n_dims = d1 + d2;                                               // shuffled sampler runs per pixel
    // RNG fallback past 10 bounces couples a pixel's samples; guided passes draw those dimensions
    // from each path's own stream instead
    const bool sequential = rc.max_light_bounces > 10 && !guided;

    // views, tile offsets and the pixel list
    std::vector<DevView> dv;
    std::vector<int> tile_first;
    int64_t total_pixels = 0;
    struct Unit { int part, tile; int64_t start, count; };
    std::vector<Unit> units;
    int job_max_spp = 1;
    for (size_t pi = 0; pi < parts.size(); pi++) {
        const yar_view& v = ctx->views[parts[pi].view];
        const int spp = v.x_pixel_sample_count * v.y_pixel_sample_count;
        if (spp <= 0 || spp > 65536) throw HipError{"samples per pixel out of range", YAR_ERR_INVALID_ARGUMENT};
        job_max_spp = std::max(job_max_spp, spp);
        int tgx, tgy, sx0, sy0, sx1, sy1;
        film_layout(v, &tgx, &tgy, &sx0, &sy0, &sx1, &sy1);
        const int base = (int)tile_first.size();
        dv.push_back(make_dev_view(v, base));
        tile_first.resize(tile_first.size() + (size_t)tgx * tgy, -1);
        for (int t : parts[pi].tiles) {
            const TileGeom g = tile_geom(v, t);
            const int64_t count = (int64_t)(g.x1 - g.x0) * (g.y1 - g.y0);
            if (total_pixels + count > (int64_t)0x7fffffff)
                throw HipError{"more than 2^31 pixels in one job", YAR_ERR_CAPACITY};
            tile_first[base + t] = (int)total_pixels;
            units.push_back({(int)pi, t, total_pixels, count});
            total_pixels += count;
        }
    }
    if (guided) gr->passes = guided_passes(job_max_spp, guide_config(gr->guide).first_pass_spp, guide_config(gr->guide).pass_count);
    int pass_max = 1;                                                          // widest guided pass
    if (guided)
        for (auto& pr : gr->passes) pass_max = std::max(pass_max, pr.second - pr.first);

    // chunks of whole tiles, bounded by paths in flight, sample-table and radiance sizes.
    // Plain jobs default to 64M paths in flight: the persistent traversal kernels need wide queues in
    // the late bounces, where most paths have ended (config 2, 1024x1024 at 64 spp: 262 Msamples/s at
    // 4M paths, 329 at 8M, 390 at 16M, 416 at 32M in two chunks, 449 at 64M in one chunk).  Guided
    // jobs revisit every chunk once per pass and keep up to 16M paths in one chunk.  Path state is
    // about 280 B per path (plain: 14 float4 rays / throughputs / direct-lighting halves / queue
    // payloads, hit, state and 5 queue slots; 64M paths hold about 19 GB) and about 800 B guided
    // (+ training vertices, network inputs and outputs).  The default never takes more than a
    // quarter of the free device memory.
    int64_t max_paths = ctx->opt.max_paths_in_flight > 0 ? ctx->opt.max_paths_in_flight : (int64_t)1 << (guided ? 24 : 26);
    if (ctx->opt.max_paths_in_flight <= 0) {
        size_t free_b = 0, total_b = 0;
        const int64_t per_path = guided ? 800 : 280;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
            max_paths = std::max<int64_t>((int64_t)1 << 16, std::min<int64_t>(max_paths, (int64_t)(free_b / 4) / per_path));
    }
    const int perm_bytes = job_max_spp <= 256 ? 1 : 2;
    const int64_t draws = 2 * (int64_t)job_max_spp + d1 * (2 * (int64_t)job_max_spp - 1) + d2 * (3 * (int64_t)job_max_spp - 1);
    if (draws >= (int64_t)kJumpLo * kJumpHi) throw HipError{"samples per pixel too large for the sampler's jump tables", YAR_ERR_CAPACITY};
    // sample table of a chunk at most 8 GB (config 2: 67M samples x 56 dimensions x 1 B = 3.8 GB)
    const int64_t max_samples = std::min<int64_t>(((int64_t)1 << 33) / std::max(1, n_dims * perm_bytes), (int64_t)1 << 27);
    struct ChunkInfo { int64_t start, pixels, samples, paths; int max_spp, min_spp; };
    std::vector<ChunkInfo> cks;
    {
        ChunkInfo cur{0, 0, 0, 0, 1, 1 << 30};
        for (const Unit& u : units) {
            const int spp = dv[u.part].spp;
            const int64_t us = u.count * spp;
            const int64_t up = guided ? u.count * pass_max : sequential ? u.count : us;
            if (cur.pixels > 0 && (cur.paths + up > max_paths || cur.samples + us > max_samples)) {
                cks.push_back(cur);
                cur = ChunkInfo{u.start, 0, 0, 0, 1, 1 << 30};
            }
            cur.pixels += u.count;
            cur.samples += us;
            cur.paths += up;
            cur.max_spp = std::max(cur.max_spp, spp);
            cur.min_spp = std::min(cur.min_spp, spp);
        }
        if (cur.pixels > 0) cks.push_back(cur);
    }
    // chunk-local spp prefix sums, one (n_pixels + 1)-array per chunk: chunk ci's array starts at
    // entry ck.start + ci
    std::vector<size_t> pf_off;
    int64_t max_chunk_pixels = 1, max_chunk_samples = 1, max_chunk_paths = 1;
    for (size_t ci = 0; ci < cks.size(); ci++) {
        const ChunkInfo& ck = cks[ci];
        pf_off.push_back((size_t)ck.start + ci);
        max_chunk_pixels = std::max(max_chunk_pixels, ck.pixels);
        max_chunk_samples = std::max(max_chunk_samples, ck.samples);
        const int64_t paths = guided ? ck.pixels * pass_max : ck.paths;
        max_chunk_paths = std::max(max_chunk_paths, paths);
    }
    const int n_passes = guided ? (int)gr->passes.size() : 1;
    const int64_t side_cap = std::max<int64_t>(
        1 << 16, std::min<int64_t>((int64_t)1 << 24, max_chunk_samples / 64 * (int64_t)cks.size() * n_passes));

    // buffers
    hipStream_t s = ctx->stream;
    const size_t np = (size_t)max_chunk_paths;
    ctx->perm.reserve((size_t)(max_chunk_samples * n_dims * perm_bytes));
    ctx->radiance.reserve((size_t)max_chunk_samples);
    ctx->spill.reserve((size_t)max_chunk_pixels);
    DevBuf<float4>* pb[] = {&ctx->ray_o, &ctx->ray_d, &ctx->coeff, &ctx->L, &ctx->dl_coeff, &ctx->dl_light, &ctx->dl_bsdf,
                            &ctx->hit_tb, &ctx->sh_o, &ctx->sh_d, &ctx->sh_c, &ctx->mi_o, &ctx->mi_d, &ctx->mi_f};
    for (auto* b : pb) b->reserve(np);
    ctx->hit_ot.reserve(np);
    ctx->pk.reserve(np);
    ctx->state.reserve(np);
    if (sequential) ctx->rng.reserve(np);
    DevBuf<uint32_t>* qb[] = {&ctx->q_a, &ctx->q_b, &ctx->sh_path, &ctx->mi_path, &ctx->dl_path};
    for (auto* b : qb) b->reserve(np);
    ctx->mi_ndw.reserve(np);
    ctx->counts.reserve(16);
    const int bin_mode = std::min(std::max(ctx->opt.bin_rays, 0), 3);
    const bool bin_rays = bin_mode > 0;
    if (bin_rays) {
        ctx->q_unsorted.reserve(np);
    }
    ctx->totals.reserve(24);
    ctx->out_film.reserve((size_t)std::max<int64_t>(1, total_pixels));
    ctx->out_var.reserve((size_t)std::max<int64_t>(1, total_pixels));
    ctx->side_rad.reserve((size_t)side_cap);
    ctx->side_info.reserve((size_t)side_cap);
    ctx->side_view.reserve((size_t)side_cap);
    ctx->chunk_slot.reserve(1);
    int64_t train_cap = 0;
    if (guided) {
        train_cap = std::max<int64_t>(65536, guide_config(gr->guide).record_capacity);
        ctx->grng.reserve(np);
        ctx->rec.reserve(np * kRecMax * kRecFloats);
        ctx->guide_path.reserve(np);
        ctx->guide_in.reserve(np * 16);
        ctx->guide_out.reserve(np * kGOut);
        ctx->train.reserve((size_t)train_cap);
        ctx->m1.reserve((size_t)std::max<int64_t>(1, total_pixels));
        ctx->m2.reserve((size_t)std::max<int64_t>(1, total_pixels));
        guide_prepare_device();
    }
    {
        // pixel list, pixel views and prefix sums written in place into pinned staging by all host
        // cores, one tile per task, then sent at link speed
        const size_t n = (size_t)total_pixels, n_pf = n + cks.size();
        const size_t o_view = (n * sizeof(int2) + 255) & ~(size_t)255;
        const size_t o_pf = (o_view + n * sizeof(int) + 255) & ~(size_t)255;
        const size_t bytes = o_pf + n_pf * sizeof(long long);
        HIPX(hipStreamSynchronize(s));                    // the previous job's sends have left the staging
        ctx->h_stage.reserve(bytes);
        int2* hpix = (int2*)ctx->h_stage.p;
        int* hview = (int*)(ctx->h_stage.p + o_view);
        long long* hpf = (long long*)(ctx->h_stage.p + o_pf);
        // each unit's chunk and the prefix sum at its first pixel
        std::vector<size_t> u_chunk(units.size());
        std::vector<long long> u_acc(units.size());
        {
            size_t ci = 0;
            long long acc = 0;
            for (size_t ui = 0; ui < units.size(); ui++) {
                while (ci + 1 < cks.size() && units[ui].start >= cks[ci + 1].start) { ci++; acc = 0; }
                u_chunk[ui] = ci;
                u_acc[ui] = acc;
                acc += (long long)units[ui].count * dv[units[ui].part].spp;
            }
        }
        parallel_for(units.size(), [&](size_t ui) {
            const Unit& u = units[ui];
            const TileGeom g = tile_geom(ctx->views[parts[u.part].view], u.tile);
            const ChunkInfo& ck = cks[u_chunk[ui]];
            const long long spp = dv[u.part].spp;
            int64_t off = u.start;
            long long* pfc = hpf + pf_off[u_chunk[ui]] - ck.start;
            long long acc = u_acc[ui];
            for (int y = g.y0; y < g.y1; y++)
                for (int x = g.x0; x < g.x1; x++, off++, acc += spp) {
                    hpix[off] = make_int2(x, y);
                    hview[off] = u.part;
                    pfc[off] = acc;
                }
        });
        for (size_t ci = 0; ci < cks.size(); ci++) hpf[pf_off[ci] + cks[ci].pixels] = cks[ci].samples;
        ctx->pixels.reserve(n);
        ctx->pix_view.reserve(n);
        ctx->pix_first.reserve(n_pf);
        if (n > 0) {
            HIPX(hipMemcpyAsync(ctx->pixels.p, hpix, n * sizeof(int2), hipMemcpyHostToDevice, s));
            HIPX(hipMemcpyAsync(ctx->pix_view.p, hview, n * sizeof(int), hipMemcpyHostToDevice, s));
        }
        if (n_pf > 0) HIPX(hipMemcpyAsync(ctx->pix_first.p, hpf, n_pf * sizeof(long long), hipMemcpyHostToDevice, s));
    }
    ctx->tile_first.upload(tile_first, s);
    ctx->dviews.upload(dv, s);

    // per (chunk, pass) parameter blocks; guided jobs run every chunk once per pass, plain jobs run
    // a chunk's passes (one, or one per sample in sequential mode) back to back
    std::vector<Chunk> hc;
    struct Step { int chunk, pass; size_t hc; };
    std::vector<Step> steps;
    auto make_chunk = [&](size_t ci) {
        const ChunkInfo& ck = cks[ci];
        Chunk c{};
        c.pixels = ctx->pixels.p + ck.start;
        c.pix_view = ctx->pix_view.p + ck.start;
        c.pix_first = ctx->pix_first.p + pf_off[ci];
        c.views = ctx->dviews.p;
        c.perm = ctx->perm.p;
        c.jt = JumpTables{ctx->jump_lo.p, ctx->jump_hi.p};
        c.n_pixels = (int)ck.pixels;
        c.max_spp = ck.max_spp;
        c.uniform_spp = ck.min_spp == ck.max_spp ? ck.max_spp : 0;
        c.d1 = d1; c.d2 = d2; c.perm_bytes = perm_bytes;
        c.sequential = sequential ? 1 : 0;
        c.n_paths = ck.paths;
        c.max_bounces = rc.max_light_bounces;
        c.rr_bounce_threshold = rc.russian_roulette_bounce_count_threshold;
        c.rr_threshold = rc.russian_roulette_threshold;
        c.film_scale = rc.film_radiance_scale;
        c.max_component = rc.max_rgb_component_value_of_film_sample;
        return c;
    };
    if (guided) {
        for (int pass = 0; pass < n_passes; pass++)
            for (size_t ci = 0; ci < cks.size(); ci++) {
                Chunk c = make_chunk(ci);
                c.guided = 1;
                c.k0 = gr->passes[pass].first;
                c.k1 = gr->passes[pass].second;
                c.alpha = pass == 0 ? 0.f : guide_config(gr->guide).guide_fraction;
                c.n_paths = (long long)c.n_pixels * (c.k1 - c.k0);
                if (c.k0 >= c.max_spp) continue;                   // no pixel of this chunk has such samples
                steps.push_back({(int)ci, pass, hc.size()});
                hc.push_back(c);
            }
    } else {
        for (size_t ci = 0; ci < cks.size(); ci++) {
            const int passes = sequential ? cks[ci].max_spp : 1;
            for (int pass = 0; pass < passes; pass++) {
                Chunk c = make_chunk(ci);
                c.pass = pass;
                steps.push_back({(int)ci, pass, hc.size()});
                hc.push_back(c);
            }
        }
    }
    ctx->chunks.upload(hc, s);
    HIPX(hipMemsetAsync(ctx->counts.p, 0, 16 * sizeof(uint32_t), s));
    HIPX(hipMemsetAsync(ctx->totals.p, 0, 24 * sizeof(unsigned long long), s));

    Paths P{};
    P.ray_o = ctx->ray_o.p; P.ray_d = ctx->ray_d.p; P.coeff = ctx->coeff.p; P.L = ctx->L.p; P.state = ctx->state.p;
    P.dl_coeff = ctx->dl_coeff.p; P.dl_light = ctx->dl_light.p; P.dl_bsdf = ctx->dl_bsdf.p; P.hit_tb = ctx->hit_tb.p;
    P.hit_ot = ctx->hit_ot.p; P.pk = ctx->pk.p; P.rng = sequential ? ctx->rng.p : nullptr; P.radiance = ctx->radiance.p;
    P.grng = guided ? ctx->grng.p : nullptr;
    P.rec = guided ? ctx->rec.p : nullptr;
    Queues q{};
    q.active0 = ctx->q_a.p; q.active1 = ctx->q_b.p;
    q.shadow_path = ctx->sh_path.p; q.shadow_o = ctx->sh_o.p; q.shadow_d = ctx->sh_d.p; q.shadow_c = ctx->sh_c.p;
    q.mis_path = ctx->mi_path.p; q.mis_o = ctx->mi_o.p; q.mis_d = ctx->mi_d.p; q.mis_f = ctx->mi_f.p;
    q.mis_ndw = ctx->mi_ndw.p; q.dl_path = ctx->dl_path.p; q.counts = ctx->counts.p; q.totals = ctx->totals.p;
    q.count_culls = ctx->opt.enable_counters != 0;
    if (bin_rays) {
        q.unsorted = ctx->q_unsorted.p;
        q.bin_rays = bin_mode;
    }
    if (guided) {
        q.guide_path = ctx->guide_path.p;
        q.guide_in = ctx->guide_in.p;
        q.guide_out = ctx->guide_out.p;
        q.train = ctx->train.p;
        q.train_cap = (uint32_t)train_cap;
    }
    FilmParams f{};
    f.out_film = ctx->out_film.p;
    f.out_var = ctx->out_var.p;
    f.spill = ctx->spill.p;
    f.tile_first = ctx->tile_first.p;
    f.side_rad = ctx->side_rad.p;
    f.side_info = ctx->side_info.p;
    f.side_view = ctx->side_view.p;
    f.side_capacity = (int)side_cap;
    if (guided) {
        f.accumulate = 1;
        f.m1 = ctx->m1.p;
        f.m2 = ctx->m2.p;
    }
    LaunchCfg lc = ctx->launch_cfg();
    lc.trav_grid = ctx->cu_count * traversal_blocks_per_cu(lc, 0);
    lc.shadow_grid = ctx->cu_count * traversal_blocks_per_cu(lc, 1);
    lc.mis_grid = ctx->cu_count * traversal_blocks_per_cu(lc, 2);
    lc.path_grid = (int)std::max<int64_t>(1, std::min<int64_t>((max_chunk_paths + 255) / 256, (int64_t)ctx->cu_count * 8));
    if (guided) {
        lc.guided = true;
        lc.net = guide_net(gr->guide);
        lc.guide_grid = ctx->cu_count;
    }
    const int iters = loop_iterations(rc.max_light_bounces);
    const bool use_graph = ctx->opt.use_graph != 0 && !guided;
    Chunk* slot = ctx->chunk_slot.p;

    // polling slots: one per iteration (eager) or per (chunk, pass) completion check (graph)
    const size_t poll_n = std::max<size_t>(hc.size() + 1, (size_t)iters + 1) + 1;
    if (ctx->h_poll_n < poll_n) {
        if (ctx->h_poll) (void)hipHostFree(ctx->h_poll);
        ctx->h_poll = nullptr;
        ctx->h_poll_n = 0;
        HIPX(hipHostMalloc((void**)&ctx->h_poll, poll_n * sizeof(uint32_t), hipHostMallocDefault));
        ctx->h_poll_n = poll_n;
    }
    uint32_t* hp = ctx->h_poll;

    // the captured loop bakes in buffer addresses and sizes: capture again when any changed
    const bool graph_stale = ctx->graph_paths != max_chunk_paths || ctx->graph_iters != iters ||
                             std::memcmp(&ctx->graph_P, &P, sizeof(P)) != 0 || std::memcmp(&ctx->graph_q, &q, sizeof(q)) != 0;
    if (use_graph && (graph_stale || !ctx->graph)) {
        ctx->drop_graph();
        hipGraph_t g;
        HIPX(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch_raygen(slot, max_chunk_paths, P, q, s);
        for (int it = 0; it < iters; it++) launch_iteration(ctx->scene, slot, P, q, it & 1, lc);
        HIPX(hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { ctx->graph = nullptr; throw HipError{std::string("hipGraphInstantiate: ") + hipGetErrorString(e)}; }
        ctx->graph_paths = max_chunk_paths;
        ctx->graph_iters = iters;
        ctx->graph_P = P;
        ctx->graph_q = q;
    }

    KernelTimer kt;
    kt.on = ctx->opt.enable_timing != 0 && !use_graph;
    kt.s = s;
    std::vector<hipEvent_t> ev((size_t)iters + 1);
    for (auto& e : ev) HIPX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    HIPX(hipEventRecord(e0, s));
    bool loop_ok = true;
    int64_t train_records = 0;
    double train_seconds = 0.0, train_nats = 0.0;
    auto run_pass_loop = [&](const Step& sp) {
        const Chunk& c = hc[sp.hc];
        if (use_graph) {
            HIPX(hipGraphLaunch(ctx->graph, s));
            // the active list the last iteration wrote must be empty
            HIPX(hipMemcpyAsync(hp + sp.hc, ctx->counts.p + (iters & 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            return;
        }
        kt.begin(4);
        launch_raygen(slot, c.n_paths, P, q, s);
        kt.end();
        // Queue kernels take their counts from device memory; the host reads each iteration's
        // active count asynchronously and stops enqueueing two iterations after it saw zero.
        int it = 0;
        for (; it < iters; it++) {
            if (it >= 2) {
                HIPX(hipEventSynchronize(ev[it - 2]));
                if (hp[it - 2] == 0) break;
            }
            const int parity = it & 1;
            if (kt.on) {
                static const int kind_of_stage[kStageCount] = {4, 0, 1, 5, 5, 4, 2, 3, 4};
                for (int stage = 0; stage < kStageCount; stage++) {
                    kt.begin(kind_of_stage[stage]);
                    launch_stage(stage, ctx->scene, slot, P, q, parity, lc);
                    kt.end();
                }
            } else {
                launch_iteration(ctx->scene, slot, P, q, parity, lc);
            }
            HIPX(hipMemcpyAsync(hp + it, ctx->counts.p + (parity ^ 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIPX(hipEventRecord(ev[it], s));
        }
        if (it == iters) {
            HIPX(hipEventSynchronize(ev[it - 1]));
            if (hp[it - 1] != 0) loop_ok = false;
        }
    };
    auto run_film = [&](const Step& sp) {
        f.base = cks[sp.chunk].start;
        kt.begin(4);
        launch_film(slot, (int)cks[sp.chunk].pixels, P, f, q, s);
        kt.end();
    };
    int table_chunk = -1;
    for (size_t si = 0; si < steps.size(); si++) {
        const Step& sp = steps[si];
        if (guided && (si == 0 || steps[si - 1].pass != sp.pass)) {
            // a new guided pass: keep about as many of its records as the store holds
            int64_t pass_samples = 0;
            const int k0 = gr->passes[sp.pass].first, k1 = gr->passes[sp.pass].second;
            for (const Unit& u : units) pass_samples += u.count * std::max(0, std::min(k1, dv[u.part].spp) - k0);
            q.keep_prob = (float)std::min(1.0, (double)train_cap / std::max<double>(1.0, 2.0 * pass_samples));
        }
        HIPX(hipMemcpyAsync(slot, ctx->chunks.p + sp.hc, sizeof(Chunk), hipMemcpyDeviceToDevice, s));
        // a chunk's table covers all its samples; regenerate it only when another chunk used the buffer
        if (sp.chunk != table_chunk) {
            table_chunk = sp.chunk;
            kt.begin(4);
            launch_sample_perms(hc[sp.hc], slot, s);
            kt.end();
        }
        run_pass_loop(sp);
        HIPX(hipGetLastError());
        const bool last_of_chunk = guided || si + 1 == steps.size() || steps[si + 1].chunk != sp.chunk;
        if (last_of_chunk) run_film(sp);
        if (guided && (si + 1 == steps.size() || steps[si + 1].pass != sp.pass)) {
            // end of a guided pass: train the proposal on the records it produced
            HIPX(hipMemcpyAsync(hp + poll_n - 1, ctx->counts.p + 10, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIPX(hipStreamSynchronize(s));
            const int64_t n_rec = std::min<int64_t>(hp[poll_n - 1], train_cap);
            train_records = hp[poll_n - 1];
            if (sp.pass + 1 < n_passes && n_rec > 0) {
                const auto t0 = std::chrono::steady_clock::now();
                train_nats = guide_train_on(gr->guide, ctx->train.p, n_rec, guide_config(gr->guide).steps_per_pass, s);
                train_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            }
        }
    }
    HIPX(hipEventRecord(e1, s));
    ctx->h_film.reserve((size_t)total_pixels * sizeof(float4));
    ctx->h_var.reserve((size_t)total_pixels * sizeof(double));
    const float4* film_own = (const float4*)ctx->h_film.p;
    const double* var_own = (const double*)ctx->h_var.p;
    if (total_pixels > 0) {
        HIPX(hipMemcpyAsync(ctx->h_film.p, ctx->out_film.p, total_pixels * sizeof(float4), hipMemcpyDeviceToHost, s));
        HIPX(hipMemcpyAsync(ctx->h_var.p, ctx->out_var.p, total_pixels * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    unsigned long long tot[24];
    uint32_t cnt[16];
    HIPX(hipMemcpyAsync(tot, ctx->totals.p, sizeof(tot), hipMemcpyDeviceToHost, s));
    HIPX(hipMemcpyAsync(cnt, ctx->counts.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    for (auto e : ev) (void)hipEventDestroy(e);
    if (use_graph)
        for (size_t i = 0; i < hc.size(); i++)
            if (hp[i] != 0) loop_ok = false;
    if (!loop_ok) throw HipError{"bounce loop did not terminate within max_light_bounces + 2 iterations"};
    double ksec[kTimerKinds];
    kt.collect(ksec);
    const uint32_t n_side = cnt[5];
    if ((int64_t)n_side > side_cap) throw HipError{"film side list overflow", YAR_ERR_CAPACITY};
    std::vector<float4> side_rad(n_side);
    std::vector<int4> side_info(n_side);
    std::vector<int> side_view(n_side);
    if (n_side) {
        HIPX(hipMemcpy(side_rad.data(), ctx->side_rad.p, n_side * sizeof(float4), hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(side_info.data(), ctx->side_info.p, n_side * sizeof(int4), hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(side_view.data(), ctx->side_view.p, n_side * sizeof(int), hipMemcpyDeviceToHost));
    }
    float gpu_ms = 0.f;
    HIPX(hipEventElapsedTime(&gpu_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    // Side entries by the tile of their source pixel, in visitation order (row-major sources, then
    // sample index), so each tile's one-pixel margin receives them as render_tile adds them.
    std::vector<std::vector<size_t>> side_of_tile(tile_first.size());
    for (size_t i = 0; i < n_side; i++) {
        const DevView& v = dv[side_view[i]];
        const int tx = (side_info[i].x - v.sx0) >> 6, ty = (side_info[i].y - v.sy0) >> 6;
        side_of_tile[(size_t)v.tile_base + (size_t)ty * v.tgx + tx].push_back(i);
    }
    parallel_for(side_of_tile.size(), [&](size_t t) {
        std::vector<size_t>& l = side_of_tile[t];
        std::sort(l.begin(), l.end(), [&](size_t a, size_t b) {
            const int4 &A = side_info[a], &B = side_info[b];
            if (A.y != B.y) return A.y < B.y;
            if (A.x != B.x) return A.x < B.x;
            return A.z < B.z;
        });
    }, 64);

    // Assemble Film_Tiles: in-tile values from the device gather, margins from the side list.
    out.assign(parts.size(), {});
    for (size_t pi = 0; pi < parts.size(); pi++) {
        out[pi].tiles.assign(parts[pi].tiles.size(), {});
        out[pi].tile_var.assign(parts[pi].tiles.size(), 0.0);
    }
    {
        // every unit (tile) fills its own Film_Tile: independent, so the units run on all host cores
        std::vector<size_t> unit_ti(units.size());
        {
            std::vector<size_t> next(parts.size(), 0);
            for (size_t ui = 0; ui < units.size(); ui++) unit_ti[ui] = next[units[ui].part]++;
        }
        parallel_for(units.size(), [&](size_t ui) {
            const Unit& u = units[ui];
            const yar_view& v = ctx->views[parts[u.part].view];
            const TileGeom g = tile_geom(v, u.tile);
            const size_t ti = unit_ti[ui];
            const int pw = g.px1 - g.px0, ph = g.py1 - g.py0;
            std::vector<yar_film_pixel>& tp = out[u.part].tiles[ti];
            tp.assign((size_t)std::max(0, pw) * std::max(0, ph), yar_film_pixel{{0, 0, 0}, 0});
            double tv = 0.0;
            int64_t off = u.start;
            for (int y = g.y0; y < g.y1; y++)
                for (int x = g.x0; x < g.x1; x++, off++) {
                    const float4 a = film_own[off];
                    yar_film_pixel& fp = tp[(size_t)(y - g.py0) * pw + (x - g.px0)];
                    fp.color_sum[0] = a.x; fp.color_sum[1] = a.y; fp.color_sum[2] = a.z; fp.weight_sum = a.w;
                    tv += var_own[off];
                }
            out[u.part].tile_var[ti] = tv;
            for (size_t si : side_of_tile[(size_t)dv[u.part].tile_base + u.tile]) {
                const int4 inf = side_info[si];
                const float4 r = side_rad[si];
                const int fx0 = inf.x + (inf.w & 3) - 1, fx1 = inf.x + ((inf.w >> 2) & 3);
                const int fy0 = inf.y + ((inf.w >> 4) & 3) - 1, fy1 = inf.y + ((inf.w >> 6) & 3);
                for (int ty = fy0; ty < fy1; ty++)
                    for (int tx = fx0; tx < fx1; tx++) {
                        if (tx < g.px0 || tx >= g.px1 || ty < g.py0 || ty >= g.py1) continue;
                        if (tx >= g.x0 && tx < g.x1 && ty >= g.y0 && ty < g.y1) continue;   // in-tile: device gather
                        yar_film_pixel& fp = tp[(size_t)(ty - g.py0) * pw + (tx - g.px0)];
                        fp.color_sum[0] += 1.f * r.x; fp.color_sum[1] += 1.f * r.y; fp.color_sum[2] += 1.f * r.z;
                        fp.weight_sum += 1.f;
                    }
            }
        });
    }
    int64_t samples = 0;
    for (const Unit& u : units) samples += u.count * dv[u.part].spp;
    st = yar_render_stats{};
    st.gpu_seconds = gpu_ms * 1e-3;
    st.samples = samples;
    st.closest_rays = (int64_t)(tot[16] + tot[18]);
    st.any_rays = (int64_t)tot[17];
    st.node_visits = (int64_t)(tot[0] + tot[4] + tot[8]);
    st.leaf_refs = (int64_t)(tot[1] + tot[5] + tot[9]);
    st.triangle_tests = (int64_t)(tot[2] + tot[6] + tot[10]);
    st.object_visits = (int64_t)(tot[3] + tot[7] + tot[11]);
    st.extend_node_visits = (int64_t)tot[0];
    st.extend_leaf_refs = (int64_t)tot[1];
    st.extend_triangle_tests = (int64_t)tot[2];
    st.extend_object_visits = (int64_t)tot[3];
    st.extend_rays = (int64_t)tot[16];
    st.extend_launches = (int64_t)tot[19];
    st.mis_culled_rays = (int64_t)tot[20];
    st.extend_seconds = ksec[0];
    st.shade_seconds = ksec[1];
    st.shadow_seconds = ksec[2];
    st.mis_seconds = ksec[3];
    st.other_seconds = ksec[4];
    st.guide_seconds = ksec[5];
    st.train_seconds = train_seconds;
    st.train_nats = train_nats;
    st.train_records = train_records;
    st.passes = guided ? (int64_t)gr->passes.size() : 1;
    st.render_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

// Film::merge_tile of every tile in index order, then Film::get_image (film.cpp:88-118), and the
// variance estimate of render_scene (reference_renderer.cpp:566-576).
// `film` is the caller's reusable work buffer: each band clears its own rows before merging.
static void merge_view(const yar_view& v, const PartResult& r, yar_film_pixel* out_film, float* out_image,
                       double* out_variance, std::vector<yar_film_pixel>& film) {
    const int n = (int)r.tiles.size();
    const int rw = v.region_x1 - v.region_x0, rh = v.region_y1 - v.region_y0;
    if (film.size() < (size_t)rw * rh) film.resize((size_t)rw * rh);
    std::vector<TileGeom> geo((size_t)n);
    for (int t = 0; t < n; t++) geo[(size_t)t] = tile_geom(v, t);
    // bands of 16 rows in parallel; within a band every pixel still takes the tiles in index order
    const int bands = (rh + 15) / 16;
    parallel_for((size_t)bands, [&](size_t b) {
        const int ry0 = v.region_y0 + (int)b * 16, ry1 = std::min(v.region_y1, ry0 + 16);
        std::fill(film.begin() + (size_t)(ry0 - v.region_y0) * rw, film.begin() + (size_t)(ry1 - v.region_y0) * rw,
                  yar_film_pixel{{0, 0, 0}, 0});
        for (int t = 0; t < n; t++) {
            const TileGeom& g = geo[(size_t)t];
            const int pw = g.px1 - g.px0;
            for (int y = std::max(g.py0, ry0); y < std::min(g.py1, ry1); y++)
                for (int x = g.px0; x < g.px1; x++) {
                    yar_film_pixel& fo = film[(size_t)(y - v.region_y0) * rw + (x - v.region_x0)];
                    const yar_film_pixel& p = r.tiles[t][(size_t)(y - g.py0) * pw + (x - g.px0)];
                    fo.color_sum[0] += p.color_sum[0]; fo.color_sum[1] += p.color_sum[1]; fo.color_sum[2] += p.color_sum[2];
                    fo.weight_sum += p.weight_sum;
                }
        }
        if (out_film)
            std::memcpy(out_film + (size_t)(ry0 - v.region_y0) * rw, film.data() + (size_t)(ry0 - v.region_y0) * rw,
                        (size_t)(ry1 - ry0) * rw * sizeof(yar_film_pixel));
        if (out_image)
            for (size_t i = (size_t)(ry0 - v.region_y0) * rw; i < (size_t)(ry1 - v.region_y0) * rw; i++) {
                const yar_film_pixel& fp = film[i];
                float cr = 0, cg = 0, cb = 0;
                if (fp.weight_sum != 0.f) {
                    const float inv = 1.f / fp.weight_sum;
                    cr = fp.color_sum[0] * inv; cg = fp.color_sum[1] * inv; cb = fp.color_sum[2] * inv;
                }
                out_image[3 * i] = std::max(0.f, cr); out_image[3 * i + 1] = std::max(0.f, cg); out_image[3 * i + 2] = std::max(0.f, cb);
            }
    });
    if (out_variance) {
        const int spp = v.x_pixel_sample_count * v.y_pixel_sample_count;
        double acc = 0.0;
        int64_t cnt = 0;
        for (int t = 0; t < n; t++) {
            acc += r.tile_var[t];
            const TileGeom g = tile_geom(v, t);
            cnt += (int64_t)(g.x1 - g.x0) * (g.y1 - g.y0);
        }
        *out_variance = spp > 1 ? acc / cnt : 0.0;
    }
}

static bool valid_view(const yar_view& v, std::string* why) {
    if (v.film_width <= 0 || v.film_height <= 0) { *why = "film size must be positive"; return false; }
    if (v.region_x0 < 0 || v.region_y0 < 0 || v.region_x0 >= v.region_x1 || v.region_y0 >= v.region_y1 ||
        v.region_x1 > v.film_width || v.region_y1 > v.film_height) { *why = "render region outside the film"; return false; }
    if (v.region_x1 > 0x10000 || v.region_y1 > 0x10000) { *why = "film larger than 65536 pixels per side"; return false; }
    if (v.x_pixel_sample_count <= 0 || v.y_pixel_sample_count <= 0) { *why = "sample counts must be positive"; return false; }
    return true;
}

}  // namespace

extern "C" {

int yar_gpu_create(const yar_scene_desc* scene, const yar_gpu_options* options, yar_gpu_ctx** out_ctx) {
    if (!scene || !out_ctx) return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_create: null argument");
    *out_ctx = nullptr;
    if (scene->object_count <= 0 || scene->mesh_count <= 0)
        return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_create: scene has no geometry");
    if (scene->rect_light_count < 0 || scene->point_light_count < 0 || scene->spot_light_count < 0 ||
        scene->directional_light_count < 0 || scene->mesh_light_count < 0 || scene->sphere_light_count < 0 ||
        (scene->spot_light_count > 0 && !scene->spot_lights) ||
        (scene->sphere_light_count > 0 && !scene->sphere_lights) ||
        scene->environment_light_count < 0 || scene->environment_light_count > 1 ||
        (scene->environment_light_count > 0 && !scene->environment_light) ||
        (scene->directional_light_count > 0 && !scene->directional_lights) ||
        (scene->mesh_light_count > 0 && !scene->mesh_lights))
        return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_create: invalid light arrays");
    if (scene->rect_light_count + scene->point_light_count + scene->spot_light_count + scene->directional_light_count +
            scene->sphere_light_count + scene->mesh_light_count + scene->environment_light_count <= 0)
        return fail(nullptr, YAR_ERR_UNSUPPORTED, "yar_gpu_create: scene has no lights");
    const int n_area = scene->rect_light_count + scene->sphere_light_count + scene->mesh_light_count;
    for (int i = 0; i < scene->sphere_light_count; i++) {
        const yar_sphere_light_desc& l = scene->sphere_lights[i];
        if (l.object < 0 || l.object >= scene->object_count ||
            scene->objects[l.object].area_light != scene->rect_light_count + i)
            return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "sphere light object does not reference the light");
        if (!(l.radius > 0.f) || !std::isfinite(l.radius))
            return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "sphere light radius must be positive");
    }
    for (int i = 0; i < scene->mesh_light_count; i++) {
        const int ob = scene->mesh_lights[i].object;
        if (ob < 0 || ob >= scene->object_count ||
            scene->objects[ob].area_light != scene->rect_light_count + scene->sphere_light_count + i)
            return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "mesh light object does not reference the light");
        static const float identity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        if (std::memcmp(scene->objects[ob].object_to_world, identity, sizeof(identity)) != 0)
            return fail(nullptr, YAR_ERR_UNSUPPORTED, "mesh light objects must have identity transforms");
    }
    for (int i = 0; i < scene->object_count; i++) {
        const yar_object_desc& o = scene->objects[i];
        if (o.mesh < 0 || o.mesh >= scene->mesh_count) return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "object mesh index out of range");
        if (o.area_light >= n_area)
            return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "object light index out of range");
        if (o.area_light < 0 && (o.material < 0 || o.material >= scene->material_count))
            return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "object material index out of range");
    }
    for (int i = 0; i < scene->mesh_count; i++) {
        const yar_mesh_desc& m = scene->meshes[i];
        if (m.triangle_count <= 0 || !m.vertices || !m.indices) return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "empty mesh");
        for (int64_t k = 0; k < 3 * (int64_t)m.triangle_count; k++)
            if (m.indices[k] < 0 || m.indices[k] >= m.vertex_count) return fail(nullptr, YAR_ERR_INVALID_ARGUMENT, "mesh index out of range");
    }
    std::unique_ptr<yar_gpu_ctx> ctx;
    const int rc = guarded(nullptr, "yar_gpu_create", [&] {
        ctx.reset(new yar_gpu_ctx);
        if (options) ctx->opt = *options;
        if (ctx->opt.block_size != 0 && (ctx->opt.block_size < 64 || ctx->opt.block_size > 256 || ctx->opt.block_size % 64 != 0))
            throw HipError{"block_size must be 0 (64), 64, 128, 192 or 256", YAR_ERR_INVALID_ARGUMENT};
        ctx->device = ctx->opt.device;
        ctx->cfg = scene->config;
        int ndev = 0;
        HIPX(hipGetDeviceCount(&ndev));
        if (ctx->device < 0 || ctx->device >= ndev) throw HipError{"no such HIP device", YAR_ERR_INVALID_ARGUMENT};
        HIPX(hipSetDevice(ctx->device));
        hipDeviceProp_t prop;
        HIPX(hipGetDeviceProperties(&prop, ctx->device));
        std::snprintf(ctx->info.device_name, sizeof(ctx->info.device_name), "%s", prop.gcnArchName);
        ctx->cu_count = std::max(1, prop.multiProcessorCount);
        HIPX(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->scratch = scratch_ptr(ctx->device);
        upload_scene(ctx.get(), scene);
        upload_jump_tables(ctx.get());
        yar_view v{};
        std::memcpy(v.camera_to_world, scene->camera.camera_to_world, sizeof(v.camera_to_world));
        v.fov_y_degrees = scene->camera.fov_y_degrees;
        v.z_is_up = scene->camera.z_is_up;
        v.film_width = scene->config.film_width;
        v.film_height = scene->config.film_height;
        v.region_x0 = scene->config.region_x0; v.region_y0 = scene->config.region_y0;
        v.region_x1 = scene->config.region_x1; v.region_y1 = scene->config.region_y1;
        v.x_pixel_sample_count = scene->config.x_pixel_sample_count;
        v.y_pixel_sample_count = scene->config.y_pixel_sample_count;
        v.rng_seed_offset = scene->config.rng_seed_offset;
        ctx->views.assign(1, v);
    });
    if (rc == YAR_OK) *out_ctx = ctx.release();
    return rc;
}

void yar_gpu_destroy(yar_gpu_ctx* ctx) { delete ctx; }

const char* yar_gpu_last_error(const yar_gpu_ctx* ctx) { return ctx ? ctx->error.c_str() : yar::get_last_error(); }

int yar_gpu_set_views(yar_gpu_ctx* ctx, const yar_view* views, int32_t count) {
    if (!ctx || !views || count <= 0) return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_set_views: invalid argument");
    std::string why;
    for (int i = 0; i < count; i++)
        if (!valid_view(views[i], &why)) return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_set_views: view " + std::to_string(i) + ": " + why);
    return guarded(ctx, "yar_gpu_set_views", [&] { ctx->views.assign(views, views + count); });
}

int yar_gpu_get_view(const yar_gpu_ctx* ctx, int32_t index, yar_view* out_view) {
    if (!ctx || !out_view || index < 0 || index >= (int)ctx->views.size()) return YAR_ERR_INVALID_ARGUMENT;
    *out_view = ctx->views[index];
    return YAR_OK;
}

int yar_film_tile_count(const yar_gpu_ctx* ctx, int32_t view, int32_t* out_count) {
    if (!ctx || !out_count || view < 0 || view >= (int)ctx->views.size()) return YAR_ERR_INVALID_ARGUMENT;
    int tgx, tgy, a, b, c2, d;
    film_layout(ctx->views[view], &tgx, &tgy, &a, &b, &c2, &d);
    *out_count = tgx * tgy;
    return YAR_OK;
}

int yar_film_tile_bounds(const yar_gpu_ctx* ctx, int32_t view, int32_t tile_index, int32_t out_sample_bounds[4],
                         int32_t out_pixel_bounds[4]) {
    int32_t n = 0;
    if (yar_film_tile_count(ctx, view, &n) != YAR_OK || tile_index < 0 || tile_index >= n) return YAR_ERR_INVALID_ARGUMENT;
    TileGeom g = tile_geom(ctx->views[view], tile_index);
    if (out_sample_bounds) { out_sample_bounds[0] = g.x0; out_sample_bounds[1] = g.y0; out_sample_bounds[2] = g.x1; out_sample_bounds[3] = g.y1; }
    if (out_pixel_bounds) { out_pixel_bounds[0] = g.px0; out_pixel_bounds[1] = g.py0; out_pixel_bounds[2] = g.px1; out_pixel_bounds[3] = g.py1; }
    return YAR_OK;
}

int yar_gpu_render_tiles(yar_gpu_ctx* ctx, int32_t view, const int32_t* tile_indices, int32_t count,
                         yar_film_pixel* const* out_tiles, double* out_tile_variance, yar_render_stats* out_stats) {
    int32_t n = 0;
    if (!ctx || !tile_indices || count <= 0 || !out_tiles || yar_film_tile_count(ctx, view, &n) != YAR_OK)
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_render_tiles: invalid argument");
    std::vector<JobPart> parts;
    {
        const int rc = guarded(ctx, "yar_gpu_render_tiles", [&] {
            parts.resize(1);
            parts[0].view = view;
            parts[0].tiles.assign(tile_indices, tile_indices + count);
            std::vector<char> seen(n, 0);
            for (int t : parts[0].tiles) {
                if (t < 0 || t >= n) throw HipError{"tile index out of range", YAR_ERR_INVALID_ARGUMENT};
                if (seen[t]) throw HipError{"tile listed twice", YAR_ERR_INVALID_ARGUMENT};
                seen[t] = 1;
            }
        });
        if (rc != YAR_OK) return rc;
    }
    return guarded(ctx, "yar_gpu_render_tiles", [&] {
        std::vector<PartResult> rr;
        yar_render_stats st{};
        HIPX(hipSetDevice(ctx->device));
        render_job(ctx, parts, rr, st);
        for (int i = 0; i < count; i++) {
            if (out_tiles[i]) std::memcpy(out_tiles[i], rr[0].tiles[i].data(), rr[0].tiles[i].size() * sizeof(yar_film_pixel));
            if (out_tile_variance) out_tile_variance[i] = rr[0].tile_var[i];
        }
        if (out_stats) *out_stats = st;
    });
}

int yar_gpu_render(yar_gpu_ctx* ctx, int32_t view, yar_film_pixel* out_film, float* out_image, double* out_variance,
                   yar_render_stats* out_stats) {
    int32_t n = 0;
    if (!ctx || yar_film_tile_count(ctx, view, &n) != YAR_OK) return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_render: invalid view");
    return guarded(ctx, "yar_gpu_render", [&] {
        std::vector<JobPart> parts(1);
        parts[0].view = view;
        for (int i = 0; i < n; i++) parts[0].tiles.push_back(i);
        std::vector<PartResult> rr;
        yar_render_stats st{};
        HIPX(hipSetDevice(ctx->device));
        render_job(ctx, parts, rr, st);
        merge_view(ctx->views[view], rr[0], out_film, out_image, out_variance, ctx->merge_work);
        if (out_stats) *out_stats = st;
    });
}

int yar_gpu_render_checkpointed(yar_gpu_ctx* ctx, int32_t view, const char* checkpoint_dir, const char* input_filename,
                                int32_t tiles_per_batch, yar_film_pixel* out_film, float* out_image, double* out_variance,
                                float* out_render_time, yar_render_stats* out_stats) {
    int32_t n = 0;
    if (!ctx || !checkpoint_dir || !input_filename || tiles_per_batch < 0 || yar_film_tile_count(ctx, view, &n) != YAR_OK)
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_render_checkpointed: invalid argument");
    return guarded(ctx, "yar_gpu_render_checkpointed", [&] {
        const auto t0 = std::chrono::steady_clock::now();
        const yar_view v = ctx->views[view];
        yar_checkpoint* ck = nullptr;
        const int rc = yar_checkpoint_open(checkpoint_dir, input_filename, n, v.x_pixel_sample_count * v.y_pixel_sample_count, &ck);
        if (rc != YAR_OK) throw HipError{yar::get_last_error(), rc};
        std::unique_ptr<yar_checkpoint, void (*)(yar_checkpoint*)> hold(ck, yar_checkpoint_free);
        const float previous = yar_checkpoint_previous_time(ck);
        PartResult all;
        all.tiles.resize(n);
        all.tile_var.assign(n, 0.0);
        std::vector<int> todo;
        for (int t = 0; t < n; t++) {
            const TileGeom g = tile_geom(v, t);
            int32_t b[4];
            if (yar_checkpoint_tile(ck, t, b, nullptr, nullptr) != YAR_OK) { todo.push_back(t); continue; }
            if (b[0] != g.px0 || b[1] != g.py0 || b[2] != g.px1 || b[3] != g.py1)
                throw HipError{"checkpoint tile " + std::to_string(t) + " has other pixel bounds than this film", YAR_ERR_INVALID_ARGUMENT};
            all.tiles[t].resize((size_t)(g.px1 - g.px0) * (g.py1 - g.py0));
                if (yar_checkpoint_tile(ck, t, b, all.tiles[t].data(), &all.tile_var[t]) != YAR_OK)
                throw HipError{yar::get_last_error(), YAR_ERR_INTERNAL};
        }
        HIPX(hipSetDevice(ctx->device));
        const int batch = tiles_per_batch > 0 ? tiles_per_batch : 64;
        yar_render_stats total{};
        for (size_t i0 = 0; i0 < todo.size(); i0 += (size_t)batch) {
            std::vector<JobPart> parts(1);
            parts[0].view = view;
            parts[0].tiles.assign(todo.begin() + i0, todo.begin() + std::min(todo.size(), i0 + (size_t)batch));
            std::vector<PartResult> rr;
            yar_render_stats st{};
            render_job(ctx, parts, rr, st);
            const float now = previous + (float)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            for (size_t k = 0; k < parts[0].tiles.size(); k++) {
                const int t = parts[0].tiles[k];
                const TileGeom g = tile_geom(v, t);
                const int32_t b[4] = {g.px0, g.py0, g.px1, g.py1};
                all.tiles[t] = std::move(rr[0].tiles[k]);
                all.tile_var[t] = rr[0].tile_var[k];
                const int wrc = yar_checkpoint_write_tile(checkpoint_dir, t, b, all.tiles[t].data(), all.tile_var[t], now);
                if (wrc != YAR_OK) throw HipError{yar::get_last_error(), wrc};
            }
            total.samples += st.samples;
            total.closest_rays += st.closest_rays;
            total.any_rays += st.any_rays;
            total.gpu_seconds += st.gpu_seconds;
        }
        merge_view(v, all, out_film, out_image, out_variance, ctx->merge_work);
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total.render_seconds = elapsed;
        total.passes = 1;
        if (out_render_time) *out_render_time = previous + (float)elapsed;
        if (out_stats) *out_stats = total;
    });
}

int yar_gpu_render_batch(yar_gpu_ctx* ctx, const int32_t* views, int32_t count, float* const* out_images,
                         double* out_variances, yar_render_stats* out_stats) {
    if (!ctx || !views || count <= 0) return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_render_batch: invalid argument");
    return guarded(ctx, "yar_gpu_render_batch", [&] {
        std::vector<JobPart> parts(count);
        for (int i = 0; i < count; i++) {
            int32_t n = 0;
            if (yar_film_tile_count(ctx, views[i], &n) != YAR_OK) throw HipError{"invalid view index", YAR_ERR_INVALID_ARGUMENT};
            parts[i].view = views[i];
            for (int t = 0; t < n; t++) parts[i].tiles.push_back(t);
        }
        std::vector<PartResult> rr;
        yar_render_stats st{};
        HIPX(hipSetDevice(ctx->device));
        render_job(ctx, parts, rr, st);
        for (int i = 0; i < count; i++)
            merge_view(ctx->views[views[i]], rr[i], nullptr, out_images ? out_images[i] : nullptr,
                       out_variances ? &out_variances[i] : nullptr, ctx->merge_work);
        if (out_stats) *out_stats = st;
    });
}

int yar_gpu_render_guided(yar_gpu_ctx* ctx, yar_guide* guide, const int32_t* views, int32_t count,
                          float* const* out_images, double* out_variances, yar_render_stats* out_stats) {
    if (!ctx || !guide || !views || count <= 0)
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_render_guided: invalid argument");
    if (guide_device(guide) != ctx->device)
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_render_guided: guide lives on another device");
    return guarded(ctx, "yar_gpu_render_guided", [&] {
        std::vector<JobPart> parts(count);
        for (int i = 0; i < count; i++) {
            int32_t n = 0;
            if (yar_film_tile_count(ctx, views[i], &n) != YAR_OK) throw HipError{"invalid view index", YAR_ERR_INVALID_ARGUMENT};
            parts[i].view = views[i];
            for (int t = 0; t < n; t++) parts[i].tiles.push_back(t);
        }
        std::vector<PartResult> rr;
        yar_render_stats st{};
        GuidedRun gr;
        gr.guide = guide;
        HIPX(hipSetDevice(ctx->device));
        render_job(ctx, parts, rr, st, &gr);
        for (int i = 0; i < count; i++)
            merge_view(ctx->views[views[i]], rr[i], nullptr, out_images ? out_images[i] : nullptr,
                       out_variances ? &out_variances[i] : nullptr, ctx->merge_work);
        if (out_stats) *out_stats = st;
    });
}

int yar_gpu_trace_rays_device(yar_gpu_ctx* ctx, const yar_ray* d_rays, int64_t count, int32_t any_hit, yar_hit* d_hits,
                              void* hip_stream);

int yar_gpu_trace_rays(yar_gpu_ctx* ctx, const yar_ray* rays, int64_t count, int32_t any_hit, yar_hit* out_hits) {
    if (!ctx || (!rays && count > 0) || (!out_hits && count > 0) || count < 0)
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_trace_rays: invalid argument");
    if (count == 0) return YAR_OK;
    return guarded(ctx, "yar_gpu_trace_rays", [&] {
        HIPX(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        ctx->tr_rays.reserve(2 * (size_t)count);
        ctx->tr_tb.reserve((size_t)count);
        ctx->tr_ot.reserve((size_t)count);
        ctx->totals.reserve(16);
        HIPX(hipMemcpyAsync(ctx->tr_rays.p, rays, (size_t)count * sizeof(yar_ray), hipMemcpyHostToDevice, s));
        HIPX(hipMemsetAsync(ctx->totals.p, 0, 16 * sizeof(unsigned long long), s));
        LaunchCfg lc = ctx->launch_cfg();
        launch_trace_rays(ctx->scene, ctx->tr_rays.p, count, any_hit, ctx->tr_tb.p, ctx->tr_ot.p, ctx->totals.p, lc);
        HIPX(hipGetLastError());
        std::vector<float4> tb(count);
        std::vector<int2> ot(count);
        unsigned long long tot[16];
        HIPX(hipMemcpyAsync(tb.data(), ctx->tr_tb.p, (size_t)count * sizeof(float4), hipMemcpyDeviceToHost, s));
        HIPX(hipMemcpyAsync(ot.data(), ctx->tr_ot.p, (size_t)count * sizeof(int2), hipMemcpyDeviceToHost, s));
        HIPX(hipMemcpyAsync(tot, ctx->totals.p, sizeof(tot), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        for (int64_t i = 0; i < count; i++) {
            yar_hit& h = out_hits[i];
            h.t = tb[i].x;
            h.object = ot[i].x;
            h.triangle = ot[i].y;
            h.barycentrics[0] = tb[i].y; h.barycentrics[1] = tb[i].z; h.barycentrics[2] = tb[i].w;
        }
        ctx->last_trace = yar_render_stats{};
        ctx->last_trace.node_visits = (int64_t)tot[12];
        ctx->last_trace.leaf_refs = (int64_t)tot[13];
        ctx->last_trace.triangle_tests = (int64_t)tot[14];
        ctx->last_trace.object_visits = (int64_t)tot[15];
        if (any_hit) ctx->last_trace.any_rays = count; else ctx->last_trace.closest_rays = count;
    });
}

}  // extern "C"

namespace {
__global__ void k_pack_hits(const float4* __restrict__ tb, const int2* __restrict__ ot, long long n, yar_hit* out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = tb[i];
    int2 b = ot[i];
    yar_hit h;
    h.t = a.x; h.object = b.x; h.triangle = b.y;
    h.barycentrics[0] = a.y; h.barycentrics[1] = a.z; h.barycentrics[2] = a.w;
    out[i] = h;
}
}  // namespace

extern "C" int yar_gpu_trace_rays_device(yar_gpu_ctx* ctx, const yar_ray* d_rays, int64_t count, int32_t any_hit,
                                         yar_hit* d_hits, void* hip_stream) {
    if (!ctx || count < 0 || (count > 0 && (!d_rays || !d_hits)))
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_trace_rays_device: invalid argument");
    if (count == 0) return YAR_OK;
    return guarded(ctx, "yar_gpu_trace_rays_device", [&] {
        HIPX(hipSetDevice(ctx->device));
        ctx->tr_tb.reserve((size_t)count);
        ctx->tr_ot.reserve((size_t)count);
        ctx->totals.reserve(16);
        LaunchCfg lc = ctx->launch_cfg();
        if (hip_stream) lc.stream = (hipStream_t)hip_stream;
        HIPX(hipMemsetAsync(ctx->totals.p, 0, 16 * sizeof(unsigned long long), lc.stream));
        launch_trace_rays(ctx->scene, reinterpret_cast<const float4*>(d_rays), count, any_hit, ctx->tr_tb.p, ctx->tr_ot.p,
                          ctx->totals.p, lc);
        hipLaunchKernelGGL(k_pack_hits, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, lc.stream, ctx->tr_tb.p,
                           ctx->tr_ot.p, (long long)count, d_hits);
        HIPX(hipGetLastError());
    });
}

extern "C" int yar_gpu_last_trace_counters(yar_gpu_ctx* ctx, yar_render_stats* out_stats) {
    if (!ctx || !out_stats) return YAR_ERR_INVALID_ARGUMENT;
    *out_stats = ctx->last_trace;
    return YAR_OK;
}

extern "C" int yar_gpu_get_info(const yar_gpu_ctx* ctx, yar_gpu_info* out_info) {
    if (!ctx || !out_info) return YAR_ERR_INVALID_ARGUMENT;
    *out_info = ctx->info;
    return YAR_OK;
}

// Allocation-free form of yar_gpu_trace_rays_device: the same kernels, with the hit staging and the
// counters in this device's `scratch` array instead of context buffers, cleared by a kernel (no
// memset node), so the call can sit inside a HIP graph capture or a loop that must not allocate.
extern "C" int yar_gpu_trace_rays_device_noalloc(yar_gpu_ctx* ctx, const yar_ray* d_rays, int64_t count, int32_t any_hit,
                                                 yar_hit* d_hits, void* hip_stream) {
    if (!ctx || count < 0 || (count > 0 && (!d_rays || !d_hits)))
        return fail(ctx, YAR_ERR_INVALID_ARGUMENT, "yar_gpu_trace_rays_device_noalloc: invalid argument");
    if (count > kNoAllocMaxRays)
        return fail(ctx, YAR_ERR_CAPACITY, "yar_gpu_trace_rays_device_noalloc: more than 2^20 rays in one call (scratch limit)");
    if (count == 0) return YAR_OK;
    return guarded(ctx, "yar_gpu_trace_rays_device_noalloc", [&] {
        HIPX(hipSetDevice(ctx->device));
        if (!ctx->scratch) throw HipError{"scratch array not set up for this device"};
        float4* tb = reinterpret_cast<float4*>(ctx->scratch);
        int2* ot = reinterpret_cast<int2*>(ctx->scratch + kNoAllocTbBytes);
        unsigned long long* totals = reinterpret_cast<unsigned long long*>(ctx->scratch + kNoAllocTbBytes + kNoAllocOtBytes);
        LaunchCfg lc = ctx->launch_cfg();
        if (hip_stream) lc.stream = (hipStream_t)hip_stream;
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, lc.stream, totals, 16);
        launch_trace_rays(ctx->scene, reinterpret_cast<const float4*>(d_rays), count, any_hit, tb, ot, totals, lc);
        hipLaunchKernelGGL(k_pack_hits, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, lc.stream, tb, ot,
                           (long long)count, d_hits);
        HIPX(hipGetLastError());
    });
}

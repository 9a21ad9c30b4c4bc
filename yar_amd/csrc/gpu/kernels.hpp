// kernels.hpp - host-visible declarations of the wavefront path-tracer kernels (render.hip) used by
// the context code (context.hip).
#pragma once

#include "device.hpp"
#include "guide.hpp"

namespace yg {

// One camera view of a render (the Scene_Overrides analog, reference_renderer.h:23): camera,
// sampler grid, seed offset and film layout (Film::Film / get_tile_bounds, film.cpp:52-86).
struct DevView {
    float4 c0, c1, c2;              // camera_to_world rows
    float W, H, horz, vert;         // film extent and tan(fov/2) terms (camera.cpp:5-14)
    int zup, spp, nx, ny;
    uint32_t seed_offset;
    int tile_base;                  // offset of this view's tiles in the tile_first array
    int sx0, sy0;                   // sample region (== render region for the radius-0.5 box filter)
    int sx1, sy1, tgx, pad;
    int rx0, ry0, rx1, ry1;         // render region
};

// Per-path state, structure of arrays indexed by path slot.
struct Paths {
    float4* ray_o;       // origin xyz
    float4* ray_d;       // direction xyz
    float4* coeff;       // path throughput (path_tracing.cpp:18)
    float4* L;           // accumulated radiance
    uint32_t* state;     // bounce | c1 << 8 | c2 << 16 | flags
    float4* dl_coeff;    // throughput at the direct-lighting vertex
    float4* dl_light;    // light-sampling half of the MIS estimate
    float4* dl_bsdf;     // BSDF-sampling half
    float4* hit_tb;      // closest hit: t, barycentrics
    int2* hit_ot;        // closest hit: object, global triangle
    uint2* pk;           // chunk pixel and sample index of the path
    uint64_t* rng;       // per-pixel fallback stream (sequential sample mode only)
    float4* radiance;    // final per-sample radiance of the chunk, at pix_first[pixel] + sample
    // guided form only
    uint64_t* grng;      // per-path stream: strategy choice, mixture sampling, fallback dimensions
    float* rec;          // kRecMax training vertices per path, kRecFloats each
};

// Training vertex of a guided path: network inputs, sampled direction, throughput before the
// continuation and the path radiance after the vertex's direct lighting.
constexpr int kRecMax = 4;
constexpr int kRecFloats = 18;     // x[9], w[3], beta[3], L[3]

enum : uint32_t {
    kDelta = 1u << 24,        // current ray was produced by a delta bounce (path_tracing.cpp:104)
    kPendingDL = 1u << 25,    // direct-lighting halves must be added by the resolve pass
    kDone = 1u << 26,         // path finished; finalize after the resolve pass
    kRecShift = 27,           // bits 27..29: training vertices recorded so far (guided form)
    kRecPending = 1u << 30,   // the newest vertex still needs its post-direct-lighting radiance
};

// Work queues.  The active path list ping-pongs between two buffers; `parity` selects the input.
// counts: [0], [1] active lists, [2] shadow, [3] mis, [4] dl, [5] film side entries,
//         [6] [7] [8] work-fetch cursors of extend, shadow and mis, [9] guide queries,
//         [10] training records.
struct Queues {
    uint32_t* active0;
    uint32_t* active1;
    uint32_t* shadow_path; float4* shadow_o; float4* shadow_d; float4* shadow_c;
    uint32_t* mis_path; float4* mis_o; float4* mis_d; float4* mis_f; float* mis_ndw;
    uint32_t* dl_path;
    uint32_t* counts;
    // guided form: continuation queries ([9] counts them) and training records ([10])
    uint32_t* guide_path;
    float* guide_in;              // 16 floats per query: inputs x[9], rec index, BSDF u[2], RR u
    float* guide_out;             // kGOut network outputs per query
    GuideRecord* train;
    uint32_t train_cap;
    float keep_prob;
    // [0..15] traversal counters (4 per kernel kind), [16] extend rays, [17] shadow rays, [18] mis rays,
    // [19] extend launches, [20] MIS rays answered by the light-rectangle test (count_culls only)
    unsigned long long* totals;
    int count_culls;
    // Direction binning of the next iteration's active list (bin_rays = 1, 2, 3: cube-map cells of
    // 2x2, 4x4, 8x8 per face): shade and guide sample append to `unsorted`; k_bin_local sorts each
    // window of 4,096 consecutive entries by direction cell in LDS and writes it to the active list
    // in place, so a traversal wave takes rays of similar direction from nearby samples.
    uint32_t* unsorted;
    int bin_rays;
};

// One chunk of pixels (whole film tiles, possibly of several views).  Lives in device memory so
// the bounce loop can be captured once in a hipGraph and replayed for every chunk and pass.
struct Chunk {
    const int2* pixels;             // film coordinates of the chunk's pixels
    const int* pix_view;            // view of each pixel
    const long long* pix_first;     // prefix sums of spp over the chunk's pixels (n_pixels + 1)
    const DevView* views;
    // Shuffle permutations of the sampler's 1D and 2D dimensions: pixel p's runs start at entry
    // pix_first[p] * (d1 + d2), one run of spp entries per dimension (1D first), entries of
    // perm_bytes (1 or 2) bytes.  The jittered values are recomputed from the pixel stream (jt).
    const uint8_t* perm;
    JumpTables jt;
    long long n_paths;              // paths of the current pass
    int n_pixels, max_spp;
    int d1, d2, perm_bytes;         // sample-vector dims (reference_renderer.cpp:87-99)
    int pass, sequential;           // sequential: one sample per pixel per pass (RNG fallback)
    int max_bounces, rr_bounce_threshold;
    float rr_threshold, film_scale, max_component;
    // guided form: this pass traces samples [k0, k1) of every pixel, slot = pixel * (k1 - k0) + j
    int guided, k0, k1;
    float alpha;                    // one-sample MIS probability of the learned mixture
    int uniform_spp;                // > 0 when every pixel of the chunk has this many samples
};

// Film accumulation of one chunk (every sample's footprint as Film_Tile::add_sample computes it).
struct FilmParams {
    float4* out_film;         // per listed pixel: its Film_Tile value inside its own tile
    double* out_var;          // per listed pixel: luminance variance of its estimate
    uint8_t* spill;           // per chunk pixel: 8-neighbour footprint bits of its samples
    const int* tile_first;    // per (view, tile): pixel-list index of its first pixel, -1 if absent
    long long base;           // pixel-list index of the chunk's first pixel
    float4* side_rad;         // samples reaching a tile's one-pixel margin: radiance
    int4* side_info;          // x, y, sample, packed footprint (p0x-x+1 | (p1x-x)<<2 | (p0y-y+1)<<4 | (p1y-y)<<6)
    int* side_view;           // view of the side entry
    int side_capacity;
    // guided passes: add the pass's samples to out_film and accumulate luminance moments
    int accumulate;
    double* m1;
    double* m2;
};

struct LaunchCfg {
    int block;                // threads per block for traversal kernels
    int scene_stack, mesh_stack;
    bool priv_stack;          // stacks in private (scratch) memory instead of LDS
    bool refill;              // per-lane refill in the persistent traversal kernels
    bool counters;
    int trav_grid;            // persistent traversal grids (blocks): k_extend, k_shadow, k_mis
    int shadow_grid, mis_grid;
    int path_grid;            // grid-stride grid of per-path kernels (blocks of 256)
    hipStream_t stream;
    bool guided;              // launch the guide stages
    GuideNet net;
    int guide_grid;
};

void launch_sample_perms(const Chunk& host_chunk, const Chunk* dev_chunk, hipStream_t s);
void launch_raygen(const Chunk* c, long long n_paths_max, const Paths& p, Queues q, hipStream_t s);
// Stages of one bounce-loop iteration, in launch order.
enum IterStage {
    kStageReset, kStageExtend, kStageShade, kStageGuideForward, kStageGuideSample, kStageBin, kStageShadow, kStageMis,
    kStageResolve, kStageCount
};
void launch_stage(int stage, const DevScene& sc, const Chunk* c, const Paths& p, Queues q, int parity, const LaunchCfg& cfg);
void launch_iteration(const DevScene& sc, const Chunk* c, const Paths& p, Queues q, int parity, const LaunchCfg& cfg);
void launch_film(const Chunk* c, int n_pixels, const Paths& p, FilmParams f, Queues q, hipStream_t s);
void launch_trace_rays(const DevScene& sc, const float4* rays, int64_t n, int any, float4* out_tb, int2* out_ot,
                       unsigned long long* totals, const LaunchCfg& cfg);
int traversal_blocks_per_cu(const LaunchCfg& cfg, int kernel);

}  // namespace yg

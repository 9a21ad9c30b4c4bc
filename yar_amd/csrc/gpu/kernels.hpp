// kernels.hpp - host-visible declarations of the wavefront path-tracer kernels (render.hip) used by
// the context code (context.hip).
#pragma once

#include "device.hpp"

namespace yg {

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
    uint64_t* rng;       // per-pixel fallback stream (sequential sample mode only)
    float4* radiance;    // final per-sample radiance of the chunk, [pixel][sample]
};

enum : uint32_t {
    kDelta = 1u << 24,        // current ray was produced by a delta bounce (path_tracing.cpp:104)
    kPendingDL = 1u << 25,    // direct-lighting halves must be added by the resolve pass
    kDone = 1u << 26,         // path finished; finalize after the resolve pass
};

// Work queues of one bounce.
struct Queues {
    uint32_t* active_in;
    uint32_t* active_out;
    uint32_t* shadow_path; float4* shadow_o; float4* shadow_d; float4* shadow_c;
    uint32_t* mis_path; float4* mis_o; float4* mis_d; float4* mis_f; float* mis_ndw;
    uint32_t* dl_path;
    uint32_t* counters;       // [0] active_out, [1] shadow, [2] mis, [3] dl, [4] side entries, [5] active_in
    unsigned long long* totals;   // 16 traversal counters: 4 per kernel kind (extend, shadow, mis, trace_rays)
};

// Camera, sampler and integrator parameters of one chunk of pixels.
struct Chunk {
    const int2* pixels;       // film coordinates of the chunk's sample pixels
    int n_pixels;
    int spp, s_par, pass;     // samples per pixel; samples in flight per pixel; pass index
    int nx, ny, d1, d2;       // stratification grid and sample-vector dimensions
    const float* table;       // per-pixel sample tables, `stride` floats each
    long long stride;
    uint32_t seed_offset;
    M34 c2w;
    float W, H, horz, vert;
    int zup;
    int max_bounces, rr_bounce_threshold;
    float rr_threshold, film_scale, max_component;
};

// Film accumulation of one chunk (whole 64x64 tiles).  Every sample's pixel footprint is computed
// exactly as Film_Tile::add_sample does (film.cpp:27-50); because film_pos = pixel + jitter is a float
// sum, a sample can reach the 3x3 neighbourhood of its pixel.  Contributions inside the source's tile
// are gathered on the device in the tile's visitation order; contributions to a tile's one-pixel
// margin go to a side list that the host adds to the Film_Tile.
struct FilmAcc {
    float4* out_film;         // per listed pixel: its Film_Tile value inside its own tile
    double* out_var;          // per listed pixel: luminance variance of its estimate
    uint8_t* spill;           // per chunk pixel: 8-neighbour footprint bits of its samples
    int base;                 // index of the chunk's first pixel in the pixel list
    const int* tile_first;    // per film tile: pixel-list index of its first pixel, -1 if not rendered
    int sx0, sy0, sx1, sy1, tgx;   // sample region and tile grid (Film::Film, film.cpp:52-71)
    int rx0, ry0, rx1, ry1;        // render region (clamps tile pixel bounds)
    float4* side_rad;         // margin contributions: radiance
    int4* side_info;          // x, y, sample, packed footprint (p0x-x+1 | (p1x-x)<<2 | (p0y-y+1)<<4 | (p1y-y)<<6)
    int side_capacity;
};

struct LaunchCfg {
    int block;                // threads per block for traversal kernels
    int scene_stack, mesh_stack;
    bool counters;
    hipStream_t stream;
};

void launch_sample_table(const Chunk& c, float* table, hipStream_t s);
void launch_raygen(const Chunk& c, const Paths& p, Queues q, hipStream_t s);
void launch_extend(const DevScene& sc, const Paths& p, Queues q, int count, const LaunchCfg& cfg);
void launch_shade(const DevScene& sc, const Chunk& c, const Paths& p, Queues q, int count, hipStream_t s);
void launch_shadow(const DevScene& sc, const Paths& p, Queues q, int count, const LaunchCfg& cfg);
void launch_mis(const DevScene& sc, const Paths& p, Queues q, int count, const LaunchCfg& cfg);
void launch_resolve(const DevScene& sc, const Chunk& c, const Paths& p, Queues q, int count, hipStream_t s);
void launch_film(const Chunk& c, const Paths& p, FilmAcc f, Queues q, hipStream_t s);
void launch_trace_rays(const DevScene& sc, const float4* rays, int64_t n, int any, float4* out_tb, int2* out_ot,
                       unsigned long long* totals, const LaunchCfg& cfg);

}  // namespace yg

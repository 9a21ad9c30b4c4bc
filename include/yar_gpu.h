/*
 * yar_gpu.h - C ABI of the MI355X path-tracing core.
 *
 * This is the drop-in boundary for the per-tile render loop of YAR's CPU reference renderer
 * ("RAY", src/ref).  Every entry point is plain C: pointers, sizes and status codes; no C++ or
 * torch types cross it.  All functions return YAR_OK (0) on success and a non-zero status
 * otherwise; the message of the last failure is available from yar_last_error() (scene, kd-tree
 * and generator calls, thread-local) or yar_gpu_last_error(ctx) (context calls).  Nothing in the
 * library calls exit() or aborts: HIP failures become YAR_ERR_HIP.
 *
 * Reference interfaces replaced (file:line in kennyalive/YAR):
 *   yar_scene_load_pbrt     Scene load_scene(const std::string&)            src/lib/scene_loader.cpp:74
 *                           (pbrt3 subset, load_pbrt_scene src/lib/pbrt_scene.cpp:1050)
 *   yar_kdtree_build_*      build_triangle_mesh_kdtree / build_scene_kdtree src/ref/kdtree_builder.cpp:425,443
 *   yar_kdtree_save         KdTree::save (.kdtree cache file)               src/ref/kdtree.cpp:140
 *   yar_gpu_create          init_scene_context                             src/ref/reference_renderer.h:48
 *   yar_gpu_set_views       Scene_Overrides (camera pose / spp override)   src/ref/reference_renderer.h:23
 *   yar_gpu_render          Image render_scene(const Scene_Context&, double* variance, float* time)
 *                                                                          src/ref/reference_renderer.h:55
 *   yar_gpu_render_tiles    static Film_Tile render_tile(...)              src/ref/reference_renderer.cpp:331
 *   yar_film_tile_bounds    Film::get_tile_bounds                          src/ref/film.cpp:73
 *   yar_gpu_trace_rays      KdTree::intersect / intersect_any (scene tree)  src/ref/kdtree.cpp:219,329
 *   yar_gpu_render_batch    (batched multi-view form; Scene_Overrides applied to many views)
 */
#ifndef YAR_GPU_H
#define YAR_GPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define YAR_OK                   0
#define YAR_ERR_INVALID_ARGUMENT 1
#define YAR_ERR_IO               2
#define YAR_ERR_PARSE            3
#define YAR_ERR_UNSUPPORTED      4
#define YAR_ERR_HIP              5
#define YAR_ERR_OUT_OF_MEMORY    6
#define YAR_ERR_CAPACITY         7
#define YAR_ERR_INTERNAL         8       /* any other failure inside the library (never an abort) */

/* ---- scene description (host memory) ---------------------------------------------------
 * Mirrors the parts of the reference Scene (src/lib/scene.h:29) the RAY path tracer reads.
 * Matrices are row-major 3x4 (Matrix3x4::a[3][4], src/lib/matrix.h:7). */

typedef struct yar_mesh_desc {          /* Triangle_Mesh, src/lib/triangle_mesh.h:13 */
    const float*   vertices;            /* 3 * vertex_count */
    const float*   normals;             /* 3 * vertex_count or NULL */
    const float*   uvs;                 /* 2 * vertex_count or NULL */
    const int32_t* indices;             /* 3 * triangle_count */
    int32_t vertex_count;
    int32_t triangle_count;
    int32_t reverse_geometric_normal_orientation;
    int32_t reserved;
} yar_mesh_desc;

#define YAR_MATERIAL_NONE    0           /* area-light objects carry no material */
#define YAR_MATERIAL_DIFFUSE 1           /* Diffuse_Material  -> Diffuse_BRDF   (bsdf.cpp:60) */
#define YAR_MATERIAL_METAL   2           /* Metal_Material    -> Metal_BRDF     (bsdf.cpp:146) */
#define YAR_MATERIAL_GLASS   3           /* Glass_Material, roughness 0 -> delta (delta_scattering.cpp:53) */
#define YAR_MATERIAL_MIRROR  4           /* Perfect_Reflector_Material -> delta reflection (delta_scattering.cpp:27) */
#define YAR_MATERIAL_PLASTIC 5           /* Plastic_Material -> Pbrt3_Plastic_BRDF (bsdf_pbrt.cpp:258) in pbrt3
                                          * scenes, Plastic_BRDF (bsdf.cpp:204) otherwise */
#define YAR_MATERIAL_COATED_DIFFUSE 6    /* Coated_Diffuse_Material (pbrt "substrate") ->
                                          * Ashikhmin_Shirley_Phong_BRDF (bsdf.cpp:427) */
#define YAR_MATERIAL_PBRT3_UBER 7        /* Pbrt3_Uber_Material -> Pbrt3_Uber_BRDF (bsdf_pbrt.cpp:14-91) for
                                            Kd/Ks, delta reflection Kr and opacity pass-through chosen per
                                            component (delta_scattering.cpp:102-152); Kt must be black */
#define YAR_MATERIAL_MIX     8           /* Mix_Material of two sub-materials (matte, metal, plastic, substrate,
                                            glass, mirror; at most one of them a delta layer): Mix_BSDF
                                            (bsdf.cpp:497-570), or with a delta layer that layer picked per
                                            sample (delta_scattering.cpp:154-189) and the other one's BSDF
                                            scaled (Scaled_BSDF, bsdf.cpp:572-629) */

typedef struct yar_material_desc {
    int32_t type;
    int32_t roughness_is_alpha;         /* pbrt "remaproughness" false */
    float   reflectance[3];             /* matte/plastic/substrate/uber Kd, glass/mirror Kr */
    float   transmittance[3];           /* glass Kt */
    float   eta[3];                     /* metal eta (RGB) */
    float   k[3];                       /* metal k (RGB) */
    float   eta_i;                      /* metal: IOR of adjacent dielectric (1 for pbrt) */
    float   u_roughness;                /* metal/substrate uroughness, glass/plastic roughness */
    float   v_roughness;                /* metal/substrate vroughness */
    float   index_of_refraction;        /* glass, uber */
    float   specular_r0;                /* plastic: specular scale r0 (luminance of pbrt Ks) */
    float   specular[3];                /* coated diffuse: Fresnel r0 per channel (pbrt substrate Ks);
                                           uber Ks */
    float   delta_reflectance[3];       /* uber Kr */
    float   opacity[3];                 /* uber opacity (1 = opaque) */
    int32_t mix_materials[2];           /* mix: the two sub-materials (indices into materials) */
    float   mix_amount[3];              /* mix: pbrt "amount", the weight of the first sub-material */
} yar_material_desc;

typedef struct yar_object_desc {        /* Scene_Object, src/lib/scene_object.h:8 */
    int32_t mesh;                       /* index into meshes */
    int32_t material;                   /* index into materials, -1 for area lights */
    int32_t area_light;                 /* index into the area lights in the order rect_lights,
                                           sphere_lights, mesh_lights (rect_light_count + i for
                                           sphere_lights[i], rect_light_count + sphere_light_count
                                           + i for mesh_lights[i]), -1 if none */
    int32_t reserved;
    float   object_to_world[12];
    float   world_to_object[12];
} yar_object_desc;

typedef struct yar_rect_light_desc {    /* Diffuse_Rectangular_Light, src/lib/light.h:69 */
    float light_to_world[12];
    float emitted_radiance[3];
    float size[2];
    int32_t sample_count;
} yar_rect_light_desc;

typedef struct yar_point_light_desc {   /* Point_Light, src/lib/light.h:35 */
    float position[3];
    float intensity[3];
} yar_point_light_desc;

typedef struct yar_spot_light_desc {    /* Spot_Light, src/lib/light.h:40 */
    float position[3];
    float direction[3];                 /* unit axis, from the light towards what it lights */
    float cone_angle;                   /* radians */
    float penumbra_angle;               /* radians, measured inwards from the cone edge */
    float intensity[3];
} yar_spot_light_desc;

typedef struct yar_directional_light_desc {   /* Directional_Light, src/lib/light.h:48 */
    float direction[3];                 /* unit vector towards the light */
    float irradiance[3];                /* on a surface perpendicular to direction */
} yar_directional_light_desc;

/* Diffuse_Triangle_Mesh_Light (src/lib/light.h:85): a one-sided emitter mesh.  Light samples are
 * taken in the mesh's own coordinates, as Diffuse_Triangle_Mesh_Light_Sampler does
 * (light_sampling.cpp:126-184), so the emitter object's transform must be the identity. */
typedef struct yar_mesh_light_desc {
    int32_t object;                     /* the emitter object */
    float   emitted_radiance[3];
} yar_mesh_light_desc;

/* Diffuse_Sphere_Light (src/lib/light.h:78): emits on its whole surface.  Light samples are taken
 * on the analytic sphere (position, radius) by uniform cone sampling (light_sampling.cpp:76-121);
 * BSDF samples count where they hit the emitter object, the tessellated sphere. */
typedef struct yar_sphere_light_desc {
    int32_t object;                     /* the emitter object */
    float   position[3];                /* world-space centre */
    float   radius;
    float   emitted_radiance[3];
    int32_t sample_count;               /* pbrt "nsamples" (the single-sample estimator ignores it) */
} yar_sphere_light_desc;

/* Environment_Light (src/lib/light.h:91) with a constant map: pbrt "infinite" without "mapname".
 * Radiance `radiance` from every direction; light samples are drawn uniformly in the light frame's
 * (phi, theta) square, as the reference's one-texel Distribution_2D does (light_sampling.cpp:27-73). */
typedef struct yar_environment_light_desc {
    float   light_to_world[12];
    float   world_to_light[12];         /* get_inverse_transform of light_to_world */
    float   radiance[3];                /* pbrt L * scale */
    int32_t sample_count;
} yar_environment_light_desc;

typedef struct yar_render_config {      /* Raytracer_Config, src/lib/raytracer_config.h:5 */
    int32_t film_width, film_height;
    int32_t region_x0, region_y0, region_x1, region_y1;   /* render region, exclusive p1 */
    int32_t x_pixel_sample_count, y_pixel_sample_count;
    int32_t max_light_bounces;
    int32_t russian_roulette_bounce_count_threshold;
    float   russian_roulette_threshold;
    float   film_radiance_scale;
    float   max_rgb_component_value_of_film_sample;        /* +inf when unset */
    float   pixel_filter_radius;                           /* box filter only */
    uint32_t rng_seed_offset;                              /* --seedoffset */
    int32_t pbrt3_scene;                                   /* enables the pbrt3 roughness remap */
} yar_render_config;

typedef struct yar_camera_desc {
    float   camera_to_world[12];        /* columns: right, up/forward, back/up, position */
    float   fov_y_degrees;
    int32_t z_is_up;
} yar_camera_desc;

typedef struct yar_scene_desc {
    const yar_mesh_desc*        meshes;       int32_t mesh_count;
    const yar_material_desc*    materials;    int32_t material_count;
    const yar_object_desc*      objects;      int32_t object_count;
    const yar_rect_light_desc*  rect_lights;  int32_t rect_light_count;
    const yar_point_light_desc* point_lights; int32_t point_light_count;
    yar_camera_desc   camera;
    yar_render_config config;
    const yar_spot_light_desc*        spot_lights;        int32_t spot_light_count;
    const yar_directional_light_desc* directional_lights; int32_t directional_light_count;
    const yar_mesh_light_desc*        mesh_lights;        int32_t mesh_light_count;
    const yar_sphere_light_desc*      sphere_lights;      int32_t sphere_light_count;
    const yar_environment_light_desc* environment_light;  int32_t environment_light_count;   /* 0 or 1 */
} yar_scene_desc;

/* ---- scene loading / generation -------------------------------------------------------- */
typedef struct yar_scene yar_scene;

/* Loads the pbrt3 subset the RAY path tracer consumes (trianglemesh; matte, metal, glass, mirror,
 * plastic, substrate; sphere shapes, tessellated as the reference does; diffuse area lights
 * (rectangles, spheres and triangle meshes), point, spot and distant lights, infinite lights without
 * a map; perspective camera, Film, Sampler, Integrator "path", PixelFilter "box") with the
 * reference's import semantics. */
int  yar_scene_load_pbrt(const char* path, yar_scene** out_scene);
const yar_scene_desc* yar_scene_get_desc(const yar_scene* scene);
void yar_scene_free(yar_scene* scene);
const char* yar_last_error(void);

/* Writes the seeded benchmark scene `config` (1..5) as a pbrt3 file.  `scale` < 1 shrinks
 * triangle counts for quick tests (1 = the full configuration). */
int  yar_scene_generate(int32_t config, uint64_t seed, double scale, const char* pbrt_path);

/* ---- kd-tree builder (host) -------------------------------------------------------------
 * Byte-compatible with the reference KdNode encoding (src/ref/kdtree.h:18): node = 2 x uint32,
 * interior word0 = axis | above_child << 2, word1 = split (f32 bits); leaf word0 = 3 | count << 2,
 * word1 = first primitive index, the rest packed in the following words. */
typedef struct yar_kdtree yar_kdtree;
int  yar_kdtree_build_mesh(const float* vertices, const int32_t* indices, int32_t triangle_count,
                           yar_kdtree** out_tree);
/* Binned fast path (no reference counterpart): nodes above 1024 triangles choose their plane among 31
 * uniform candidates per axis with the same SAH cost, partition and clipping; the tree differs from the
 * reference's but lists every triangle that overlaps a leaf, so traversal finds the same hits. */
int  yar_kdtree_build_mesh_binned(const float* vertices, const int32_t* indices, int32_t triangle_count,
                                  yar_kdtree** out_tree);
/* boxes: 6 floats per primitive (min xyz, max xyz); builds the object-level (scene) tree. */
int  yar_kdtree_build_boxes(const float* boxes, int32_t count, yar_kdtree** out_tree);
uint32_t        yar_kdtree_node_count(const yar_kdtree* tree);
const uint32_t* yar_kdtree_nodes(const yar_kdtree* tree);       /* 2 * node_count words */
void            yar_kdtree_bounds(const yar_kdtree* tree, float out_bounds[6]);
int32_t         yar_kdtree_max_depth(const yar_kdtree* tree);    /* deepest leaf (root = 0) */
int  yar_kdtree_save(const yar_kdtree* tree, const char* path);  /* reference .kdtree layout */
void yar_kdtree_free(yar_kdtree* tree);

/* ---- GPU context --------------------------------------------------------------------- */
typedef struct yar_gpu_ctx yar_gpu_ctx;

typedef struct yar_gpu_options {
    int32_t device;                     /* HIP device ordinal */
    int32_t block_size;                 /* traversal threads per block: 0 (= 64), 64, 128, 192 or 256 */
    int64_t max_paths_in_flight;        /* 0 = default: 64M plain, 16M guided, at most a quarter of free HBM */
    int32_t traversal_stack_depth;      /* short-stack entries per tree level, 0 = 8 (kd-restart beyond) */
    int32_t enable_counters;            /* count node visits / triangle tests per ray */
    int32_t use_graph;                  /* capture the bounce loop in a hipGraph */
    int32_t enable_timing;              /* per-kernel HIP-event timing in yar_render_stats */
    int32_t stack_mode;                 /* traversal stack storage: 0 = default, 1 = LDS, 2 = private */
    int32_t lane_refill;                /* 1 = traversal lanes take new rays individually, 0/2 = batches of 64 */
    int32_t disable_mis_cull;           /* 1 = trace every MIS BSDF ray (no analytic light-rectangle test) */
    int32_t bin_rays;                   /* 1, 2, 3 = sort each bounce's rays by direction cell (2x2, 4x4, 8x8 per cube face) within windows of 4096 before traversal */
    int32_t kdtree_builder;             /* 0 = exact (the reference's trees), 1 = binned SAH fast path */
    int32_t reserved[3];
} yar_gpu_options;

typedef struct yar_view {               /* Scene_Overrides analog, reference_renderer.h:23 */
    float    camera_to_world[12];
    float    fov_y_degrees;
    int32_t  z_is_up;
    int32_t  film_width, film_height;
    int32_t  region_x0, region_y0, region_x1, region_y1;
    int32_t  x_pixel_sample_count, y_pixel_sample_count;
    uint32_t rng_seed_offset;
    int32_t  reserved;
} yar_view;

typedef struct yar_film_pixel {         /* Film_Pixel, src/ref/film.h:26 (16 bytes) */
    float color_sum[3];
    float weight_sum;
} yar_film_pixel;

typedef struct yar_ray {                /* Ray + tmax, 32 bytes */
    float origin[3];
    float tmax;                         /* closest: initial Intersection::t; any-hit: ray_tmax */
    float direction[3];
    uint32_t reserved;
} yar_ray;

typedef struct yar_hit {                /* Intersection subset, 24 bytes */
    float   t;                          /* +inf on miss (closest); 1/0 in t for any-hit */
    int32_t object;                     /* scene object index, -1 on miss */
    int32_t triangle;                   /* mesh-local triangle index */
    float   barycentrics[3];
} yar_hit;

typedef struct yar_render_stats {
    double  render_seconds;             /* host wall time of the call */
    double  gpu_seconds;                /* HIP-event time of the bounce loop */
    int64_t samples;                    /* camera samples (pixels * spp) */
    int64_t closest_rays;               /* scene-tree closest-hit queries (trace_ray + MIS) */
    int64_t any_rays;                   /* shadow (any-hit) queries */
    int64_t node_visits;                /* only with enable_counters: all traversal kernels */
    int64_t leaf_refs;
    int64_t triangle_tests;
    int64_t object_visits;
    int64_t extend_node_visits;         /* only with enable_counters: extend (path-ray) kernel */
    int64_t extend_leaf_refs;
    int64_t extend_triangle_tests;
    int64_t extend_object_visits;
    int64_t extend_rays;                /* rays traced by the extend kernel */
    int64_t extend_launches;
    double  extend_seconds, shade_seconds, shadow_seconds, mis_seconds, other_seconds;  /* enable_timing */
    double  guide_seconds;              /* enable_timing: guide network + guided sampling stages */
    double  train_seconds;              /* guided: host wall time of the training between passes */
    int64_t train_records;              /* guided: training records collected over all passes */
    int64_t passes;                     /* guided: sample passes (1 for plain renders) */
    double  train_nats;                 /* guided: last batch's log-likelihood gain over uniform */
    int64_t mis_culled_rays;            /* only with enable_counters: MIS BSDF rays toward a light that the
                                           analytic light-rectangle test answered without a traversal */
} yar_render_stats;

int  yar_gpu_create(const yar_scene_desc* scene, const yar_gpu_options* options, yar_gpu_ctx** out_ctx);
void yar_gpu_destroy(yar_gpu_ctx* ctx);
const char* yar_gpu_last_error(const yar_gpu_ctx* ctx);

/* View 0 is the scene's own camera/config; set_views replaces the list (n >= 1). */
int  yar_gpu_set_views(yar_gpu_ctx* ctx, const yar_view* views, int32_t count);
int  yar_gpu_get_view(const yar_gpu_ctx* ctx, int32_t index, yar_view* out_view);

/* Renders the whole render region of a view.  out_film (region_w*region_h Film_Pixels, merged
 * in tile-index order like Film::merge_tile) and out_image (region_w*region_h*3 floats,
 * Film::get_image) may each be NULL. */
int  yar_gpu_render(yar_gpu_ctx* ctx, int32_t view, yar_film_pixel* out_film, float* out_image,
                    double* out_variance, yar_render_stats* out_stats);

/* Tile count / bounds of a view's film (Tile_Size 64, film.cpp:5). Bounds are {x0,y0,x1,y1}. */
int  yar_film_tile_count(const yar_gpu_ctx* ctx, int32_t view, int32_t* out_count);
int  yar_film_tile_bounds(const yar_gpu_ctx* ctx, int32_t view, int32_t tile_index,
                          int32_t out_sample_bounds[4], int32_t out_pixel_bounds[4]);
/* Renders the given tiles; out_tiles[i] must hold pixel_bounds area Film_Pixels of tile i, and
 * out_tile_variance[i] receives the tile's summed per-pixel luminance variance. */
int  yar_gpu_render_tiles(yar_gpu_ctx* ctx, int32_t view, const int32_t* tile_indices, int32_t count,
                          yar_film_pixel* const* out_tiles, double* out_tile_variance,
                          yar_render_stats* out_stats);

/* Batched form: renders several views in one persistent bounce loop.  out_images[i] receives the
 * view's resolved image (region area * 3 floats) and may be NULL. */
int  yar_gpu_render_batch(yar_gpu_ctx* ctx, const int32_t* views, int32_t count,
                          float* const* out_images, double* out_variances, yar_render_stats* out_stats);

/* Scene-tree ray queries. Host buffers; any_hit != 0 runs intersect_any (t = 1 if occluded). */
int  yar_gpu_trace_rays(yar_gpu_ctx* ctx, const yar_ray* rays, int64_t count, int32_t any_hit,
                        yar_hit* out_hits);
/* Same, on device-resident buffers and the given HIP stream (NULL = the context's stream). */
int  yar_gpu_trace_rays_device(yar_gpu_ctx* ctx, const yar_ray* d_rays, int64_t count, int32_t any_hit,
                               yar_hit* d_hits, void* hip_stream);
/* Allocation-free form of yar_gpu_trace_rays_device.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It calls no allocating routine; the per-ray hit staging and the traversal
 * counters live in a static device array of 24 MB (each device the library runs on holds its own copy
 * until the process ends), so count must not exceed 2^20 rays (YAR_ERR_CAPACITY otherwise).  Every
 * call on one device shares that array: calls that overlap on one device race. */
int  yar_gpu_trace_rays_device_noalloc(yar_gpu_ctx* ctx, const yar_ray* d_rays, int64_t count, int32_t any_hit,
                                       yar_hit* d_hits, void* hip_stream);
/* Per-ray traversal counters of the last trace_rays call (requires enable_counters). */
int  yar_gpu_last_trace_counters(yar_gpu_ctx* ctx, yar_render_stats* out_stats);

/* Device/scene summary: bytes of HBM used by the flattened scene, tree depths. */
typedef struct yar_gpu_info {
    int64_t scene_bytes;
    int64_t triangle_count;
    int32_t object_count;
    int32_t scene_tree_depth;
    int32_t mesh_tree_depth;
    int32_t stack_depth;
    double  build_seconds;
    double  upload_seconds;
    char    device_name[64];
} yar_gpu_info;
int  yar_gpu_get_info(const yar_gpu_ctx* ctx, yar_gpu_info* out_info);

/* ---- multi-session checkpoints and image output -------------------------------------------
 * The reference's checkpoint directory format (src/ref/reference_renderer.cpp:162-324): a text
 * file "checkpoint" (input file name, total tile count, samples per pixel) and one binary file
 * "tile_NNNN" per finished tile (render time f32, tile variance f64, pixel bounds 4 x i32, Film_Pixels),
 * written as "temp_tile_NNNN" and renamed. */
typedef struct yar_checkpoint yar_checkpoint;
/* start_or_resume_checkpoint (reference_renderer.cpp:179): creates the directory and its metadata
 * when absent or empty, else checks the metadata against the project and loads every finished tile. */
int  yar_checkpoint_open(const char* dir, const char* input_filename, int32_t total_tile_count,
                         int32_t samples_per_pixel, yar_checkpoint** out_checkpoint);
int32_t yar_checkpoint_finished_count(const yar_checkpoint* checkpoint);
int  yar_checkpoint_finished_tiles(const yar_checkpoint* checkpoint, int32_t* out_indices);   /* ascending */
int  yar_checkpoint_tile(const yar_checkpoint* checkpoint, int32_t tile_index, int32_t out_pixel_bounds[4],
                         yar_film_pixel* out_pixels, double* out_variance);
float yar_checkpoint_previous_time(const yar_checkpoint* checkpoint);   /* time spent in earlier sessions */
void yar_checkpoint_free(yar_checkpoint* checkpoint);
/* write_tile_to_checkpoint_directory (reference_renderer.cpp:283). */
int  yar_checkpoint_write_tile(const char* dir, int32_t tile_index, const int32_t pixel_bounds[4],
                               const yar_film_pixel* pixels, double tile_variance, float render_time);

/* render_scene with a checkpoint directory (reference_renderer.cpp:479-584): loads the finished
 * tiles, renders the others on the GPU in batches of tiles_per_batch (0 = 64), writes each tile's
 * file as its batch completes, then merges every tile in index order.  out_render_time = time of
 * earlier sessions + this call. */
int  yar_gpu_render_checkpointed(yar_gpu_ctx* ctx, int32_t view, const char* checkpoint_dir, const char* input_filename,
                                 int32_t tiles_per_batch, yar_film_pixel* out_film, float* out_image,
                                 double* out_variance, float* out_render_time, yar_render_stats* out_stats);

typedef struct yar_exr_attributes {     /* EXR_Attributes + EXR_Write_Params, reference_renderer.h:29-46 */
    const char* input_file;
    int32_t spp;
    float   variance;
    float   load_time;                  /* written only with enable_varying_attributes */
    float   render_time;
    int32_t enable_varying_attributes;
    int32_t dump_attributes;            /* also write "<name> <value>" lines to <image>.txt */
} yar_exr_attributes;
/* write_openexr_image (reference_renderer.cpp:673) / Image::write_exr (src/lib/image.cpp:163): HALF
 * channels B, G, R, uncompressed or ZIP, with the yar_* attributes; rgb is width*height*3 floats,
 * top row first. */
int  yar_image_write_exr(const char* path, const float* rgb, int32_t width, int32_t height,
                         const yar_exr_attributes* attributes, int32_t compress_zip);

/* ---- learned directional proposal (guided form) ------------------------------------------
 * No counterpart in the reference: a small MLP (9 inputs -> 16-frequency encoding, 297 features ->
 * 4 x 64 ReLU -> 40 outputs = 8 von Mises-Fisher lobes) with bf16 weights and fp32 accumulation,
 * trained online with Adam on the tracer's own radiance samples.  Inputs are position in scene
 * bounds, outgoing direction and shading normal, each mapped to [0,1]^3.  The parameters travel in
 * a canonical fp32 order (W1[304][64], b1, W2[64][64], b2, W3, b3, W4, b4, W5[64][48], b5; padding
 * rows/columns of W1 and W5 are zero). */
typedef struct yar_guide yar_guide;

typedef struct yar_guide_config {
    uint32_t seed;                      /* parameter initialisation and batch selection */
    int32_t  batch_size;                /* training samples per Adam step (0 = 16384) */
    float    learning_rate;             /* 0 = 5e-3 */
    float    beta1, beta2, epsilon;     /* 0 = 0.9, 0.99, 1e-8 */
    int32_t  steps_per_pass;            /* Adam steps after each guided render pass (0 = 32) */
    float    guide_fraction;            /* one-sample MIS probability of the learned mixture (0 = 0.5) */
    int32_t  first_pass_spp;            /* samples per pixel of the first (untrained) pass (0 = 8) */
    int32_t  record_capacity;           /* training records kept per pass (0 = 2^22) */
    int32_t  pass_count;                /* sample passes per render (0 = 2): the first pass, then the rest
                                           split into passes of doubling size, training after each */
    int32_t  reserved[5];
} yar_guide_config;

int64_t yar_guide_param_count(void);
int  yar_guide_create(int32_t device, const yar_guide_config* config, yar_guide** out_guide);
void yar_guide_destroy(yar_guide* guide);
const char* yar_guide_last_error(const yar_guide* guide);
int  yar_guide_get_params(yar_guide* guide, float* out_params);
/* Replaces the parameters (padding forced to zero) and resets the Adam state. */
int  yar_guide_set_params(yar_guide* guide, const float* params);
/* Network outputs (40 floats per query) for n queries of 9 inputs. */
int  yar_guide_forward(yar_guide* guide, const float* inputs, int64_t n, float* out_outputs);
/* Mean loss -wt log q(dir | inputs) over the n samples and its gradient in canonical order. */
int  yar_guide_gradients(yar_guide* guide, const float* inputs, const float* dirs, const float* weights, int64_t n,
                         float* out_gradient, float* out_loss);
/* `steps` Adam steps on batches drawn from the given samples; out_loss = loss of the last batch. */
int  yar_guide_train(yar_guide* guide, const float* inputs, const float* dirs, const float* weights, int64_t n,
                     int32_t steps, float* out_loss);
/* Mixture density at given directions / directions sampled with 3 uniforms per query. */
int  yar_guide_pdf(yar_guide* guide, const float* inputs, const float* dirs, int64_t n, float* out_pdf);
int  yar_guide_sample(yar_guide* guide, const float* inputs, const float* uniforms, int64_t n, float* out_dirs,
                      float* out_pdf);

/* Guided form of yar_gpu_render_batch: renders the views in pass_count sample passes (first_pass_spp
 * samples per pixel with BSDF sampling alone, then the rest in passes of doubling size), trains the
 * guide on the radiance samples of the passes so far between passes, and in the later passes traces
 * the continuation at diffuse vertices with one-sample MIS between the guide's mixture and the
 * BSDF.  The image converges to the plain render's mean. */
int  yar_gpu_render_guided(yar_gpu_ctx* ctx, yar_guide* guide, const int32_t* views, int32_t count,
                           float* const* out_images, double* out_variances, yar_render_stats* out_stats);

#ifdef __cplusplus
}
#endif
#endif /* YAR_GPU_H */

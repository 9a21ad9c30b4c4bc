"""ctypes mirror of include/yar_gpu.h (the C ABI of the path-tracing core)."""
import ctypes as C

YAR_OK = 0
YAR_MATERIAL_NONE, YAR_MATERIAL_DIFFUSE, YAR_MATERIAL_METAL, YAR_MATERIAL_GLASS = 0, 1, 2, 3
YAR_MATERIAL_MIRROR, YAR_MATERIAL_PLASTIC, YAR_MATERIAL_COATED_DIFFUSE, YAR_MATERIAL_PBRT3_UBER = 4, 5, 6, 7
YAR_MATERIAL_MIX = 8


class MeshDesc(C.Structure):
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("uvs", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_int32)),
                ("vertex_count", C.c_int32), ("triangle_count", C.c_int32),
                ("reverse_geometric_normal_orientation", C.c_int32), ("reserved", C.c_int32)]


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("roughness_is_alpha", C.c_int32),
                ("reflectance", C.c_float * 3), ("transmittance", C.c_float * 3),
                ("eta", C.c_float * 3), ("k", C.c_float * 3), ("eta_i", C.c_float),
                ("u_roughness", C.c_float), ("v_roughness", C.c_float),
                ("index_of_refraction", C.c_float), ("specular_r0", C.c_float),
                ("specular", C.c_float * 3), ("delta_reflectance", C.c_float * 3), ("opacity", C.c_float * 3),
                ("mix_materials", C.c_int32 * 2), ("mix_amount", C.c_float * 3)]


class ObjectDesc(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("material", C.c_int32), ("area_light", C.c_int32),
                ("reserved", C.c_int32), ("object_to_world", C.c_float * 12),
                ("world_to_object", C.c_float * 12)]


class RectLightDesc(C.Structure):
    _fields_ = [("light_to_world", C.c_float * 12), ("emitted_radiance", C.c_float * 3),
                ("size", C.c_float * 2), ("sample_count", C.c_int32)]


class PointLightDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("intensity", C.c_float * 3)]


class SpotLightDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("cone_angle", C.c_float),
                ("penumbra_angle", C.c_float), ("intensity", C.c_float * 3)]


class DirectionalLightDesc(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("irradiance", C.c_float * 3)]


class MeshLightDesc(C.Structure):
    _fields_ = [("object", C.c_int32), ("emitted_radiance", C.c_float * 3)]


class SphereLightDesc(C.Structure):
    _fields_ = [("object", C.c_int32), ("position", C.c_float * 3), ("radius", C.c_float),
                ("emitted_radiance", C.c_float * 3), ("sample_count", C.c_int32)]


class EnvironmentLightDesc(C.Structure):
    _fields_ = [("light_to_world", C.c_float * 12), ("world_to_light", C.c_float * 12),
                ("radiance", C.c_float * 3), ("sample_count", C.c_int32)]


class RenderConfig(C.Structure):
    _fields_ = [("film_width", C.c_int32), ("film_height", C.c_int32),
                ("region_x0", C.c_int32), ("region_y0", C.c_int32),
                ("region_x1", C.c_int32), ("region_y1", C.c_int32),
                ("x_pixel_sample_count", C.c_int32), ("y_pixel_sample_count", C.c_int32),
                ("max_light_bounces", C.c_int32), ("russian_roulette_bounce_count_threshold", C.c_int32),
                ("russian_roulette_threshold", C.c_float), ("film_radiance_scale", C.c_float),
                ("max_rgb_component_value_of_film_sample", C.c_float),
                ("pixel_filter_radius", C.c_float), ("rng_seed_offset", C.c_uint32),
                ("pbrt3_scene", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("camera_to_world", C.c_float * 12), ("fov_y_degrees", C.c_float), ("z_is_up", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("meshes", C.POINTER(MeshDesc)), ("mesh_count", C.c_int32),
                ("materials", C.POINTER(MaterialDesc)), ("material_count", C.c_int32),
                ("objects", C.POINTER(ObjectDesc)), ("object_count", C.c_int32),
                ("rect_lights", C.POINTER(RectLightDesc)), ("rect_light_count", C.c_int32),
                ("point_lights", C.POINTER(PointLightDesc)), ("point_light_count", C.c_int32),
                ("camera", CameraDesc), ("config", RenderConfig),
                ("spot_lights", C.POINTER(SpotLightDesc)), ("spot_light_count", C.c_int32),
                ("directional_lights", C.POINTER(DirectionalLightDesc)), ("directional_light_count", C.c_int32),
                ("mesh_lights", C.POINTER(MeshLightDesc)), ("mesh_light_count", C.c_int32),
                ("sphere_lights", C.POINTER(SphereLightDesc)), ("sphere_light_count", C.c_int32),
                ("environment_light", C.POINTER(EnvironmentLightDesc)), ("environment_light_count", C.c_int32)]


class GpuOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("block_size", C.c_int32), ("max_paths_in_flight", C.c_int64),
                ("traversal_stack_depth", C.c_int32), ("enable_counters", C.c_int32),
                ("use_graph", C.c_int32), ("enable_timing", C.c_int32), ("stack_mode", C.c_int32),
                ("lane_refill", C.c_int32), ("disable_mis_cull", C.c_int32), ("bin_rays", C.c_int32),
                ("kdtree_builder", C.c_int32), ("reserved", C.c_int32 * 3)]


class View(C.Structure):
    _fields_ = [("camera_to_world", C.c_float * 12), ("fov_y_degrees", C.c_float), ("z_is_up", C.c_int32),
                ("film_width", C.c_int32), ("film_height", C.c_int32),
                ("region_x0", C.c_int32), ("region_y0", C.c_int32),
                ("region_x1", C.c_int32), ("region_y1", C.c_int32),
                ("x_pixel_sample_count", C.c_int32), ("y_pixel_sample_count", C.c_int32),
                ("rng_seed_offset", C.c_uint32), ("reserved", C.c_int32)]


class FilmPixel(C.Structure):
    _fields_ = [("color_sum", C.c_float * 3), ("weight_sum", C.c_float)]


class Ray(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("tmax", C.c_float), ("direction", C.c_float * 3),
                ("reserved", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("object", C.c_int32), ("triangle", C.c_int32),
                ("barycentrics", C.c_float * 3)]


class RenderStats(C.Structure):
    _fields_ = [("render_seconds", C.c_double), ("gpu_seconds", C.c_double), ("samples", C.c_int64),
                ("closest_rays", C.c_int64), ("any_rays", C.c_int64), ("node_visits", C.c_int64),
                ("leaf_refs", C.c_int64), ("triangle_tests", C.c_int64), ("object_visits", C.c_int64),
                ("extend_node_visits", C.c_int64), ("extend_leaf_refs", C.c_int64),
                ("extend_triangle_tests", C.c_int64), ("extend_object_visits", C.c_int64),
                ("extend_rays", C.c_int64), ("extend_launches", C.c_int64),
                ("extend_seconds", C.c_double), ("shade_seconds", C.c_double),
                ("shadow_seconds", C.c_double), ("mis_seconds", C.c_double), ("other_seconds", C.c_double),
                ("guide_seconds", C.c_double), ("train_seconds", C.c_double), ("train_records", C.c_int64),
                ("passes", C.c_int64), ("train_nats", C.c_double), ("mis_culled_rays", C.c_int64)]


class GuideConfig(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("batch_size", C.c_int32), ("learning_rate", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("steps_per_pass", C.c_int32), ("guide_fraction", C.c_float), ("first_pass_spp", C.c_int32),
                ("record_capacity", C.c_int32), ("pass_count", C.c_int32),
                ("reserved", C.c_int32 * 5)]


class GpuInfo(C.Structure):
    _fields_ = [("scene_bytes", C.c_int64), ("triangle_count", C.c_int64), ("object_count", C.c_int32),
                ("scene_tree_depth", C.c_int32), ("mesh_tree_depth", C.c_int32), ("stack_depth", C.c_int32),
                ("build_seconds", C.c_double), ("upload_seconds", C.c_double), ("device_name", C.c_char * 64)]


# numpy dtypes matching the 32-byte ray and 24-byte hit records
def ray_dtype():
    import numpy as np
    return np.dtype([("origin", "<f4", 3), ("tmax", "<f4"), ("direction", "<f4", 3), ("reserved", "<u4")])


def hit_dtype():
    import numpy as np
    return np.dtype([("t", "<f4"), ("object", "<i4"), ("triangle", "<i4"), ("barycentrics", "<f4", 3)])


def film_dtype():
    import numpy as np
    return np.dtype([("color_sum", "<f4", 3), ("weight_sum", "<f4")])


class ExrAttributes(C.Structure):
    _fields_ = [("input_file", C.c_char_p), ("spp", C.c_int32), ("variance", C.c_float), ("load_time", C.c_float),
                ("render_time", C.c_float), ("enable_varying_attributes", C.c_int32), ("dump_attributes", C.c_int32)]


_P = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "yar_last_error": (C.c_char_p, []),
    "yar_scene_load_pbrt": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "yar_scene_get_desc": (C.POINTER(SceneDesc), [_P]),
    "yar_scene_free": (None, [_P]),
    "yar_scene_generate": (C.c_int, [C.c_int32, C.c_uint64, C.c_double, C.c_char_p]),
    "yar_kdtree_build_mesh": (C.c_int, [_P, _P, C.c_int32, C.POINTER(_P)]),
    "yar_kdtree_build_mesh_binned": (C.c_int, [_P, _P, C.c_int32, C.POINTER(_P)]),
    "yar_kdtree_build_boxes": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "yar_kdtree_node_count": (C.c_uint32, [_P]),
    "yar_kdtree_nodes": (C.POINTER(C.c_uint32), [_P]),
    "yar_kdtree_bounds": (None, [_P, C.POINTER(C.c_float)]),
    "yar_kdtree_max_depth": (C.c_int32, [_P]),
    "yar_kdtree_save": (C.c_int, [_P, C.c_char_p]),
    "yar_kdtree_free": (None, [_P]),
    "yar_gpu_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(GpuOptions), C.POINTER(_P)]),
    "yar_gpu_destroy": (None, [_P]),
    "yar_gpu_last_error": (C.c_char_p, [_P]),
    "yar_gpu_set_views": (C.c_int, [_P, C.POINTER(View), C.c_int32]),
    "yar_gpu_get_view": (C.c_int, [_P, C.c_int32, C.POINTER(View)]),
    "yar_gpu_render": (C.c_int, [_P, C.c_int32, _P, _P, C.POINTER(C.c_double), C.POINTER(RenderStats)]),
    "yar_film_tile_count": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "yar_film_tile_bounds": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "yar_gpu_render_tiles": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(_P),
                                       C.POINTER(C.c_double), C.POINTER(RenderStats)]),
    "yar_gpu_render_batch": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32, C.POINTER(_P),
                                       C.POINTER(C.c_double), C.POINTER(RenderStats)]),
    "yar_gpu_trace_rays": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P]),
    "yar_gpu_trace_rays_device": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P]),
    "yar_gpu_trace_rays_device_noalloc": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P]),
    "yar_gpu_last_trace_counters": (C.c_int, [_P, C.POINTER(RenderStats)]),
    "yar_gpu_get_info": (C.c_int, [_P, C.POINTER(GpuInfo)]),
    "yar_gpu_render_guided": (C.c_int, [_P, _P, C.POINTER(C.c_int32), C.c_int32, _P, _P, _P]),
    "yar_checkpoint_open": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "yar_checkpoint_finished_count": (C.c_int32, [_P]),
    "yar_checkpoint_finished_tiles": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "yar_checkpoint_tile": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _P, C.POINTER(C.c_double)]),
    "yar_checkpoint_previous_time": (C.c_float, [_P]),
    "yar_checkpoint_free": (None, [_P]),
    "yar_checkpoint_write_tile": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), _P, C.c_double, C.c_float]),
    "yar_gpu_render_checkpointed": (C.c_int, [_P, C.c_int32, C.c_char_p, C.c_char_p, C.c_int32, _P, _P,
                                              C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(RenderStats)]),
    "yar_image_write_exr": (C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32, C.POINTER(ExrAttributes), C.c_int32]),
    "yar_guide_param_count": (C.c_int64, []),
    "yar_guide_create": (C.c_int, [C.c_int32, C.POINTER(GuideConfig), C.POINTER(_P)]),
    "yar_guide_destroy": (None, [_P]),
    "yar_guide_last_error": (C.c_char_p, [_P]),
    "yar_guide_get_params": (C.c_int, [_P, _P]),
    "yar_guide_set_params": (C.c_int, [_P, _P]),
    "yar_guide_forward": (C.c_int, [_P, _P, C.c_int64, _P]),
    "yar_guide_gradients": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P]),
    "yar_guide_train": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int32, _P]),
    "yar_guide_pdf": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "yar_guide_sample": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P]),
}

"""yar_amd - MI355X path-tracing core for YAR's reference renderer (RAY).

Thin Python mirror of the C ABI in include/yar_gpu.h.  The library (libyar_gpu.so, hand-written
HIP kernels for gfx950 plus the host pbrt3 loader and kd-tree builder) must be built in-tree
first (``python -m yar_amd.build``); nothing here falls back to a CPU implementation.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (FilmPixel, GpuInfo, GpuOptions, GuideConfig, Hit, Ray, RenderStats, SceneDesc, View,
                   film_dtype, hit_dtype, ray_dtype)

__all__ = ["lib", "YarError", "Scene", "Context", "Guide", "GuideConfig", "Checkpoint", "checkpoint_write_tile", "write_exr",
           "generate_scene", "build_kdtree", "build_box_kdtree",
           "GpuOptions", "View", "RenderStats", "ray_dtype", "hit_dtype", "film_dtype"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libyar_gpu.so")
_lib = None


class YarError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[yar status {code}] {msg}")
        self.code = code


def lib():
    """Loads libyar_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m yar_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _abi.SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, ctx=None):
    if rc != 0:
        L = lib()
        msg = L.yar_gpu_last_error(ctx) if ctx is not None else L.yar_last_error()
        raise YarError(rc, (msg or b"").decode(errors="replace"))


def generate_scene(config, path, seed=1, scale=1.0):
    """Writes the seeded benchmark scene `config` (1..5) as a pbrt3 file."""
    _check(lib().yar_scene_generate(int(config), int(seed), float(scale), os.fsencode(path)))
    return path


class Scene:
    """A loaded pbrt3 scene (Scene / load_scene, src/lib/scene_loader.cpp:74)."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib().yar_scene_load_pbrt(os.fsencode(path), C.byref(h)))
        self._h = h
        self.path = path

    @property
    def desc(self):
        d = lib().yar_scene_get_desc(self._h).contents
        d._owner = self            # keeps the scene (and the arrays the descriptor points at) alive
        return d

    def desc_ptr(self):
        return lib().yar_scene_get_desc(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().yar_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def triangle_count(self):
        d = self.desc
        return sum(d.meshes[i].triangle_count for i in range(d.mesh_count))


class KdTree:
    def __init__(self, handle):
        self._h = handle

    @property
    def words(self):
        L = lib()
        n = L.yar_kdtree_node_count(self._h)
        return np.ctypeslib.as_array(L.yar_kdtree_nodes(self._h), shape=(2 * n,)).copy()

    @property
    def bounds(self):
        b = (C.c_float * 6)()
        lib().yar_kdtree_bounds(self._h, b)
        return np.array(b[:], dtype=np.float32)

    @property
    def depth(self):
        return lib().yar_kdtree_max_depth(self._h)

    def save(self, path):
        _check(lib().yar_kdtree_save(self._h, os.fsencode(path)))

    def __del__(self):
        try:
            lib().yar_kdtree_free(self._h)
        except Exception:
            pass


def build_kdtree(vertices, indices, binned=False):
    """Mesh kd-tree: the reference builder's exact tree, or the binned SAH fast path."""
    v = np.ascontiguousarray(vertices, dtype=np.float32)
    i = np.ascontiguousarray(indices, dtype=np.int32)
    h = C.c_void_p()
    f = lib().yar_kdtree_build_mesh_binned if binned else lib().yar_kdtree_build_mesh
    _check(f(v.ctypes.data, i.ctypes.data, i.size // 3, C.byref(h)))
    return KdTree(h)


def build_box_kdtree(boxes):
    b = np.ascontiguousarray(boxes, dtype=np.float32)
    h = C.c_void_p()
    _check(lib().yar_kdtree_build_boxes(b.ctypes.data, b.size // 6, C.byref(h)))
    return KdTree(h)


class Context:
    """GPU context on one device (init_scene_context + render_scene, reference_renderer.h:48-55)."""

    def __init__(self, scene, options=None, **kw):
        opt = options or GpuOptions()
        for k, v in kw.items():
            setattr(opt, k, v)
        h = C.c_void_p()
        desc = scene.desc_ptr() if isinstance(scene, Scene) else C.pointer(scene)
        _check(lib().yar_gpu_create(desc, C.byref(opt), C.byref(h)))
        self._h = h
        self.scene = scene

    def close(self):
        if getattr(self, "_h", None):
            lib().yar_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        inf = GpuInfo()
        _check(lib().yar_gpu_get_info(self._h, C.byref(inf)), self._h)
        return inf

    def view(self, index=0):
        v = View()
        _check(lib().yar_gpu_get_view(self._h, index, C.byref(v)), self._h)
        return v

    def set_views(self, views):
        arr = (View * len(views))(*views)
        _check(lib().yar_gpu_set_views(self._h, arr, len(views)), self._h)

    def render(self, view=0, want_film=True):
        """Returns (image HxWx3 float32, film structured array or None, variance, stats)."""
        v = self.view(view)
        w, hgt = v.region_x1 - v.region_x0, v.region_y1 - v.region_y0
        img = np.empty((hgt, w, 3), dtype=np.float32)
        film = np.empty((hgt, w), dtype=film_dtype()) if want_film else None
        var = C.c_double()
        st = RenderStats()
        _check(lib().yar_gpu_render(self._h, view, film.ctypes.data if want_film else None, img.ctypes.data,
                                    C.byref(var), C.byref(st)), self._h)
        return img, film, var.value, st

    def tile_count(self, view=0):
        n = C.c_int32()
        _check(lib().yar_film_tile_count(self._h, view, C.byref(n)), self._h)
        return n.value

    def tile_bounds(self, tile, view=0):
        sb = (C.c_int32 * 4)()
        pb = (C.c_int32 * 4)()
        _check(lib().yar_film_tile_bounds(self._h, view, tile, sb, pb), self._h)
        return tuple(sb), tuple(pb)

    def render_tiles(self, tiles, view=0):
        """Film_Tile buffers (pixel-bounds sized structured arrays) and tile variance sums."""
        bufs = []
        for t in tiles:
            _, pb = self.tile_bounds(t, view)
            bufs.append(np.zeros((pb[3] - pb[1], pb[2] - pb[0]), dtype=film_dtype()))
        ptrs = (C.c_void_p * len(tiles))(*[b.ctypes.data for b in bufs])
        idx = (C.c_int32 * len(tiles))(*tiles)
        var = (C.c_double * len(tiles))()
        st = RenderStats()
        _check(lib().yar_gpu_render_tiles(self._h, view, idx, len(tiles), ptrs, var, C.byref(st)), self._h)
        return bufs, list(var), st

    def render_checkpointed(self, checkpoint_dir, input_filename, view=0, tiles_per_batch=0, want_film=True):
        """render_scene with a checkpoint directory: (image, film or None, variance, render time, stats)."""
        v = self.view(view)
        w, hgt = v.region_x1 - v.region_x0, v.region_y1 - v.region_y0
        img = np.empty((hgt, w, 3), dtype=np.float32)
        film = np.empty((hgt, w), dtype=film_dtype()) if want_film else None
        var = C.c_double()
        t = C.c_float()
        st = RenderStats()
        _check(lib().yar_gpu_render_checkpointed(self._h, view, os.fsencode(checkpoint_dir), os.fsencode(input_filename),
                                                 tiles_per_batch, film.ctypes.data if want_film else None, img.ctypes.data,
                                                 C.byref(var), C.byref(t), C.byref(st)), self._h)
        return img, film, var.value, t.value, st

    def render_batch(self, views):
        imgs = []
        for vi in views:
            v = self.view(vi)
            imgs.append(np.empty((v.region_y1 - v.region_y0, v.region_x1 - v.region_x0, 3), dtype=np.float32))
        ptrs = (C.c_void_p * len(views))(*[i.ctypes.data for i in imgs])
        idx = (C.c_int32 * len(views))(*views)
        var = (C.c_double * len(views))()
        st = RenderStats()
        _check(lib().yar_gpu_render_batch(self._h, idx, len(views), ptrs, var, C.byref(st)), self._h)
        return imgs, list(var), st

    def render_guided(self, guide, views):
        """Guided form of render_batch (the guide trains between sample passes)."""
        imgs = []
        for vi in views:
            v = self.view(vi)
            imgs.append(np.empty((v.region_y1 - v.region_y0, v.region_x1 - v.region_x0, 3), dtype=np.float32))
        ptrs = (C.c_void_p * len(views))(*[i.ctypes.data for i in imgs])
        idx = (C.c_int32 * len(views))(*views)
        var = (C.c_double * len(views))()
        st = RenderStats()
        _check(lib().yar_gpu_render_guided(self._h, guide._h, idx, len(views), ptrs, var, C.byref(st)), self._h)
        return imgs, list(var), st

    def trace_rays(self, rays, any_hit=False):
        """rays: structured array of ray_dtype(); returns structured array of hit_dtype()."""
        r = np.ascontiguousarray(rays, dtype=ray_dtype())
        out = np.empty(r.shape, dtype=hit_dtype())
        _check(lib().yar_gpu_trace_rays(self._h, r.ctypes.data, r.size, 1 if any_hit else 0, out.ctypes.data), self._h)
        return out

    def trace_rays_device(self, d_rays_ptr, count, d_hits_ptr, any_hit=False, stream=None):
        _check(lib().yar_gpu_trace_rays_device(self._h, C.c_void_p(d_rays_ptr), count, 1 if any_hit else 0,
                                               C.c_void_p(d_hits_ptr), C.c_void_p(stream) if stream else None),
               self._h)

    def trace_rays_device_noalloc(self, d_rays_ptr, count, d_hits_ptr, any_hit=False, stream=None):
        """Allocation-free device query (at most 2^20 rays; calls overlapping on one device race)."""
        _check(lib().yar_gpu_trace_rays_device_noalloc(self._h, C.c_void_p(d_rays_ptr), count, 1 if any_hit else 0,
                                                       C.c_void_p(d_hits_ptr), C.c_void_p(stream) if stream else None),
               self._h)

    def last_trace_counters(self):
        st = RenderStats()
        _check(lib().yar_gpu_last_trace_counters(self._h, C.byref(st)), self._h)
        return st


class Checkpoint:
    """A checkpoint directory in the reference's format (start_or_resume_checkpoint,
    src/ref/reference_renderer.cpp:179)."""

    def __init__(self, directory, input_filename, total_tile_count, samples_per_pixel):
        h = C.c_void_p()
        _check(lib().yar_checkpoint_open(os.fsencode(directory), os.fsencode(input_filename), total_tile_count,
                                         samples_per_pixel, C.byref(h)))
        self._h = h

    def finished_tiles(self):
        n = lib().yar_checkpoint_finished_count(self._h)
        out = (C.c_int32 * max(1, n))()
        _check(lib().yar_checkpoint_finished_tiles(self._h, out))
        return list(out[:n])

    def tile(self, index):
        """(pixel bounds x0, y0, x1, y1; Film_Pixels as a structured array; tile variance)."""
        b = (C.c_int32 * 4)()
        _check(lib().yar_checkpoint_tile(self._h, index, b, None, None))
        px = np.zeros((b[3] - b[1], b[2] - b[0]), dtype=film_dtype())
        var = C.c_double()
        _check(lib().yar_checkpoint_tile(self._h, index, b, px.ctypes.data, C.byref(var)))
        return tuple(b), px, var.value

    @property
    def previous_time(self):
        return lib().yar_checkpoint_previous_time(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().yar_checkpoint_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def checkpoint_write_tile(directory, tile_index, pixel_bounds, pixels, variance, render_time):
    """write_tile_to_checkpoint_directory (src/ref/reference_renderer.cpp:283)."""
    b = (C.c_int32 * 4)(*pixel_bounds)
    px = np.ascontiguousarray(pixels, dtype=film_dtype())
    _check(lib().yar_checkpoint_write_tile(os.fsencode(directory), tile_index, b, px.ctypes.data, float(variance),
                                           float(render_time)))


def write_exr(path, image, input_file="", spp=0, variance=0.0, load_time=0.0, render_time=0.0,
              varying_attributes=False, dump_attributes=False, compress=False):
    """write_openexr_image (src/ref/reference_renderer.cpp:673): HALF B, G, R with the yar_* attributes."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    a = _abi.ExrAttributes(os.fsencode(input_file), int(spp), float(variance), float(load_time), float(render_time),
                           1 if varying_attributes else 0, 1 if dump_attributes else 0)
    _check(lib().yar_image_write_exr(os.fsencode(path), img.ctypes.data, img.shape[1], img.shape[0], C.byref(a),
                                     1 if compress else 0))
    return path


class Guide:
    """Learned directional proposal (yar_guide_*): the MLP + vMF mixture with its own training."""

    def __init__(self, device=0, config=None, **kw):
        cfg = config or _abi.GuideConfig()
        for k, v in kw.items():
            setattr(cfg, k, v)
        h = C.c_void_p()
        rc = lib().yar_guide_create(int(device), C.byref(cfg), C.byref(h))
        if rc != 0:
            raise YarError(rc, (lib().yar_last_error() or b"").decode(errors="replace"))
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise YarError(rc, (lib().yar_guide_last_error(self._h) or b"").decode(errors="replace"))

    def close(self):
        if getattr(self, "_h", None):
            lib().yar_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count():
        return int(lib().yar_guide_param_count())

    def params(self):
        out = np.empty(self.param_count(), dtype=np.float32)
        self._check(lib().yar_guide_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.param_count()
        self._check(lib().yar_guide_set_params(self._h, p.ctypes.data))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        out = np.empty((x.shape[0], 40), dtype=np.float32)
        self._check(lib().yar_guide_forward(self._h, x.ctypes.data, x.shape[0], out.ctypes.data))
        return out

    def gradients(self, x, dirs, weights):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        g = np.empty(self.param_count(), dtype=np.float32)
        loss = C.c_float()
        self._check(lib().yar_guide_gradients(self._h, x.ctypes.data, d.ctypes.data, w.ctypes.data, x.shape[0],
                                              g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def train(self, x, dirs, weights, steps):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        loss = C.c_float()
        self._check(lib().yar_guide_train(self._h, x.ctypes.data, d.ctypes.data, w.ctypes.data, x.shape[0],
                                          int(steps), C.byref(loss)))
        return loss.value

    def pdf(self, x, dirs):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        out = np.empty(x.shape[0], dtype=np.float32)
        self._check(lib().yar_guide_pdf(self._h, x.ctypes.data, d.ctypes.data, x.shape[0], out.ctypes.data))
        return out

    def sample(self, x, u):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 3)
        dirs = np.empty((x.shape[0], 3), dtype=np.float32)
        pdf = np.empty(x.shape[0], dtype=np.float32)
        self._check(lib().yar_guide_sample(self._h, x.ctypes.data, u.ctypes.data, x.shape[0], dirs.ctypes.data,
                                           pdf.ctypes.data))
        return dirs, pdf

"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) - test infrastructure only."""
import ctypes as C
import os

import numpy as np

from yar_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
PCG_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpcg_ref.so")
THREADS = int(os.environ.get("ORACLE_THREADS", str(len(os.sched_getaffinity(0)))))

_o = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            from oracle import build
            build.build()
        o = C.CDLL(ORACLE_SO)
        P = C.c_void_p
        o.oref_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int32, P]
        o.oref_pixel_samples.argtypes = [C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         P, P, P, P]
        o.oref_watertight.argtypes = [P, P, P]
        o.oref_watertight.restype = C.c_float
        for n in ("oref_kdtree_build_mesh",):
            getattr(o, n).argtypes = [P, P, C.c_int32, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32), P,
                                      C.POINTER(C.c_int32)]
        o.oref_kdtree_build_boxes.argtypes = [P, C.c_int32, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32), P,
                                              C.POINTER(C.c_int32)]
        o.oref_free.argtypes = [P]
        o.oref_scene_create.argtypes = [C.POINTER(_abi.SceneDesc), C.c_int32]
        o.oref_scene_create.restype = P
        o.oref_scene_free.argtypes = [P]
        o.oref_scene_tree.argtypes = [P, C.c_int32, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32), P]
        o.oref_trace.argtypes = [P, P, C.c_int64, C.c_int32, C.c_int32, P, P, C.c_int32]
        o.oref_render.argtypes = [P, C.POINTER(_abi.View), P, P, C.POINTER(C.c_double), P, P, C.c_int32]
        _o = o
    return _o


def pcg32(initstate, initseq, n):
    out = np.empty(n, dtype=np.uint32)
    lib().oref_pcg32(initstate, initseq, n, out.ctypes.data)
    return out


def pixel_samples(x, y, seed_offset, nx, ny, d1, d2):
    spp = nx * ny
    img = np.empty(2 * spp, np.float32)
    s1 = np.empty(max(1, d1 * spp), np.float32)
    s2 = np.empty(max(1, 2 * d2 * spp), np.float32)
    st = C.c_uint64()
    lib().oref_pixel_samples(x, y, seed_offset, nx, ny, d1, d2, img.ctypes.data, s1.ctypes.data, s2.ctypes.data,
                             C.byref(st))
    return img.reshape(spp, 2), s1[:d1 * spp].reshape(d1, spp), s2[:2 * d2 * spp].reshape(d2, spp, 2), st.value


def watertight(origin, direction, p0, p1, p2):
    ray = np.array(list(origin) + list(direction), np.float32)
    tri = np.array(list(p0) + list(p1) + list(p2), np.float32)
    b = np.zeros(3, np.float32)
    t = lib().oref_watertight(ray.ctypes.data, tri.ctypes.data, b.ctypes.data)
    return np.float32(t), b


def _take_words(ptr, n):
    arr = np.ctypeslib.as_array(ptr, shape=(2 * n,)).copy()
    lib().oref_free(ptr)
    return arr


def kdtree_mesh(vertices, indices):
    v = np.ascontiguousarray(vertices, np.float32)
    i = np.ascontiguousarray(indices, np.int32)
    w = C.POINTER(C.c_uint32)()
    n = C.c_uint32()
    b = np.zeros(6, np.float32)
    d = C.c_int32()
    lib().oref_kdtree_build_mesh(v.ctypes.data, i.ctypes.data, i.size // 3, C.byref(w), C.byref(n), b.ctypes.data,
                                 C.byref(d))
    return _take_words(w, n.value), b, d.value


def kdtree_boxes(boxes):
    bx = np.ascontiguousarray(boxes, np.float32)
    w = C.POINTER(C.c_uint32)()
    n = C.c_uint32()
    b = np.zeros(6, np.float32)
    d = C.c_int32()
    lib().oref_kdtree_build_boxes(bx.ctypes.data, bx.size // 6, C.byref(w), C.byref(n), b.ctypes.data, C.byref(d))
    return _take_words(w, n.value), b, d.value


class OracleScene:
    def __init__(self, scene, threads=THREADS):
        self.threads = threads
        self._h = lib().oref_scene_create(scene.desc_ptr(), threads)
        self._scene = scene

    def __del__(self):
        try:
            lib().oref_scene_free(self._h)
        except Exception:
            pass

    def tree(self, mesh=-1):
        w = C.POINTER(C.c_uint32)()
        n = C.c_uint32()
        b = np.zeros(6, np.float32)
        lib().oref_scene_tree(self._h, mesh, C.byref(w), C.byref(n), b.ctypes.data)
        return _take_words(w, n.value), b

    def trace(self, rays, any_hit=False, brute=False, counters=False):
        r = np.ascontiguousarray(rays, dtype=_abi.ray_dtype())
        out = np.empty(r.shape, dtype=_abi.hit_dtype())
        cnt = np.zeros((r.size, 4), np.int64) if counters else None
        lib().oref_trace(self._h, r.ctypes.data, r.size, 1 if any_hit else 0, 1 if brute else 0, out.ctypes.data,
                         cnt.ctypes.data if counters else None, self.threads)
        return (out, cnt) if counters else out

    def render(self, view, per_sample=False, stats=False):
        w, h = view.region_x1 - view.region_x0, view.region_y1 - view.region_y0
        spp = view.x_pixel_sample_count * view.y_pixel_sample_count
        film = np.zeros((h, w), dtype=_abi.film_dtype())
        img = np.zeros((h, w, 3), np.float32)
        ps = np.zeros((h, w, spp, 3), np.float32) if per_sample else None
        st = np.zeros(6, np.int64)
        var = C.c_double()
        lib().oref_render(self._h, C.byref(view), film.ctypes.data, img.ctypes.data, C.byref(var),
                          ps.ctypes.data if per_sample else None, st.ctypes.data if stats else None, self.threads)
        return img, film, var.value, ps, st


def pcg_ref_available():
    return os.path.exists(PCG_REF_SO)


def pcg32_ref(initstate, initseq, n):
    """The reference's own pcg_basic.c (built unchanged into oracle/_ref)."""

    class State(C.Structure):
        _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]

    r = C.CDLL(PCG_REF_SO)
    r.pcg32_srandom_r.argtypes = [C.POINTER(State), C.c_uint64, C.c_uint64]
    r.pcg32_random_r.argtypes = [C.POINTER(State)]
    r.pcg32_random_r.restype = C.c_uint32
    s = State()
    r.pcg32_srandom_r(C.byref(s), initstate, initseq)
    return np.array([r.pcg32_random_r(C.byref(s)) for _ in range(n)], dtype=np.uint32)

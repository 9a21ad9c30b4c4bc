"""Allocation-free ray queries (yar_gpu_trace_rays_device_noalloc) against the allocating device entry
point yar_gpu_trace_rays_device: same kernels, scratch in a static device array instead of context
buffers.  Hit records must be bit-identical, on two calls in a row with the second one smaller."""
import ctypes as C

import numpy as np
import pytest

import yar_amd
from tests import raygen

pytestmark = pytest.mark.gpu


def _noalloc(ctx, d_rays, n, d_hits, any_hit, stream):
    return yar_amd.lib().yar_gpu_trace_rays_device_noalloc(ctx._h, C.c_void_p(d_rays), n, 1 if any_hit else 0,
                                                           C.c_void_p(d_hits), C.c_void_p(stream))


class _Hip:
    """Device buffers through the HIP runtime the library itself is linked against (libamdhip64.so.7,
    already loaded with libyar_gpu.so); a second runtime in the process (torch's bundled one) cannot
    share the device with it."""

    def __init__(self):
        yar_amd.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.bufs = []

    def alloc(self, nbytes, fill):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        assert self.h.hipMemset(p, fill, nbytes) == 0
        assert self.h.hipDeviceSynchronize() == 0          # the context's stream does not wait on the null stream
        self.bufs.append(p)
        return p.value

    def upload(self, arr):
        p = self.alloc(arr.nbytes, 0)
        assert self.h.hipMemcpy(C.c_void_p(p), arr.ctypes.data, arr.nbytes, 1) == 0     # host to device
        return p

    def download(self, p, nbytes):
        out = np.empty(nbytes, np.uint8)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, C.c_void_p(p), nbytes, 2) == 0          # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.mark.parametrize("any_hit", [False, True], ids=["closest", "any"])
def test_noalloc_matches_device_entry_point(any_hit, c2_small_path):
    scene = yar_amd.Scene(c2_small_path)
    ctx = yar_amd.Context(scene)
    hip = _Hip()
    rays = raygen.random_rays(scene, 300_000, seed=21)
    if any_hit:
        rays["tmax"] = np.random.default_rng(4).uniform(0.01, 5.0, rays.size).astype(np.float32)
    try:
        for n in (300_000, 70_001):                      # the second call is smaller than the first
            print(f"noalloc {'any' if any_hit else 'closest'}: {n} rays", flush=True)
            d_rays = hip.upload(np.ascontiguousarray(rays[:n]))
            ref = hip.alloc(n * 24, 0xAB)
            got = hip.alloc(n * 24, 0xCD)
            ctx.trace_rays_device(d_rays, n, ref, any_hit=any_hit)
            rc = _noalloc(ctx, d_rays, n, got, any_hit, None)
            assert rc == 0, yar_amd.lib().yar_gpu_last_error(ctx._h)
            a, b = hip.download(ref, n * 24), hip.download(got, n * 24)
            assert np.array_equal(a, b)
            hits = b.view(yar_amd.hit_dtype())
            host = ctx.trace_rays(rays[:n], any_hit=any_hit)
            assert np.array_equal(b, host.view(np.uint8))
            if not any_hit:
                assert (hits["object"] >= 0).mean() > 0.05
            hip.free()
    finally:
        hip.free()


def test_noalloc_rejects_batches_beyond_the_scratch(c1_path):
    scene = yar_amd.Scene(c1_path)
    ctx = yar_amd.Context(scene)
    rc = _noalloc(ctx, 16, (1 << 20) + 1, 16, False, None)      # rejected before any device access
    assert rc == 7                                               # YAR_ERR_CAPACITY
    assert b"2^20" in yar_amd.lib().yar_gpu_last_error(ctx._h)

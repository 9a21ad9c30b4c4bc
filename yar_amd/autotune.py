"""Autotuner for the wavefront tracer's launch parameters.

Renders a crop of a view with each candidate configuration (traversal block size, stack storage and
depth, paths in flight) and keeps the fastest by the HIP-event time of the bounce loop.  Results
are identical across candidates (the traversal order does not depend on them), so only time is
compared.
"""
import itertools
import json
import time

from . import Context, GpuOptions


def candidates(full=False, paths=None):
    """Traversal block size x stack storage/depth x lane refill, optionally x paths in flight."""
    blocks = [64, 128, 256] if full else [64, 128]
    out = []
    for b in blocks:
        for depth in ([4, 8, 16, 32] if full else [4, 8]):
            out.append(dict(block_size=b, stack_mode=1, traversal_stack_depth=depth))
        out.append(dict(block_size=b, stack_mode=2, traversal_stack_depth=16))
    out = [dict(c, lane_refill=r) for c in out for r in (1, 2)]
    if paths:
        out = [dict(c, max_paths_in_flight=p) for c in out for p in paths]
    return out


def _crop_view(ctx, view_index, crop):
    v = ctx.view(view_index)
    if crop:
        cx = (v.region_x0 + v.region_x1) // 2
        cy = (v.region_y0 + v.region_y1) // 2
        half = crop // 2
        v.region_x0, v.region_x1 = max(v.region_x0, cx - half), min(v.region_x1, cx - half + crop)
        v.region_y0, v.region_y1 = max(v.region_y0, cy - half), min(v.region_y1, cy - half + crop)
    return v


def tune(scene, view_index=0, crop=256, repeats=2, full=False, extra=None, log=None, paths=None, views=None):
    """Returns (best options dict, list of (options, seconds)) for `scene`.  With `views` the
    candidates are timed on a batched render of those views instead of a crop of one view."""
    results = []
    for cand in candidates(full, paths) + list(extra or []):
        opts = GpuOptions()
        for k, val in cand.items():
            setattr(opts, k, val)
        opts.enable_timing = 0
        ctx = Context(scene, opts)
        if views:
            ctx.set_views(views)
            idx = list(range(len(views)))
            run = lambda: ctx.render_batch(idx)[2]
        else:
            ctx.set_views([_crop_view(ctx, view_index, crop)])
            run = lambda: ctx.render(0, want_film=False)[3]
        run()                                                # warm up (allocations, code load)
        best = float("inf")
        for _ in range(repeats):
            best = min(best, run().gpu_seconds)
        ctx.close()
        results.append((cand, best))
        if log:
            log(f"autotune {json.dumps(cand)}: {best * 1e3:.2f} ms")
    best_cand = min(results, key=lambda r: r[1])[0]
    return best_cand, results


def options_from(cand, **overrides):
    opts = GpuOptions()
    for k, val in dict(cand, **overrides).items():
        setattr(opts, k, val)
    return opts

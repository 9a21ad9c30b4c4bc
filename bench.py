#!/usr/bin/env python3
"""Benchmark of the plain wavefront GPU path tracer on the config-2 workload.

A step is one full render of the seeded config-2 scene (400 displaced icospheres + a 2 km ground
grid, diffuse / GGX conductor / smooth glass, 2 rectangular lights + 1 point light) at 1024x1024,
64 spp (8x8 stratified), max depth 8, through the C ABI (yar_gpu_render).  The scene and its
kd-trees are resident in HBM before the timed region; each step writes the merged film back to the
host.  Prints one JSON line: Msamples/s (value), rays/s, the traversal kernel's roofline and the CPU
oracle timed on the host cores on a bounded crop of the same workload.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s & rays/s at 1 GPU; equal-time MSE, guided vs plain wavefront"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per traversal event with this build's HBM layout: kd node 8 B, leaf index 4 B,
# triangle 48 B (3 x float4 vertices), object record 80 B (world_to_object + mesh bounds).
BYTES_NODE, BYTES_LEAF_REF, BYTES_TRI, BYTES_OBJ = 8, 4, 48, 80


def _scene(config, scale, seed):
    import yar_amd
    cache = os.environ.get("YAR_SCENE_CACHE", os.path.join(tempfile.gettempdir(), "yar_scenes"))
    os.makedirs(cache, exist_ok=True)
    path = os.path.join(cache, f"config{config}_s{seed}_x{scale}.pbrt")
    if not os.path.exists(path):
        yar_amd.generate_scene(config, path + ".tmp", seed=seed, scale=scale)
        os.replace(path + ".tmp", path)
    return yar_amd.Scene(path)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads():
    """Every host core the process may use (the CPU baseline runs the oracle on all of them)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_baseline(scene, view, crop, threads):
    from tests import oracle_lib
    import yar_amd
    v = yar_amd.View()
    C = type(view)
    for name, _ in C._fields_:
        setattr(v, name, getattr(view, name))
    cx = (view.region_x0 + view.region_x1) // 2
    cy = (view.region_y0 + view.region_y1) // 2
    v.region_x0, v.region_y0 = cx - crop // 2, cy - crop // 2
    v.region_x1, v.region_y1 = v.region_x0 + crop, v.region_y0 + crop
    osc = oracle_lib.OracleScene(scene, threads=threads)
    t0 = time.perf_counter()
    osc.render(v)
    dt = time.perf_counter() - t0
    spp = v.x_pixel_sample_count * v.y_pixel_sample_count
    samples = crop * crop * spp
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{crop}x{crop} centre crop of the same render ({samples} samples, {spp} spp), "
                      f"{dt:.2f} s on {threads} threads = all host cores (oracle/cpu_ref.cpp, g++ -O2 -ffp-contract=off)"}


def _lum(img):
    import numpy as np
    return 0.2126729 * img[..., 0] + 0.7151522 * img[..., 1] + 0.0721750 * img[..., 2]


def _config1_leg(threads, cpu):
    """Config 1 (256x256, 16 spp, depth 5): GPU samples/s of the whole render beside the CPU oracle
    on the same render."""
    import yar_amd
    scene = _scene(1, 1.0, 1)
    ctx = yar_amd.Context(scene)
    ctx.render(0, want_film=False)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        _, _, _, st = ctx.render(0, want_film=False)
        ts.append(time.perf_counter() - t0)
    gpu = st.samples / min(ts) / 1e6
    out = {"workload": "config 1 plain render, 256x256, 16 spp, max depth 5", "gpu_msamples_s": round(gpu, 2),
           "gpu_ms": round(min(ts) * 1e3, 2)}
    if cpu:
        from tests import oracle_lib
        osc = oracle_lib.OracleScene(scene, threads=threads)
        t0 = time.perf_counter()
        osc.render(ctx.view(0))
        dt = time.perf_counter() - t0
        c = st.samples / dt / 1e6
        out.update({"cpu_msamples_s": round(c, 3), "cpu_threads": threads, "cpu_s": round(dt, 2),
                    "gpu_vs_cpu": round(gpu / c, 1)})
    ctx.close()
    return out


def _batch_leg():
    """Config 3: the config-1 scene from 32 seeded poses (sides 128..1024, 4/16/64 spp) in one
    batched render with the bounce loop replayed from a hipGraph."""
    import yar_amd
    from yar_amd import views as yv
    from yar_amd import autotune
    scene = _scene(1, 1.0, 1)
    probe = yar_amd.Context(scene)
    views = yv.batch_views(probe.view(0), 32)
    probe.close()
    # launch shape from the autotuner, timed on a reduced batch of the same poses; paths in flight
    # keep the library default (a reduced batch is too small to judge the chunk size)
    small = [yv.with_film(v, 256, 256, v.x_pixel_sample_count, seed_offset=v.rng_seed_offset) for v in views[:6]]
    best, _ = autotune.tune(scene, views=small, repeats=2)
    ctx = yar_amd.Context(scene, use_graph=1, **best)
    ctx.set_views(views)
    idx = list(range(len(views)))
    ctx.render_batch(idx)
    t0 = time.perf_counter()
    _, _, st = ctx.render_batch(idx)
    dt = time.perf_counter() - t0
    ctx.close()
    return {"workload": "config 3: config-1 scene, 32 views, sides {128..1024}, spp {4,16,64}, hipGraph bounce loop",
            "autotuned": best,
            "samples": st.samples, "ms": round(dt * 1e3, 1), "msamples_s": round(st.samples / dt / 1e6, 2),
            "mrays_s": round((st.closest_rays + st.any_rays) / dt / 1e6, 1)}


def _mse_legs(ctx, plain_views, ref_views, runs=2, guide_seed=1):
    """Renders the plain views, their references and the guided views; MSE against the references and
    the guided/plain time ratio to reach the same MSE (MSE ~ 1/time for both estimators)."""
    import numpy as np
    import yar_amd
    n = len(plain_views)
    ctx.set_views(plain_views + ref_views)
    ref, _, _ = ctx.render_batch(list(range(n, 2 * n)))
    idx = list(range(n))
    ctx.render_batch(idx)
    tp = []
    for _ in range(runs):
        t0 = time.perf_counter()
        plain, _, _ = ctx.render_batch(idx)
        tp.append(time.perf_counter() - t0)
    tg = []
    for _ in range(runs):
        guide = yar_amd.Guide(seed=guide_seed)
        t0 = time.perf_counter()
        guided, _, gst = ctx.render_guided(guide, idx)
        tg.append(time.perf_counter() - t0)
    tp, tg = min(tp), min(tg)
    mp = float(np.mean([((a - b) ** 2).mean() for a, b in zip(plain, ref)]))
    mg = float(np.mean([((a - b) ** 2).mean() for a, b in zip(guided, ref)]))
    return {"plain_ms": round(tp * 1e3, 1), "guided_ms": round(tg * 1e3, 1), "mse_plain": mp, "mse_guided": mg,
            "equal_spp_mse_ratio": round(mg / mp, 4), "time_ratio_to_equal_mse": round((mg * tg) / (mp * tp), 3),
            "train_nats": round(gst.train_nats, 3), "passes": gst.passes, "train_records": gst.train_records}


def _guided_c3_leg(side=128, grid=8, ref_grid=64):
    """Config 3 target: the 32 seeded poses of the batched workload (films reduced to side^2), guided
    vs plain at grid^2 spp against a ref_grid^2-spp plain reference."""
    import yar_amd
    from yar_amd import views as yv
    scene = _scene(1, 1.0, 1)
    ctx = yar_amd.Context(scene)
    base = yv.batch_views(ctx.view(0), 32)
    vs = [yv.with_film(v, side, side, grid, seed_offset=v.rng_seed_offset) for v in base]
    refs = [yv.with_film(v, side, side, ref_grid, seed_offset=7000 + i) for i, v in enumerate(base)]
    out = {"workload": f"config 3: the 32 batch poses at {side}x{side}, {grid * grid} spp, reference {ref_grid ** 2} spp plain"}
    out.update(_mse_legs(ctx, vs, refs))
    ctx.close()
    return out


def _guided_c4_leg(side=256, grid=8, ref_grid=64):
    """Config 4 (light through a slit): a side^2 centre crop, guided vs plain at grid^2 spp against a
    ref_grid^2-spp plain reference."""
    import yar_amd
    from yar_amd import views as yv
    scene = _scene(4, 1.0, 1)
    ctx = yar_amd.Context(scene)
    v0 = ctx.view(0)
    cx, cy = (v0.region_x0 + v0.region_x1) // 2, (v0.region_y0 + v0.region_y1) // 2
    region = (cx - side // 2, cy - side // 2, cx + side // 2, cy + side // 2)
    v = yv.with_film(v0, v0.film_width, v0.film_height, grid, seed_offset=0, region=region)
    r = yv.with_film(v0, v0.film_width, v0.film_height, ref_grid, seed_offset=9000, region=region)
    out = {"workload": f"config 4: {side}x{side} centre crop of the 1024x1024 film, {grid * grid} spp, "
                       f"reference {ref_grid ** 2} spp plain"}
    out.update(_mse_legs(ctx, [v], [r]))
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-crop", type=int, default=128, help="side of the CPU-baseline crop (0 disables)")
    ap.add_argument("--no-counters", action="store_true", help="skip the counted render (roofline bytes)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0)
    ap.add_argument("--opt", action="append", default=[], metavar="FIELD=INT",
                    help="set a yar_gpu_options field for the config-2 render (e.g. lane_refill=1)")
    ap.add_argument("--no-extra", action="store_true", help="skip the config-1, batch and guided legs")
    ap.add_argument("--leg", choices=["config1", "batch", "guided_c3", "guided_c4"], help="run only this leg (profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with a torch copy kernel (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the path shards by "
              f"film tiles or views with no collectives (one context per GPU, Film_Tiles merged in tile order)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return
    if args.leg:
        threads = _host_threads()
        fn = {"config1": lambda: _config1_leg(threads, False), "batch": _batch_leg, "guided_c3": _guided_c3_leg,
              "guided_c4": _guided_c4_leg}[args.leg]
        print(json.dumps(fn()), flush=True)
        return

    import yar_amd
    t0 = time.perf_counter()
    scene = _scene(args.config, args.scale, args.seed)
    t_load = time.perf_counter() - t0
    ntri = scene.triangle_count()
    opts = dict(enable_timing=1)
    if args.block:
        opts["block_size"] = args.block
    if args.stack:
        opts["traversal_stack_depth"] = args.stack
    for kv in args.opt:
        k, v = kv.split("=", 1)
        opts[k] = int(v)
    ctx = yar_amd.Context(scene, **opts)
    info = ctx.info()
    view = ctx.view(0)
    spp = view.x_pixel_sample_count * view.y_pixel_sample_count
    w = view.region_x1 - view.region_x0
    h = view.region_y1 - view.region_y0
    print(f"# scene: {ntri} triangles, {info.object_count} objects, kd build {info.build_seconds:.2f} s, "
          f"load {t_load:.2f} s, tree depths scene {info.scene_tree_depth} mesh {info.mesh_tree_depth}, "
          f"stacks {info.stack_depth >> 16}/{info.stack_depth & 0xffff}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        img, _, _, st = ctx.render(0, want_film=False)
        print(f"# warmup {i}: {st.render_seconds * 1e3:.1f} ms", file=sys.stderr, flush=True)

    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        img, _, _, st = ctx.render(0, want_film=False)
        stats.append(st)
        print(f"# step {i}: {st.render_seconds * 1e3:.1f} ms (gpu {st.gpu_seconds * 1e3:.1f} ms)",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    samples = stats[0].samples
    rays = stats[0].closest_rays + stats[0].any_rays
    value = samples / (elapsed / args.steps) / 1e6
    rays_per_s = rays / (elapsed / args.steps)

    ext_sec = sum(s.extend_seconds for s in stats) / len(stats)
    ext_launches = stats[0].extend_launches
    roofline = None
    culled = None
    if not args.no_counters:
        cctx = yar_amd.Context(scene, **dict(opts, enable_timing=0, enable_counters=1))
        _, _, _, cs = cctx.render(0, want_film=False)
        cctx.close()
        culled = cs.mis_culled_rays
        ext_bytes = (BYTES_NODE * cs.extend_node_visits + BYTES_LEAF_REF * cs.extend_leaf_refs +
                     BYTES_TRI * cs.extend_triangle_tests + BYTES_OBJ * cs.extend_object_visits)
        achieved = ext_bytes / ext_sec / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "extend_traffic.json")
        if os.path.exists(tpath):
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        counters = None
        cpath = os.path.join(ROOT, "profiles", "extend_counters.json")
        if os.path.exists(cpath):
            counters = json.load(open(cpath))
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": "k_extend (closest-hit kd traversal of path rays)",
                    "limiter": "instruction issue: dependent kd-node loads under SIMD divergence (HBM traffic "
                               "and bandwidth are far below peak; see counters)",
                    "counters": counters,
                    "algorithmic_bytes_per_launch": int(ext_bytes / max(1, ext_launches)),
                    "avg_launch_ms": round(ext_sec / max(1, ext_launches) * 1e3, 4),
                    "launches_per_step": ext_launches,
                    "per_ray": {"nodes": round(cs.extend_node_visits / cs.extend_rays, 2),
                                "leaf_refs": round(cs.extend_leaf_refs / cs.extend_rays, 2),
                                "triangle_tests": round(cs.extend_triangle_tests / cs.extend_rays, 2),
                                "objects": round(cs.extend_object_visits / cs.extend_rays, 2)}}

    cpu = None
    threads = _host_threads()
    if args.cpu_crop > 0:
        cpu = _cpu_baseline(scene, view, args.cpu_crop, threads)
    extra = {}
    if not args.no_extra:
        extra["config1"] = _config1_leg(threads, args.cpu_crop > 0)
        print(f"# config 1: {extra['config1']}", file=sys.stderr, flush=True)
        extra["batch"] = _batch_leg()
        print(f"# batch: {extra['batch']}", file=sys.stderr, flush=True)
        extra["guided_c3"] = _guided_c3_leg()
        print(f"# guided config 3: {extra['guided_c3']}", file=sys.stderr, flush=True)
        extra["guided_c4"] = _guided_c4_leg()
        print(f"# guided config 4: {extra['guided_c4']}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded config-2 scene generator (yar_scene_generate), no checkpoints",
        "config": {"workload": f"config {args.config} plain wavefront render, {ntri} triangles, {w}x{h}, "
                               f"{spp} spp, max depth {ctx.scene.desc.config.max_light_bounces}",
                   "samples_per_step": samples, "triangles": ntri, "scale": args.scale, "seed": args.seed},
        "rays_per_s": round(rays_per_s, 1), "mrays_per_s": round(rays_per_s / 1e6, 2),
        # the reference counts every MIS BSDF ray toward a light as traced; this build answers some of
        # them with an analytic light-rectangle test (count from the counted render of the same step)
        "mrays_per_s_reference_definition": (round((rays + culled) / (elapsed / args.steps) / 1e6, 2)
                                             if culled is not None else None),
        "rays_per_sample": round(rays / samples, 3),
        "gpu_ms_per_step": round(sum(s.gpu_seconds for s in stats) / len(stats) * 1e3, 2),
        "kernel_ms_per_step": {"extend": round(ext_sec * 1e3, 2),
                               "shade": round(sum(s.shade_seconds for s in stats) / len(stats) * 1e3, 2),
                               "shadow": round(sum(s.shadow_seconds for s in stats) / len(stats) * 1e3, 2),
                               "mis": round(sum(s.mis_seconds for s in stats) / len(stats) * 1e3, 2),
                               "other": round(sum(s.other_seconds for s in stats) / len(stats) * 1e3, 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_vs_cpu"] = round(value / cpu["value"], 1)
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

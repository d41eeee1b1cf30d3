#!/usr/bin/env python3
"""Benchmark of the wavefront render path on one MI355X.

Workload (configuration 1 of the measurement plan): the sphere field -- a
direct-tested ground sphere of radius 1000 and 484 spheres of radius 0.2 on a
jittered 22x22 grid (80% Lambertian / 15% metal with fuzz U(0, 0.5) / 5%
dielectric 1.5) -- plus three TLAS instances of one BLAS, a displaced
geodesic icosphere of 1,003,520 triangles.  1920x1080, 64 spp, depth 12,
Russian roulette from depth 3, PCG32 streams, indirect clamp 10.

One step = one full frame (64 spp) rendered into the device framebuffer by
the hipGraph-captured staged wavefront; the scene is resident in HBM before
the timed region.  `value` is Mrays/s = (closest-hit + shadow queries) / s;
Msamples/s and the persistent-kernel rate are reported beside it.

The roofline object is for the dominant kernel (closest-hit trace), timed with
HIP events on the renderer's stream; its algorithmic bytes per query follow
the reference algorithm's work counts (node tests, triangle/sphere tests,
instance entries) measured by the CPU oracle on the same scene and BVH.
The CPU baseline is the oracle (the scalar FP32 restatement of the reference
traversal and shading) on the host cores, over a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wavefront-path-tracer_amd", "python"))

import numpy as np  # noqa: E402

import wpt  # noqa: E402

METRIC = "Mrays/s + Msamples/s at 1080p 64spp; batched vs plain; fwd+adjoint ms/iter"
WIDTH, HEIGHT, SPP, DEPTH, RR_START, CLAMP = 1920, 1080, 64, 12, 3, 10.0
HBM_PEAK_GBS = 8000.0
# Bytes per unit of work in the reference algorithm's data model: a BVH node
# record (AABB + two child words), a triangle's three vertices, a sphere
# (center, radius, material), a mesh-instance record (transforms + ids).
NODE_B, TRI_B, SPHERE_B, INSTANCE_B = 32, 36, 20, 96
RAY_IN_B, HIT_OUT_B = 24, 28


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, camera, threads, sample_spp):
    """Oracle FP32 render of `sample_spp` samples of the full frame on the
    host cores; returns the rate, the sample description and the work counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    oracle = pyoracle.Oracle(scene)
    cfg = wpt.render_config(WIDTH, HEIGHT, spp=sample_spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP)
    t0 = time.perf_counter()
    _, st = oracle.render(camera, cfg, threads=threads)
    dt = time.perf_counter() - t0
    queries = st.closest_queries + st.shadow_queries
    return {
        "value": round(queries / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{WIDTH}x{HEIGHT} x {sample_spp} spp of the same scene/camera/seeds ({st.paths} paths, "
                  f"{queries} queries, {dt:.2f} s), FP32 scalar restatement, {threads} std::threads",
        "msamples_per_s": round(st.paths / dt / 1e6, 4),
    }, st


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--mode", choices=["graph", "persistent"], default="graph")
    ap.add_argument("--skip-persistent", action="store_true", help="do not time the persistent kernel beside")
    ap.add_argument("--skip-cpu", action="store_true", help="do not time the CPU baseline")
    ap.add_argument("--cpu-spp", type=int, default=8, help="samples per pixel of the CPU baseline sample")
    ap.add_argument("--profile-spp", type=int, default=4, help="samples rendered with per-kernel events")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is a design, "
              f"not a measured configuration", file=sys.stderr)
        sys.exit(2)
    if wpt.device_count() == 0:
        print("bench.py needs a gfx950 device", file=sys.stderr)
        sys.exit(1)

    t0 = time.perf_counter()
    host = wpt.HostScene("sphere-field")
    t1 = time.perf_counter()
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    t2 = time.perf_counter()
    info = scene.info()
    log(f"scene: generated in {t1 - t0:.2f} s, BVH build + upload {t2 - t1:.2f} s, "
        f"{host.desc().mesh_triangle_count} BLAS triangles x {info.instance_count} instances, "
        f"{info.sphere_count} spheres ({info.direct_sphere_count} direct)")
    camera = host.camera(WIDTH, HEIGHT)
    mode = wpt.MODE_GRAPH if args.mode == "graph" else wpt.MODE_PERSISTENT
    cfg = wpt.render_config(WIDTH, HEIGHT, spp=args.spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP, mode=mode)
    renderer = wpt.Renderer(scene, camera, cfg)

    for i in range(args.warmup):
        renderer.reset()
        renderer.render_samples(0, args.spp)
        log(f"warmup step {i + 1}/{args.warmup} done")
    device_ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        renderer.reset()
        renderer.render_samples(0, args.spp)
        device_ms.append(renderer.stats().trace_milliseconds)
        log(f"step {i + 1}/{args.steps}: {device_ms[-1]:.1f} ms device")
    wall = time.perf_counter() - t_start
    stats = renderer.stats()
    queries = stats.closest_queries + stats.shadow_queries
    ms_per_step = wall * 1000.0 / args.steps
    value = queries * args.steps / wall / 1e6
    msamples = WIDTH * HEIGHT * args.spp * args.steps / wall / 1e6
    image, samples = renderer.framebuffer()
    log(f"graph: {value:.1f} Mrays/s, {msamples:.1f} Msamples/s, {ms_per_step:.1f} ms/frame, "
        f"{queries / (WIDTH * HEIGHT * args.spp):.3f} queries/path, mean radiance {image.mean() / samples:.4f}")

    # Per-kernel device time (events around each launch, no graph).
    kernel_ms, launches = renderer.profile_kernels(0, args.profile_spp)
    prof_stats = renderer.stats()
    trace_ms = kernel_ms[2] if mode == wpt.MODE_GRAPH else kernel_ms[4]
    trace_launches = int(launches[2] if mode == wpt.MODE_GRAPH else launches[4])
    log(f"kernels over {args.profile_spp} spp: primary {kernel_ms[0]:.2f} ms, reset {kernel_ms[1]:.2f} ms, "
        f"trace {kernel_ms[2]:.2f} ms ({int(launches[2])} launches), shade {kernel_ms[3]:.2f} ms, "
        f"persistent {kernel_ms[4]:.2f} ms, shadow {kernel_ms[5]:.2f} ms ({int(launches[5])} launches), "
        f"compact {kernel_ms[6]:.2f} ms")

    persistent = None
    if not args.skip_persistent and mode == wpt.MODE_GRAPH:
        pcfg = wpt.render_config(WIDTH, HEIGHT, spp=args.spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP,
                                 mode=wpt.MODE_PERSISTENT)
        pr = wpt.Renderer(scene, camera, pcfg)
        pr.reset()
        pr.render_samples(0, args.spp)
        tp = time.perf_counter()
        pr.reset()
        pr.render_samples(0, args.spp)
        pwall = time.perf_counter() - tp
        pst = pr.stats()
        pimg, _ = pr.framebuffer()
        persistent = {
            "mrays_per_s": round((pst.closest_queries + pst.shadow_queries) / pwall / 1e6, 2),
            "msamples_per_s": round(WIDTH * HEIGHT * args.spp / pwall / 1e6, 2),
            "ms_per_frame": round(pwall * 1000.0, 2),
            "image_identical_to_graph": bool(np.array_equal(pimg, image)),
        }
        log(f"persistent: {persistent}")
        pr.close()

    cpu = None
    work = None
    if not args.skip_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu, work = cpu_baseline(scene, camera, threads, args.cpu_spp)
        log(f"cpu baseline: {cpu}")

    roofline = None
    if work is not None and trace_ms > 0:
        per_query = (NODE_B * work.closest_node_tests + TRI_B * work.closest_triangle_tests +
                     SPHERE_B * work.closest_sphere_tests + INSTANCE_B * work.closest_instance_entries) / \
            max(1, work.closest_queries) + RAY_IN_B + HIT_OUT_B
        prof_queries = prof_stats.closest_queries
        achieved = per_query * prof_queries / (trace_ms / 1000.0) / 1e9
        roofline = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "kTrace" if mode == wpt.MODE_GRAPH else "kPersistent",
            "bytes_per_query": round(per_query, 1),
            "avg_launch_ms": round(trace_ms / max(1, trace_launches), 4),
            "launches": trace_launches,
        }

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {
            "workload": "sphere-field 1920x1080 64spp depth12 rr3 (3x 1,003,520-triangle icosphere + 485 spheres)",
            "width": WIDTH, "height": HEIGHT, "spp": args.spp, "max_depth": DEPTH, "rr_start": RR_START,
            "rng": "pcg32", "clamp": CLAMP, "mode": args.mode,
        },
        "msamples_per_s": round(msamples, 2),
        "queries_per_frame": int(queries),
        "device_ms_per_frame": round(float(np.mean(device_ms)), 2),
        "persistent": persistent,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

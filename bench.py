#!/usr/bin/env python3
"""Benchmark of the wavefront render path on one MI355X.

Workload (configuration 1 of the measurement plan): the sphere field -- a
direct-tested ground sphere of radius 1000 and 484 spheres of radius 0.2 on a
jittered 22x22 grid (80% Lambertian / 15% metal with fuzz U(0, 0.5) / 5%
dielectric 1.5) -- plus three TLAS instances of one BLAS, a displaced
geodesic icosphere of 1,003,520 triangles.  1920x1080, 64 spp, depth 12,
Russian roulette from depth 3, PCG32 streams, indirect clamp 10.

One step = one full frame (64 spp) rendered into the device framebuffer by
the chosen scheduler (default: the persistent kernel with path
regeneration); the scene is resident in HBM before the timed region.
`value` is Mrays/s = (closest-hit + shadow queries) / s; Msamples/s and the
rates of the other two schedulers (hipGraph-captured staged wavefront,
cooperative persistent kernel), with a check that their images are
identical, are reported beside it.

The roofline object is for the dominant kernel (the regenerating kernel, or
the closest-hit trace of the staged wavefront), timed with HIP events on the
renderer's stream; its algorithmic bytes per query follow the reference
algorithm's work counts (node tests, triangle/sphere tests, instance
entries) measured by the CPU oracle on the same scene and BVH.
The CPU baseline is the oracle (the scalar FP32 restatement of the reference
traversal and shading) on the host cores, over a bounded sample.

Beside the headline: the small Cornell box case (256x256, 16 spp, GPU and
the full frame on the host cores), batches of uneven tasks against tasks
one by one, the path-replay adjoint against the forward pass, and the
shading-heavy terrain with 4M caustic photons.
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
# HBM bytes per launch of the headline kernel on this workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a 256 MiB copy kernel
# (profiles/pmc_traffic_kBatch_sphere_field_1080p_64spp.md): one launch = one
# 64-sample frame of the regenerating scheduler.
MEASURED_TRAFFIC = {("kBatch", 64): 38.86e9}


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


def measure_cornell(threads, frames=20):
    """The small plain case: the unit Cornell box (12 triangles, emissive
    quad, Lambertian and dielectric spheres), 256x256, 16 spp, depth 8,
    Russian roulette from 3, on the GPU (regenerating scheduler) and in full
    on the host cores (the oracle), with the images compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    host = wpt.HostScene("cornell-box")
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    cam = host.camera(256, 256)
    cfg = wpt.render_config(256, 256, spp=16, max_depth=8, rr_start=3, clamp=CLAMP, mode=wpt.MODE_REGENERATE)
    r = wpt.Renderer(scene, cam, cfg)
    r.render_frame()
    t0 = time.perf_counter()
    for _ in range(frames):
        r.reset()
        r.render_samples(0, 16)
    wall = (time.perf_counter() - t0) / frames
    st = r.stats()
    img, _ = r.framebuffer()
    r.close()
    oracle = pyoracle.Oracle(scene)
    t0 = time.perf_counter()
    want, ost = oracle.render(cam, cfg, threads=threads)
    cpu_s = time.perf_counter() - t0
    queries = st.closest_queries + st.shadow_queries
    out = {
        "workload": "cornell-box 256x256 16spp depth8 rr3 (12 triangles, 2 spheres), regenerating scheduler",
        "mrays_per_s": round(queries / wall / 1e6, 2),
        "msamples_per_s": round(256 * 256 * 16 / wall / 1e6, 2),
        "ms_per_frame": round(wall * 1000.0, 3),
        "cpu_mrays_per_s": round((ost.closest_queries + ost.shadow_queries) / cpu_s / 1e6, 3),
        "cpu_ms_per_frame": round(cpu_s * 1000.0, 1),
        "cpu_cores": threads,
        "identical_to_cpu": bool(np.array_equal(img, want)),
    }
    log(f"cornell-box: {out}")
    return out


def batch_workload(n_tasks, seed=2024):
    """Batching workload: four resident scenes -- the Cornell box, the sphere
    field with its 1,003,520-triangle icosphere, the dielectric prism room
    (standing in for the shading-heavy scene, whose textures, environment and
    photon map are not built) and a sphere-field variant whose icosphere has
    200,000 triangles (frequency 100) -- and `n_tasks` square tiles of side
    {32, 64, 128, 256, 512} at seeded positions of a 1920x1080 frame, with
    spp uniform in [1, 64], mixed across the scenes."""
    specs = [("cornell-box", 0, 8), ("sphere-field", 0, 12), ("prism", 0, 10), ("sphere-field", 100, 12)]
    hosts = [wpt.HostScene(name, icosphere_frequency=freq) for name, freq, _ in specs]
    scenes = [wpt.Scene(h.desc(), device=True, owner=h) for h in hosts]
    cams = [h.camera(WIDTH, HEIGHT) for h in hosts]
    rng = np.random.default_rng(seed)
    tasks = []
    for _ in range(n_tasks):
        s = int(rng.integers(len(specs)))
        side = int(rng.choice([32, 64, 128, 256, 512]))
        x, y = int(rng.integers(0, WIDTH - side + 1)), int(rng.integers(0, HEIGHT - side + 1))
        spp = int(rng.integers(1, 65))
        tasks.append(wpt.batch_task(s, cams[s], WIDTH, HEIGHT, x, y, side, side, 0, spp, max_depth=specs[s][2],
                                    rr_start=RR_START, rng=wpt.RNG_PCG32, clamp=CLAMP))
    return hosts, scenes, tasks


def measure_batch(n_tasks):
    hosts, scenes, tasks = batch_workload(n_tasks)
    samples = sum(t.tile_width * t.tile_height * t.sample_count for t in tasks)
    batch = wpt.Batch(scenes)
    # Autotune on the first 32 tasks at up to 4 spp, then render everything.
    probe = [wpt.batch_task(t.scene, t.camera, t.width, t.height, t.x, t.y, t.tile_width, t.tile_height, 0,
                            min(4, t.sample_count), t.max_depth, t.rr_start_depth, t.rng_strategy, t.indirect_clamp)
             for t in tasks[:32]]
    ta = time.perf_counter()
    chosen = batch.autotune(probe)
    tune_s = time.perf_counter() - ta
    log(f"batch: autotuned in {tune_s:.1f} s -> block {chosen.block_size}, grid {chosen.blocks}")
    batch.render(tasks[:8])  # warm-up
    tb = time.perf_counter()
    outs, lat = batch.render(tasks, latencies=True)
    batch_s = time.perf_counter() - tb
    bst = batch.stats()
    log(f"batch: {n_tasks} tasks, {samples / 1e6:.1f} M samples in {batch_s * 1000:.1f} ms "
        f"(device {bst.milliseconds:.1f} ms)")

    # The same tasks one by one through wpt_render_task, as the reference's
    # RenderTaskRenderer renders them: one renderer per scene and tile size,
    # its queues sized to the tile.
    renderers = {}
    for t in tasks:
        key = (t.scene, t.tile_width)
        if key not in renderers:
            cfg = wpt.render_config(WIDTH, HEIGHT, spp=64, max_depth=t.max_depth, rr_start=RR_START, clamp=CLAMP,
                                    tile=(t.tile_width, t.tile_height))
            renderers[key] = wpt.Renderer(scenes[t.scene], hosts[t.scene].camera(WIDTH, HEIGHT), cfg)
    for (_, side), r in renderers.items():
        r.render_task(wpt.Task(0, 0, 0, 1, 0, 0, side, side))
    done = []
    identical = True
    tp = time.perf_counter()
    for i, t in enumerate(tasks):
        out = renderers[(t.scene, t.tile_width)].render_task(
            wpt.Task(0, 0, t.sample_start, t.sample_count, t.x, t.y, t.tile_width, t.tile_height))
        done.append((time.perf_counter() - tp) * 1000.0)
        identical &= bool(np.array_equal(out, outs[i]))
    plain_s = time.perf_counter() - tp
    for r in renderers.values():
        r.close()
    lat_plain = np.array(done)
    result = {
        "workload": f"{n_tasks} tasks over 4 resident scenes (cornell-box, sphere-field 1.0M-tri icosphere, prism, "
                    f"sphere-field 200k-tri icosphere); square tiles {{32..512}}, spp U[1,64], seed 2024",
        "samples": int(samples),
        "batched_msamples_per_s": round(samples / batch_s / 1e6, 2),
        "plain_msamples_per_s": round(samples / plain_s / 1e6, 2),
        "speedup": round(plain_s / batch_s, 3),
        "batched_ms": round(batch_s * 1000.0, 2),
        "plain_ms": round(plain_s * 1000.0, 2),
        "batched_latency_ms_p50_p99": [round(float(np.percentile(lat, 50)), 2), round(float(np.percentile(lat, 99)), 2)],
        "plain_latency_ms_p50_p99": [round(float(np.percentile(lat_plain, 50)), 2),
                                     round(float(np.percentile(lat_plain, 99)), 2)],
        "autotuned": {"block_size": int(chosen.block_size), "blocks": int(chosen.blocks),
                      "tuning_s": round(tune_s, 2)},
        "identical_to_plain": identical,
    }
    log(f"batch vs plain: {result}")
    batch.close()
    return result


def diff_parameters(layout, seed, truth):
    """Parameter vector [texture 512x512 RGB | 64 roughness | 8 emission].
    The initial guess (truth=False) is a flat grey texture with noise; the
    ground truth is a smooth pattern of seeded sinusoids."""
    rng = np.random.default_rng(seed)
    side = int(round((layout.texture_floats // 3) ** 0.5))
    if truth:
        yy, xx = np.mgrid[0:side, 0:side] / side
        tex = np.zeros((side, side, 3))
        for c in range(3):
            acc = np.zeros((side, side))
            for _ in range(4):
                fx, fy, ph = rng.uniform(1, 6), rng.uniform(1, 6), rng.uniform(0, 2 * np.pi)
                acc += np.sin(2 * np.pi * (fx * xx + fy * yy) + ph)
            tex[..., c] = 0.5 + 0.1 * acc
        tex = np.clip(tex, 0.05, 0.95)
        rough = rng.uniform(0.05, 0.95, layout.roughness_count)
        emit = rng.uniform(0.5, 1.5, layout.emission_count)
    else:
        tex = np.clip(0.5 + rng.uniform(-0.05, 0.05, (side, side, 3)), 0, 1)
        rough = rng.uniform(0.3, 0.7, layout.roughness_count)
        emit = rng.uniform(0.8, 1.2, layout.emission_count)
    return np.concatenate([tex.reshape(-1), rough, emit]).astype(np.float32)


def measure_diff(steps):
    """Differentiation workload: the Cornell box tessellated into 64
    reflective + 8 emissive triangles, a 512x512 RGB albedo texture on the
    back wall, per-triangle GGX roughness and per-light emission scales.
    Target: 512x512 at 256 spp from the seed-B parameters; optimisation:
    L2 loss at 512x512, 16 spp, depth 6, path-replay adjoint, Adam from the
    seed-A parameters."""
    import torch
    size, spp, depth = 512, 16, 6
    host = wpt.HostScene("diff-cornell")
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    cam = host.camera(size, size)
    r = wpt.DiffRenderer(scene, cam, size, size, spp, depth, (512, 512), 5)
    lay = r.layout
    truth = torch.tensor(diff_parameters(lay, 0xB, True), device="cuda:0")
    target_renderer = wpt.DiffRenderer(scene, cam, size, size, 256, depth, (512, 512), 5)
    target = target_renderer.forward(truth, sample_start=1 << 20)
    target_renderer.close()
    params = torch.tensor(diff_parameters(lay, 0xA, False), device="cuda:0")
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    # Warm-up, then time forward alone and forward + adjoint.
    img = r.forward(params)
    r.adjoint(params, 2.0 * (img - target) / img.numel())
    fwd_ms, adj_ms, losses = [], [], []
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        img = r.forward(params, sample_start=step * spp)
        g_img = 2.0 * (img - target) / img.numel()
        losses.append(float((g_img * g_img).sum() * img.numel() / 4.0))
        grad = r.adjoint(params, g_img)
        st = r.stats()
        fwd_ms.append(st.forward_milliseconds)
        adj_ms.append(st.adjoint_milliseconds)
        r.adam(params, grad, m, v, step, 0.01)
    wall = time.perf_counter() - t0
    err = lambda a, b: float((a - b).abs().mean())
    t, rc = lay.texture_floats, lay.roughness_count
    result = {
        "workload": "diff-cornell 512x512, 16 spp, depth 6; 512x512 RGB albedo texture + 64 roughness + 8 emission; "
                    "target 256 spp from seed B, Adam lr 0.01 from seed A",
        "forward_ms": round(float(np.median(fwd_ms)), 3),
        "adjoint_ms": round(float(np.median(adj_ms)), 3),
        "fwd_adjoint_ms_per_iter": round(float(np.median(np.array(fwd_ms) + np.array(adj_ms))), 3),
        "ratio_to_forward": round(float(np.median((np.array(fwd_ms) + np.array(adj_ms)) / np.array(fwd_ms))), 3),
        "wall_ms_per_iter": round(wall * 1000.0 / steps, 3),
        "adam_steps": steps,
        "loss_first_last": [losses[0], losses[-1]],
        "param_mae_start_end": {
            "texture": [err(torch.tensor(diff_parameters(lay, 0xA, False)[:t]), truth[:t].cpu()),
                        err(params[:t], truth[:t])],
            "roughness": [err(torch.tensor(diff_parameters(lay, 0xA, False)[t:t + rc]), truth[t:t + rc].cpu()),
                          err(params[t:t + rc], truth[t:t + rc])],
            "emission": [err(torch.tensor(diff_parameters(lay, 0xA, False)[t + rc:]), truth[t + rc:].cpu()),
                         err(params[t + rc:], truth[t + rc:])],
        },
    }
    log(f"diff: {result}")
    r.close()
    return result


def measure_terrain(spp, photons, frames):
    """Shading-heavy workload with caustics: the terrain preset at full size
    (2,006,680 triangles, 16 x 4096^2 RGBA8 textures, 512 emissive
    triangles, a 2048 x 1024 HDR sky with a sun, dielectric crystals) with a
    spectral photon map, 1920x1080, depth 10, regenerating scheduler; BLASes
    built on the device."""
    t0 = time.perf_counter()
    host = wpt.HostScene("terrain")
    t1 = time.perf_counter()
    desc = wpt.SceneDesc.from_buffer_copy(host.desc())
    desc.bvh_builder = wpt.BVH_DEVICE_PLOC
    scene = wpt.Scene(desc, device=True, owner=host)
    t2 = time.perf_counter()
    info = scene.info()
    log(f"terrain: generated in {t1 - t0:.2f} s, scene build + upload {t2 - t1:.2f} s "
        f"(device BLAS build {info.blas_build_milliseconds:.1f} ms)")
    status, rc, _, _ = wpt.parse_command_line(
        ["--scene", "terrain", "--caustics", "--caustic-photons", str(photons), "--resolution",
         f"{WIDTH}x{HEIGHT}", "--spp", str(spp), "--max-depth", "10", "--rr-depth", "3", "--rng", "pcg32",
         "--regenerate"])
    assert status == wpt.OK
    r = wpt.Renderer(scene, host.camera(WIDTH, HEIGHT), rc)
    photon = r.stats()
    r.reset()
    r.render_samples(0, spp)  # warmup
    tp = time.perf_counter()
    for _ in range(frames):
        r.reset()
        r.render_samples(0, spp)
    wall = (time.perf_counter() - tp) / frames
    st = r.stats()
    img, n = r.framebuffer()
    r.close()
    out = {
        "workload": f"terrain 1920x1080 {spp}spp depth10 rr3: 2,006,680 triangles (terrain + 4 rock meshes "
                    f"+ foliage + crystals + 512 lamp triangles, 162 instances), 16 x 4096^2 RGBA8 textures, "
                    f"2048x1024 HDR sky, {photons} spectral caustic photons",
        "mrays_per_s": round((st.closest_queries + st.shadow_queries) / wall / 1e6, 2),
        "msamples_per_s": round(WIDTH * HEIGHT * spp / wall / 1e6, 2),
        "ms_per_frame": round(wall * 1000.0, 2),
        "device_ms_per_frame": round(st.trace_milliseconds, 2),
        "queries_per_path": round((st.closest_queries + st.shadow_queries) / (WIDTH * HEIGHT * spp), 3),
        "photon_map_ms": round(photon.caustic_build_milliseconds, 2),
        "photons_stored": int(photon.caustic_photons_stored),
        "device_blas_build_ms": round(info.blas_build_milliseconds, 2),
        "mean_radiance": round(float(img.mean() / max(1, n)), 5),
    }
    log(f"terrain: {out}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--mode", choices=["regenerate", "graph", "persistent"], default="regenerate",
                    help="scheduler of the timed frames; the other two are timed beside it")
    ap.add_argument("--skip-persistent", action="store_true", help="do not time the other schedulers beside")
    ap.add_argument("--skip-cpu", action="store_true", help="do not time the CPU baseline")
    ap.add_argument("--cpu-spp", type=int, default=8, help="samples per pixel of the CPU baseline sample")
    ap.add_argument("--profile-spp", type=int, default=4, help="samples rendered with per-kernel events")
    ap.add_argument("--skip-batch", action="store_true", help="do not measure batched vs one-by-one tasks")
    ap.add_argument("--batch-tasks", type=int, default=256, help="tasks of the batching workload")
    ap.add_argument("--skip-diff", action="store_true", help="do not measure the differentiable renderer")
    ap.add_argument("--skip-terrain", action="store_true", help="do not measure the shading-heavy caustic workload")
    ap.add_argument("--skip-cornell", action="store_true", help="do not measure the small Cornell box case")
    ap.add_argument("--terrain-spp", type=int, default=64)
    ap.add_argument("--terrain-photons", type=int, default=4 * 1024 * 1024)
    ap.add_argument("--diff-steps", type=int, default=200, help="Adam steps of the inverse-rendering workload")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB three times: the reference kernel of "
                         "known traffic against which the FETCH_SIZE/WRITE_SIZE counters are scaled")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is a design, "
              f"not a measured configuration", file=sys.stderr)
        sys.exit(2)
    if args.calibration_copy_mib or not args.skip_diff:
        # torch ships its own HIP runtime beside the one libwpt links; it
        # must initialise the device before libwpt's first device call.
        import torch
        torch.zeros(1, device="cuda:0")
        torch.cuda.synchronize()
    if wpt.device_count() == 0:
        print("bench.py needs a gfx950 device", file=sys.stderr)
        sys.exit(1)

    if args.calibration_copy_mib:
        import torch
        n = args.calibration_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: 3 x {args.calibration_copy_mib} MiB device-to-device")
        del src, dst

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
    modes = {"regenerate": wpt.MODE_REGENERATE, "graph": wpt.MODE_GRAPH, "persistent": wpt.MODE_PERSISTENT}
    mode = modes[args.mode]
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
    log(f"{args.mode}: {value:.1f} Mrays/s, {msamples:.1f} Msamples/s, {ms_per_step:.1f} ms/frame, "
        f"{queries / (WIDTH * HEIGHT * args.spp):.3f} queries/path, mean radiance {image.mean() / samples:.4f}")

    # Device time of the dominant kernel.  The regenerating scheduler is one
    # kBatch launch per frame: its HIP-event time is the frames' own.  The
    # staged scheduler's kernels are timed one by one over a few samples.
    if mode == wpt.MODE_REGENERATE:
        trace_ms = float(np.sum(device_ms))
        trace_launches = args.steps
        prof_closest, prof_shadow = stats.closest_queries * args.steps, stats.shadow_queries * args.steps
    else:
        kernel_ms, launches = renderer.profile_kernels(0, args.profile_spp)
        prof_stats = renderer.stats()
        trace_slot = 2 if mode == wpt.MODE_GRAPH else 4
        trace_ms = kernel_ms[trace_slot]
        trace_launches = int(launches[trace_slot])
        prof_closest, prof_shadow = prof_stats.closest_queries, prof_stats.shadow_queries
        log(f"kernels over {args.profile_spp} spp: primary {kernel_ms[0]:.2f} ms, reset {kernel_ms[1]:.2f} ms, "
            f"trace {kernel_ms[2]:.2f} ms ({int(launches[2])} launches), shade {kernel_ms[3]:.2f} ms, "
            f"persistent {kernel_ms[4]:.2f} ms ({int(launches[4])} launches), shadow {kernel_ms[5]:.2f} ms "
            f"({int(launches[5])} launches)")
    renderer.close()

    others = {}
    if not args.skip_persistent:
        for name, m in modes.items():
            if m == mode:
                continue
            ocfg = wpt.render_config(WIDTH, HEIGHT, spp=args.spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP,
                                     mode=m)
            orr = wpt.Renderer(scene, camera, ocfg)
            orr.reset()
            orr.render_samples(0, args.spp)
            tp = time.perf_counter()
            orr.reset()
            orr.render_samples(0, args.spp)
            owall = time.perf_counter() - tp
            ost = orr.stats()
            oimg, _ = orr.framebuffer()
            others[name] = {
                "mrays_per_s": round((ost.closest_queries + ost.shadow_queries) / owall / 1e6, 2),
                "msamples_per_s": round(WIDTH * HEIGHT * args.spp / owall / 1e6, 2),
                "ms_per_frame": round(owall * 1000.0, 2),
                "image_identical": bool(np.array_equal(oimg, image)),
            }
            log(f"{name}: {others[name]}")
            orr.close()

    threads = min(16, os.cpu_count() or 1)
    cpu = None
    if not args.skip_cpu:
        cpu, work = cpu_baseline(scene, camera, threads, args.cpu_spp)
        log(f"cpu baseline: {cpu}")
    else:
        # Work counts for the byte model from one sample of the frame.
        _, work = cpu_baseline(scene, camera, threads, 1)

    cornell = None
    if not args.skip_cornell:
        cornell = measure_cornell(threads)

    batch = None
    if not args.skip_batch:
        batch = measure_batch(args.batch_tasks)

    diff = None
    if not args.skip_diff:
        diff = measure_diff(args.diff_steps)

    terrain = None
    if not args.skip_terrain:
        terrain = measure_terrain(args.terrain_spp, args.terrain_photons, 1)

    roofline = None
    if work is not None and trace_ms > 0:
        def traversal_bytes(nodes, tris, spheres, instances, count):
            return (NODE_B * nodes + TRI_B * tris + SPHERE_B * spheres + INSTANCE_B * instances) / max(1, count)
        closest_b = traversal_bytes(work.closest_node_tests, work.closest_triangle_tests, work.closest_sphere_tests,
                                    work.closest_instance_entries, work.closest_queries)
        shadow_b = traversal_bytes(work.shadow_node_tests, work.shadow_triangle_tests, work.shadow_sphere_tests,
                                   work.shadow_instance_entries, work.shadow_queries)
        if mode == wpt.MODE_GRAPH:
            # kTrace: closest-hit traversal plus its ray in and hit out.
            per_query = closest_b + RAY_IN_B + HIT_OUT_B
            algorithmic = per_query * prof_closest
        else:
            # One kernel does every query; paths stay in registers.
            per_query = closest_b
            algorithmic = closest_b * prof_closest + shadow_b * prof_shadow
        achieved = algorithmic / (trace_ms / 1000.0) / 1e9
        kernel = {wpt.MODE_GRAPH: "kTrace", wpt.MODE_PERSISTENT: "kPersistent", wpt.MODE_REGENERATE: "kBatch"}[mode]
        roofline = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "kernel": kernel,
            "traffic": MEASURED_TRAFFIC.get((kernel, args.spp)),
            "bytes_per_closest_query": round(per_query, 1),
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
            "rng": "pcg32", "clamp": CLAMP, "scheduler": args.mode,
        },
        "msamples_per_s": round(msamples, 2),
        "queries_per_frame": int(queries),
        "device_ms_per_frame": round(float(np.mean(device_ms)), 2),
        "other_schedulers": others,
        "plain_small": cornell,
        "batch": batch,
        "diff": diff,
        "shading_heavy": terrain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

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
# HBM bytes per launch of the dominant kernel on this workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a 256 MiB copy kernel
# (FETCH_SIZE reads 131086 KiB for it: x2.000; WRITE_SIZE 262144 KiB: x1.000),
# with the counters that say what bounds it.  kTrace: average of the 12
# launches (bounces) of one 64-sample pass of the staged scheduler
# (profiles/pmc_kTrace_sphere_field_1080p_64spp.md); kBatch: one 64-sample
# frame of the regenerating scheduler
# (profiles/pmc_kBatch_sphere_field_1080p_64spp.md).
MEASURED_TRAFFIC = {("kTrace", 64): 46.84e9 / 12, ("kBatch", 64): 39.21e9}
MEASURED_COUNTERS = {
    ("kTrace", 64): {"l2_hit": 0.515, "valu_busy": 0.757, "waves_per_cu": 12.14, "wait_any": 0.526,
                     "lanes_active": 0.257, "l1_hit": 0.904, "ta_busy": 0.485, "l1_to_l2_latency_cycles": 306},
    ("kBatch", 64): {"l2_hit": 0.615, "valu_busy": 0.303, "waves_per_cu": 7.87, "wait_any": 0.672},
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, camera, threads, sample_spp, runs=5):
    """Oracle FP32 render of `sample_spp` samples of the full frame on the
    host cores, `runs` times; returns the median rate, the sample
    description and the work counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    oracle = pyoracle.Oracle(scene)
    cfg = wpt.render_config(WIDTH, HEIGHT, spp=sample_spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, st = oracle.render(camera, cfg, threads=threads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    queries = st.closest_queries + st.shadow_queries
    return {
        "value": round(queries / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{WIDTH}x{HEIGHT} x {sample_spp} spp of the same scene/camera/seeds ({st.paths} paths, "
                  f"{queries} queries), median of {runs} runs ({dt:.2f} s; {min(times):.2f}-{max(times):.2f} s), "
                  f"FP32 scalar restatement, {threads} std::threads",
        "msamples_per_s": round(st.paths / dt / 1e6, 4),
    }, st


def primary_rays(camera, width, height):
    """Camera rays through the pixel centres (SoA float32, 6 x w*h), row 0 at
    the top: position + normalize(f + r*x + u*y) with
    x = (2 (i + 0.5) / w - 1) * aspect * tan(fov/2), y = (1 - 2 (j + 0.5) / h) * tan(fov/2)."""
    p, f, r, u = (np.array(list(v), np.float64) for v in (camera.position, camera.forward, camera.right, camera.up))
    t = np.tan(camera.vertical_fov * 0.5)
    j, i = np.mgrid[0:height, 0:width]
    x = (2.0 * (i.reshape(-1) + 0.5) / width - 1.0) * camera.aspect_ratio * t
    y = (1.0 - 2.0 * (j.reshape(-1) + 0.5) / height) * t
    d = f[None, :] + r[None, :] * x[:, None] + u[None, :] * y[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((6, width * height), np.float32)
    rays[0:3] = p[:, None]
    rays[3:6] = d.T
    return rays


def measure_traversal(scene, camera, threads, reps=5):
    """Closest-hit traversal alone (wpt_trace_closest against the oracle's
    traceRay restatement on the host cores) on two ray sets of the headline
    scene: the reference acceleration test's rays (8 fixed + 2048 from
    mt19937 seed 12345, tests/scene_acceleration_tests.cu:184-201) and one
    primary wavefront of the 1920x1080 frame.  Median of `reps` runs each;
    hit records compared bit for bit."""
    import concurrent.futures
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    oracle = pyoracle.Oracle(scene)
    out = {}
    sets = {"reference_rays_2056": pyoracle.reference_rays(12345, 2048),
            "primary_wavefront_1080p": primary_rays(camera, WIDTH, HEIGHT)}
    pool = concurrent.futures.ThreadPoolExecutor(threads)
    for name, rays in sets.items():
        n = rays.shape[1]
        d_rays = torch.from_numpy(rays).cuda()
        hits = scene.trace_closest(d_rays)
        torch.cuda.synchronize()
        gpu = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            hits = scene.trace_closest(d_rays)
            b.record()
            b.synchronize()
            gpu.append(a.elapsed_time(b) / 1000.0)
        gpu_hits = wpt.hits_to_numpy(hits)
        chunks = np.array_split(np.arange(n), threads)
        cpu = []
        for _ in range(reps):
            t0 = time.perf_counter()
            parts = list(pool.map(lambda c: oracle.trace(np.ascontiguousarray(rays[:, c])), chunks))
            cpu.append(time.perf_counter() - t0)
        cpu_hits = np.concatenate(parts)
        g, c = float(np.median(gpu)), float(np.median(cpu))
        out[name] = {"rays": int(n), "gpu_mrays_per_s": round(n / g / 1e6, 2), "gpu_ms": round(g * 1000.0, 4),
                     "cpu_mrays_per_s": round(n / c / 1e6, 3), "cpu_ms": round(c * 1000.0, 2), "cpu_threads": threads,
                     "hits_identical": gpu_hits.tobytes() == cpu_hits.tobytes(),
                     "hit_fraction": round(float((gpu_hits["kind"] != 0).mean()), 4)}
        log(f"traversal {name}: {out[name]}")
    pool.shutdown()
    return out


def measure_flat_accelerators(reps=5):
    """Flat-triangle accelerators (the reference's nexus_bvh2, nexus_bvh8,
    median_bvh and spatial_split_bvh policies, src/scene/scene.cu:241-294)
    on the sphere field's 1,003,520-triangle icosphere uploaded as flat
    triangles: host build time, and closest-hit rates on a 1920x1080
    pinhole wavefront aimed at the mesh and on 2M scattered rays (origins
    inside the mesh's box, uniform directions).  Every accelerator must give
    the same hit records, bit for bit."""
    import ctypes
    import torch

    host = wpt.HostScene("sphere-field")
    src = host.desc()
    n = src.mesh_triangle_count
    raw = (ctypes.c_uint8 * (n * ctypes.sizeof(wpt.Triangle))).from_address(
        ctypes.addressof(src.mesh_triangles.contents))
    verts = np.frombuffer(raw, np.uint8).reshape(n, -1)[:, :36].copy().view(np.float32).reshape(n, 3, 3)
    lo, hi = verts.min(axis=(0, 1)).astype(np.float64), verts.max(axis=(0, 1)).astype(np.float64)
    centre, extent = 0.5 * (lo + hi), float(np.max(hi - lo))
    # Pinhole wavefront from outside the box, looking at its centre.
    eye = centre + np.array([0.35, 0.25, 1.0]) * 1.2 * extent
    f = centre - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, [0.0, 1.0, 0.0])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    t = np.tan(np.radians(40.0) * 0.5)
    j, i = np.mgrid[0:HEIGHT, 0:WIDTH]
    x = ((2.0 * (i.reshape(-1) + 0.5) / WIDTH - 1.0) * (WIDTH / HEIGHT) * t)[:, None]
    y = ((1.0 - 2.0 * (j.reshape(-1) + 0.5) / HEIGHT) * t)[:, None]
    d = f[None, :] + r[None, :] * x + u[None, :] * y
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    camera_rays = np.empty((6, WIDTH * HEIGHT), np.float32)
    camera_rays[0:3] = eye[:, None]
    camera_rays[3:6] = d.T
    rng = np.random.default_rng(12345)
    m = 2 * 1024 * 1024
    scattered = np.empty((6, m), np.float32)
    scattered[0:3] = (lo[:, None] + (hi - lo)[:, None] * rng.random((3, m)))
    v = rng.normal(size=(3, m))
    scattered[3:6] = v / np.linalg.norm(v, axis=0)
    sets = {"camera_1080p": torch.from_numpy(camera_rays).cuda(), "scattered_2m": torch.from_numpy(scattered).cuda()}

    names = {wpt.FLAT_SAH_BVH2: "sah_bvh2", wpt.FLAT_MEDIAN_BVH2: "median_bvh2",
             wpt.FLAT_SPATIAL_SPLIT_BVH2: "spatial_split_bvh2", wpt.FLAT_WIDE8: "wide8"}
    out, first = {}, {}
    identical = True
    for accel, name in names.items():
        desc = wpt.SceneDesc()
        desc.flat_triangles, desc.flat_triangle_count = src.mesh_triangles, n
        desc.materials, desc.material_count = src.materials, src.material_count
        desc.flat_accelerator = accel
        t0 = time.perf_counter()
        scene = wpt.Scene(desc, device=True, owner=host)
        build_s = time.perf_counter() - t0
        info = scene.info()
        row = {"build_and_upload_s": round(build_s, 2), "leaf_slots": int(info.flat_leaf_slots),
               "wide_nodes": int(info.flat_wide_node_count), "max_stack_depth": int(info.max_stack_depth)}
        for set_name, rays in sets.items():
            hits = scene.trace_closest(rays)
            torch.cuda.synchronize()
            times = []
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                hits = scene.trace_closest(rays)
                b.record()
                b.synchronize()
                times.append(a.elapsed_time(b) / 1000.0)
            h = wpt.hits_to_numpy(hits)
            if set_name in first:
                identical = identical and h.tobytes() == first[set_name]
            else:
                first[set_name] = h.tobytes()
            g = float(np.median(times))
            row[set_name] = {"mrays_per_s": round(rays.shape[1] / g / 1e6, 1), "ms": round(g * 1000.0, 3),
                             "hit_fraction": round(float((h["kind"] != 0).mean()), 4)}
        scene.close()
        out[name] = row
        log(f"flat accelerator {name}: {row}")
    out["hits_identical"] = identical
    out["triangles"] = int(n)
    return out


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


def terrain_scene():
    """The shading-heavy terrain preset at full size, BLASes built on the
    device, and its render configuration (caustics with `photons` spectral
    photons, 1920x1080, depth 10, Russian roulette from 3, PCG32)."""
    host = wpt.HostScene("terrain")
    desc = wpt.SceneDesc.from_buffer_copy(host.desc())
    desc.bvh_builder = wpt.BVH_DEVICE_PLOC
    scene = wpt.Scene(desc, device=True, owner=host)
    return host, scene


def terrain_config(photons, spp=64, mode="--regenerate"):
    status, rc, _, _ = wpt.parse_command_line(
        ["--scene", "terrain", "--caustics", "--caustic-photons", str(photons), "--resolution",
         f"{WIDTH}x{HEIGHT}", "--spp", str(spp), "--max-depth", "10", "--rr-depth", "3", "--rng", "pcg32", mode])
    assert status == wpt.OK
    return rc


def batch_workload(n_tasks, photons, seed=2024):
    """Batching workload: four resident scenes -- the Cornell box, the sphere
    field with its 1,003,520-triangle icosphere, the shading-heavy terrain
    (2,006,680 triangles, 16 x 4096^2 textures, HDR sky, `photons` caustic
    photons in its own photon map) and a sphere-field variant whose
    icosphere has 200,000 triangles (frequency 100) -- and `n_tasks` square
    tiles of side {32, 64, 128, 256, 512} at seeded positions of a 1920x1080
    frame, with spp uniform in [1, 64], mixed across the scenes."""
    specs = [("cornell-box", 0, 8), ("sphere-field", 0, 12), ("terrain", 0, 10), ("sphere-field", 100, 12)]
    hosts, scenes = [], []
    for name, freq, _ in specs:
        if name == "terrain":
            h, sc = terrain_scene()
        else:
            h = wpt.HostScene(name, icosphere_frequency=freq)
            sc = wpt.Scene(h.desc(), device=True, owner=h)
        hosts.append(h)
        scenes.append(sc)
    configs = [None, None, terrain_config(photons), None]
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
    return hosts, scenes, configs, tasks


def tasks_one_by_one(hosts, scenes, configs, tasks, mode, outs):
    """The batch's tasks through wpt_render_task one at a time, as the
    reference's RenderTaskRenderer renders them: one renderer per scene and
    tile size (its queues sized to the tile; a caustic scene's renderer
    builds its own photon map), all created before the clock starts.
    Returns the wall time, per-task completion times and whether every task
    equals the batch's output."""
    renderers = {}
    for t in tasks:
        key = (t.scene, t.tile_width)
        if key not in renderers:
            if configs[t.scene] is not None:
                cfg = wpt.RenderConfig.from_buffer_copy(configs[t.scene])
            else:
                cfg = wpt.render_config(WIDTH, HEIGHT, spp=64, max_depth=t.max_depth, rr_start=RR_START, clamp=CLAMP)
            cfg.tiled, cfg.tile_width, cfg.tile_height, cfg.mode = 1, t.tile_width, t.tile_height, mode
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
    wall = time.perf_counter() - tp
    for r in renderers.values():
        r.close()
    return wall, np.array(done), identical


def measure_batch(n_tasks, photons):
    hosts, scenes, configs, tasks = batch_workload(n_tasks, photons)
    samples = sum(t.tile_width * t.tile_height * t.sample_count for t in tasks)
    batch = wpt.Batch(scenes)  # staged passes
    stored, photon_ms = batch.set_caustics(2, configs[2])
    # Autotune on the first 32 tasks at up to 4 spp, then render everything.
    probe = [wpt.batch_task(t.scene, t.camera, t.width, t.height, t.x, t.y, t.tile_width, t.tile_height, 0,
                            min(4, t.sample_count), t.max_depth, t.rr_start_depth, t.rng_strategy, t.indirect_clamp)
             for t in tasks[:32]]
    ta = time.perf_counter()
    chosen = batch.autotune(probe)
    tune_s = time.perf_counter() - ta
    log(f"batch: terrain photon map {stored} photons stored in {photon_ms:.1f} ms; autotuned in {tune_s:.1f} s "
        f"-> scheduler {chosen.scheduler}, block {chosen.block_size}, grid {chosen.blocks}")
    batch.render(tasks[:8])  # warm-up
    tb = time.perf_counter()
    outs, lat = batch.render(tasks, latencies=True)
    batch_s = time.perf_counter() - tb
    bst = batch.stats()
    batch.close()
    scheduler = {wpt.BATCH_STAGED: "staged", wpt.BATCH_REGENERATE: "regenerate"}[chosen.scheduler]
    log(f"batch ({scheduler}): {n_tasks} tasks, {samples / 1e6:.1f} M samples in {batch_s * 1000:.1f} ms "
        f"(device {bst.milliseconds:.1f} ms)")
    # The same batch with the other scheduler: one persistent launch over
    # every task (regenerating kernel), autotuned on the same probe.
    other_scheduler = wpt.BATCH_REGENERATE if chosen.scheduler == wpt.BATCH_STAGED else wpt.BATCH_STAGED
    other = wpt.Batch(scenes, scheduler=other_scheduler)
    other.set_caustics(2, configs[2])
    other.autotune(probe)
    other.render(tasks[:8])
    to = time.perf_counter()
    other_outs, other_lat = other.render(tasks, latencies=True)
    other_s = time.perf_counter() - to
    other.close()
    other_same = all(np.array_equal(a, b) for a, b in zip(outs, other_outs))
    log(f"batch ({'regenerate' if other_scheduler == wpt.BATCH_REGENERATE else 'staged'}): {other_s * 1000:.1f} ms, "
        f"identical {other_same}")
    # Tasks one by one with the same kernel (regenerating scheduler), then
    # with the staged hipGraph wavefront.
    regen_s, regen_done, regen_same = tasks_one_by_one(hosts, scenes, configs, tasks, wpt.MODE_REGENERATE, outs)
    log(f"batch: one by one, regenerating kernel: {regen_s * 1000:.1f} ms")
    graph_s, graph_done, graph_same = tasks_one_by_one(hosts, scenes, configs, tasks, wpt.MODE_GRAPH, outs)
    log(f"batch: one by one, staged wavefront: {graph_s * 1000:.1f} ms")

    def pct(a):
        return [round(float(np.percentile(a, 50)), 2), round(float(np.percentile(a, 99)), 2)]

    result = {
        "workload": f"{n_tasks} tasks over 4 resident scenes (cornell-box, sphere-field 1.0M-tri icosphere, terrain "
                    f"2.0M tris + 16 4096^2 textures + HDR sky + {photons} caustic photons, sphere-field 200k-tri "
                    f"icosphere); square tiles {{32..512}}, spp U[1,64], seed 2024",
        "samples": int(samples),
        "batched_msamples_per_s": round(samples / batch_s / 1e6, 2),
        "batched_ms": round(batch_s * 1000.0, 2),
        "batched_latency_ms_p50_p99": pct(lat),
        "batched_scheduler": scheduler,
        "batched_other_scheduler": {
            "scheduler": "regenerate" if other_scheduler == wpt.BATCH_REGENERATE else "staged",
            "ms": round(other_s * 1000.0, 2), "msamples_per_s": round(samples / other_s / 1e6, 2),
            "latency_ms_p50_p99": pct(other_lat), "identical": other_same},
        "one_by_one_same_kernel": {
            "msamples_per_s": round(samples / regen_s / 1e6, 2), "ms": round(regen_s * 1000.0, 2),
            "latency_ms_p50_p99": pct(regen_done), "speedup_of_batch": round(regen_s / batch_s, 3),
            "identical": regen_same},
        "one_by_one_staged": {
            "msamples_per_s": round(samples / graph_s / 1e6, 2), "ms": round(graph_s * 1000.0, 2),
            "latency_ms_p50_p99": pct(graph_done), "speedup_of_batch": round(graph_s / batch_s, 3),
            "identical": graph_same},
        "terrain_photons_stored": int(stored),
        "autotuned": {"block_size": int(chosen.block_size), "blocks": int(chosen.blocks), "scheduler": scheduler,
                      "tuning_s": round(tune_s, 2)},
    }
    log(f"batch vs plain: {result}")
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
    result["roughness_observability"] = roughness_observability(r, truth, target, lay)
    result["decorrelated_loss"] = decorrelated_adam(r, truth, target, lay, steps, spp)
    log(f"diff: {result}")
    r.close()
    return result


def decorrelated_adam(r, truth, target, lay, steps, spp):
    """The same Adam run with the gradient of (X1 - T)(X2 - T) for two
    independent 16-spp renders X1, X2 (the adjoint of X1's paths with X2's
    residual) instead of (X - T)^2.  The squared loss of one noisy render
    adds the render's variance, which falls as roughness grows, so its
    expected gradient pulls roughness up; the product of independent
    residuals is an unbiased estimate of the squared error of the mean."""
    import torch
    params = torch.tensor(diff_parameters(lay, 0xA, False), device="cuda:0")
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        other = r.forward(params, sample_start=(1 << 22) + step * spp)
        img = r.forward(params, sample_start=step * spp)
        grad = r.adjoint(params, 2.0 * (other - target) / img.numel())
        r.adam(params, grad, m, v, step, 0.01)
    wall = time.perf_counter() - t0
    t, rc = lay.texture_floats, lay.roughness_count
    err = lambda a, b: float((a - b).abs().mean())
    return {"adam_steps": steps, "wall_ms_per_iter": round(wall * 1000.0 / steps, 3),
            "param_mae_end": {"texture": err(params[:t], truth[:t]),
                              "roughness": err(params[t:t + rc], truth[t:t + rc]),
                              "emission": err(params[t + rc:], truth[t + rc:])}}


def roughness_observability(r, truth, target, lay, seeds=8):
    """Why Adam leaves the roughness where it is: the loss change from
    moving every roughness by +-0.1 from the truth (other parameters at the
    truth, common random numbers), beside the same for a 0.05 texture shift
    and a 5% emission change, and the signal-to-noise ratio of the
    roughness gradient at truth + 0.1 over `seeds` independent sample sets
    (|mean| / standard deviation, averaged over the slots)."""
    import torch
    t, rc = lay.texture_floats, lay.roughness_count
    start = 1 << 24

    def loss(p, seed=0):
        img = r.forward(p, sample_start=start + seed * r.spp)
        return float(((img.double() - target.double()) ** 2).mean())

    def moved(section, amount, scale=False):
        p = truth.clone()
        sl = {"texture": slice(0, t), "roughness": slice(t, t + rc), "emission": slice(t + rc, None)}[section]
        p[sl] = p[sl] * (1.0 + amount) if scale else (p[sl] + amount).clamp(0.01, 1.0)
        return p

    base = loss(truth)
    out = {"loss_at_truth": base,
           "roughness_minus_0.1": loss(moved("roughness", -0.1)) - base,
           "roughness_plus_0.1": loss(moved("roughness", 0.1)) - base,
           "texture_plus_0.05": loss(moved("texture", 0.05)) - base,
           "emission_times_1.05": loss(moved("emission", 0.05, scale=True)) - base}
    p = moved("roughness", 0.1)
    grads = []
    for k in range(seeds):
        img = r.forward(p, sample_start=start + (k + 1) * r.spp)
        g = r.adjoint(p, 2.0 * (img - target) / img.numel())
        grads.append(g[t:t + rc].double().cpu().numpy())
    g = np.array(grads)
    seen = np.abs(g).max(axis=0) > 0  # slots some path reaches
    snr = np.abs(g.mean(axis=0)[seen]) / np.maximum(g.std(axis=0, ddof=1)[seen], 1e-30)
    out["roughness_slots_seen"] = int(seen.sum())
    out["roughness_gradient_snr_mean"] = float(snr.mean())
    # Above the truth the loss grows with roughness: the mean gradient is positive.
    out["roughness_gradient_positive_fraction"] = float((g.mean(axis=0)[seen] > 0).mean())
    return out


def measure_terrain(spp, photons, frames, schedulers=("graph", "persistent", "regenerate")):
    """Shading-heavy workload with caustics: the terrain preset at full size
    (2,006,680 triangles, 16 x 4096^2 RGBA8 textures, 512 emissive
    triangles, a 2048 x 1024 HDR sky with a sun, dielectric crystals) with a
    spectral photon map, 1920x1080, depth 10, each scheduler timed on the
    same frame (images compared); BLASes built on the device."""
    t0 = time.perf_counter()
    host, scene = terrain_scene()
    t2 = time.perf_counter()
    info = scene.info()
    log(f"terrain: generated, built and uploaded in {t2 - t0:.2f} s "
        f"(device BLAS build {info.blas_build_milliseconds:.1f} ms)")
    camera = host.camera(WIDTH, HEIGHT)
    modes = {"graph": wpt.MODE_GRAPH, "persistent": wpt.MODE_PERSISTENT, "regenerate": wpt.MODE_REGENERATE}
    runs, first = {}, None
    photon = None
    breakdown = None
    for name, mode in modes.items():
        if name not in schedulers:
            continue
        rc = terrain_config(photons, spp)
        rc.mode = mode
        r = wpt.Renderer(scene, camera, rc)
        if photon is None:
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
        if mode == wpt.MODE_GRAPH:
            # Per-kernel device time of one more frame (events around each launch).
            k_ms, k_launches = r.profile_kernels(0, spp)
            names = ["primary", "reset", "trace", "shade", "persistent", "shadow", "resolve"]
            breakdown = {n_: round(float(k_ms[k]), 2) for k, n_ in enumerate(names) if k_launches[k]}
        r.close()
        if first is None:
            first = (img, n, st)
        runs[name] = {"ms_per_frame": round(wall * 1000.0, 2),
                      "mrays_per_s": round((st.closest_queries + st.shadow_queries) / wall / 1e6, 2),
                      "image_identical": bool(np.array_equal(img, first[0]))}
        log(f"terrain {name}: {runs[name]}")
    best = min(runs, key=lambda k: runs[k]["ms_per_frame"])
    img, n, st = first
    out = {
        "workload": f"terrain 1920x1080 {spp}spp depth10 rr3: 2,006,680 triangles (terrain + 4 rock meshes "
                    f"+ foliage + crystals + 512 lamp triangles, 162 instances), 16 x 4096^2 RGBA8 textures, "
                    f"2048x1024 HDR sky, {photons} spectral caustic photons",
        "scheduler": best,
        "mrays_per_s": runs[best]["mrays_per_s"],
        "msamples_per_s": round(WIDTH * HEIGHT * spp / (runs[best]["ms_per_frame"] / 1000.0) / 1e6, 2),
        "ms_per_frame": runs[best]["ms_per_frame"],
        "schedulers": runs,
        "staged_kernel_ms_per_frame": breakdown,
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
    ap.add_argument("--mode", choices=["regenerate", "graph", "persistent"], default="graph",
                    help="scheduler of the timed frames; the other two are timed beside it")
    ap.add_argument("--samples-per-pass", type=int, default=0,
                    help="samples per pass of the staged and persistent schedulers (0: the library's default)")
    ap.add_argument("--skip-persistent", action="store_true", help="do not time the other schedulers beside")
    ap.add_argument("--skip-cpu", action="store_true", help="do not time the CPU baseline")
    ap.add_argument("--cpu-spp", type=int, default=8, help="samples per pixel of the CPU baseline sample")
    ap.add_argument("--profile-spp", type=int, default=0,
                    help="samples rendered with per-kernel events (0: the whole frame, so launches have their real size)")
    ap.add_argument("--skip-batch", action="store_true", help="do not measure batched vs one-by-one tasks")
    ap.add_argument("--batch-tasks", type=int, default=256, help="tasks of the batching workload")
    ap.add_argument("--skip-diff", action="store_true", help="do not measure the differentiable renderer")
    ap.add_argument("--skip-terrain", action="store_true", help="do not measure the shading-heavy caustic workload")
    ap.add_argument("--skip-cornell", action="store_true", help="do not measure the small Cornell box case")
    ap.add_argument("--skip-flat", action="store_true",
                    help="do not measure the flat-triangle accelerators")
    ap.add_argument("--skip-traversal", action="store_true",
                    help="do not measure closest-hit traversal alone on the reference and primary ray sets")
    ap.add_argument("--terrain-spp", type=int, default=64)
    ap.add_argument("--terrain-photons", type=int, default=4 * 1024 * 1024)
    ap.add_argument("--terrain-schedulers", default="graph,persistent,regenerate",
                    help="comma-separated schedulers timed on the terrain")
    ap.add_argument("--diff-steps", type=int, default=200, help="Adam steps of the inverse-rendering workload")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB three times: the reference kernel of "
                         "known traffic against which the FETCH_SIZE/WRITE_SIZE counters are scaled")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is a design, "
              f"not a measured configuration", file=sys.stderr)
        sys.exit(2)
    # torch ships its own HIP runtime beside the one libwpt links; it must
    # initialise the device before libwpt's first device call.
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
    cfg = wpt.render_config(WIDTH, HEIGHT, spp=args.spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP, mode=mode,
                            samples_per_pass=args.samples_per_pass)
    renderer = wpt.Renderer(scene, camera, cfg)
    launch = renderer.launch_info()
    log(f"launch: {launch}")

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
    breakdown = None
    if mode == wpt.MODE_REGENERATE:
        trace_ms = float(np.sum(device_ms))
        trace_launches = args.steps
        prof_closest, prof_shadow = stats.closest_queries * args.steps, stats.shadow_queries * args.steps
    else:
        profile_spp = args.profile_spp or args.spp
        kernel_ms, launches = renderer.profile_kernels(0, profile_spp)
        prof_stats = renderer.stats()
        trace_slot = 2 if mode == wpt.MODE_GRAPH else 4
        trace_ms = kernel_ms[trace_slot]
        trace_launches = int(launches[trace_slot])
        prof_closest, prof_shadow = prof_stats.closest_queries, prof_stats.shadow_queries
        log(f"kernels over {profile_spp} spp: primary {kernel_ms[0]:.2f} ms, reset {kernel_ms[1]:.2f} ms, "
            f"trace {kernel_ms[2]:.2f} ms ({int(launches[2])} launches), shade {kernel_ms[3]:.2f} ms, "
            f"persistent {kernel_ms[4]:.2f} ms ({int(launches[4])} launches), shadow {kernel_ms[5]:.2f} ms "
            f"({int(launches[5])} launches)")
        names = ["primary", "reset", "trace", "shade", "persistent", "shadow", "resolve"]
        breakdown = {"spp": profile_spp,
                     "ms": {n: round(float(kernel_ms[k]), 3) for k, n in enumerate(names) if launches[k]},
                     "launches": {n: int(launches[k]) for k, n in enumerate(names) if launches[k]}}
    renderer.close()

    others = {}
    if not args.skip_persistent:
        for name, m in modes.items():
            if m == mode:
                continue
            ocfg = wpt.render_config(WIDTH, HEIGHT, spp=args.spp, max_depth=DEPTH, rr_start=RR_START, clamp=CLAMP,
                                     mode=m, samples_per_pass=args.samples_per_pass)
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
        _, work = cpu_baseline(scene, camera, threads, 1, runs=1)

    traversal = None
    if not args.skip_traversal:
        traversal = measure_traversal(scene, camera, threads)

    flat = None
    if not args.skip_flat:
        flat = measure_flat_accelerators()

    cornell = None
    if not args.skip_cornell:
        cornell = measure_cornell(threads)

    batch = None
    if not args.skip_batch:
        batch = measure_batch(args.batch_tasks, args.terrain_photons)

    diff = None
    if not args.skip_diff:
        diff = measure_diff(args.diff_steps)

    terrain = None
    if not args.skip_terrain:
        terrain = measure_terrain(args.terrain_spp, args.terrain_photons, 1, args.terrain_schedulers.split(","))

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
            "hbm_gbs_measured": None if (kernel, args.spp) not in MEASURED_TRAFFIC else
                round(MEASURED_TRAFFIC[(kernel, args.spp)] / (trace_ms / max(1, trace_launches) / 1000.0) / 1e9, 1),
            "limiter": "chains of dependent node and triangle fetches (L1 serves 90%, L2 half of the rest at "
                       "306 cycles) with a quarter of the lanes active: VALU busy 76%, the address path half busy; "
                       "the algorithmic bytes count every node visit and the measured HBM traffic is a small part",
            "counters": MEASURED_COUNTERS.get((kernel, args.spp)),
            "bytes_per_closest_query": round(per_query, 1),
            "avg_launch_ms": round(trace_ms / max(1, trace_launches), 4),
            "launches": trace_launches,
            "kernel_breakdown": breakdown,
            "occupancy": {k: launch[k] for k in ("block_size", "blocks", "blocks_per_cu", "waves_per_simd",
                                                 "registers", "scratch_bytes", "lds_bytes_per_block",
                                                 "lds_bytes_per_cu")},
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
        "traversal": traversal,
        "flat_accelerators": flat,
        "batch": batch,
        "diff": diff,
        "shading_heavy": terrain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

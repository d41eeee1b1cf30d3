"""GPU tests of the batch consumer (wpt_batch_*): uneven tasks of several
resident scenes, as staged passes over many tasks or in one persistent
launch.  Each task's sums must equal
wpt_render_task's on the same task bit for bit (the reference renders tasks
one by one, src/renderer/render_task.cu:26-114), and small tasks the
oracle's; exact-once coverage follows tests/distributed_task_tests.cpp:40-51."""
import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu

FRAME = (160, 96)
SCHEDULERS = pytest.mark.parametrize("scheduler", [wpt.BATCH_STAGED, wpt.BATCH_REGENERATE],
                                     ids=["staged", "regenerate"])


def make_scenes():
    hosts = [wpt.HostScene("cornell-box"), wpt.HostScene("sphere-field", icosphere_frequency=16),
             wpt.HostScene("prism")]
    scenes = [wpt.Scene(h.desc(), device=True, owner=h) for h in hosts]
    return hosts, scenes


def random_tasks(hosts, count, seed, max_spp=6):
    rng = np.random.default_rng(seed)
    w, h = FRAME
    tasks = []
    for _ in range(count):
        s = int(rng.integers(len(hosts)))
        tw, th = int(rng.choice([8, 16, 32, 48])), int(rng.choice([8, 16, 32]))
        x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
        start = int(rng.integers(0, 4))
        spp = int(rng.integers(1, max_spp + 1))
        depth = int(rng.choice([4, 8, 12]))
        tasks.append(wpt.batch_task(s, hosts[s].camera(w, h), w, h, x, y, tw, th, start, spp, max_depth=depth,
                                    rr_start=3, rng=int(rng.integers(2)), clamp=float(rng.choice([0.0, 10.0]))))
    return tasks


def render_one_by_one(scenes, task):
    cfg = wpt.render_config(task.width, task.height, spp=task.sample_start + task.sample_count,
                            max_depth=task.max_depth, rr_start=task.rr_start_depth, rng=task.rng_strategy,
                            clamp=task.indirect_clamp)
    r = wpt.Renderer(scenes[task.scene], task.camera, cfg)
    t = wpt.Task(0, 0, task.sample_start, task.sample_count, task.x, task.y, task.tile_width, task.tile_height)
    out = r.render_task(t)
    queries = r.stats()
    r.close()
    return out, queries


@SCHEDULERS
def test_batch_equals_tasks_one_by_one(torch_cuda, scheduler):
    hosts, scenes = make_scenes()
    tasks = random_tasks(hosts, 24, seed=7)
    batch = wpt.Batch(scenes, block_size=256, scheduler=scheduler)
    outs, lat = batch.render(tasks, latencies=True)
    st = batch.stats()
    closest = shadow = 0
    for i, t in enumerate(tasks):
        want, q = render_one_by_one(scenes, t)
        closest += q.closest_queries
        shadow += q.shadow_queries
        assert np.array_equal(outs[i], want), f"task {i}: max |diff| {np.abs(outs[i] - want).max()}"
    assert st.closest_queries == closest and st.shadow_queries == shadow
    assert st.samples == sum(t.tile_width * t.tile_height * t.sample_count for t in tasks)
    assert np.all(lat > 0) and lat.max() <= st.milliseconds * 1.05 + 0.05


@SCHEDULERS
def test_batch_small_tasks_match_oracle(torch_cuda, oracle_mod, scheduler):
    hosts, scenes = make_scenes()
    tasks = random_tasks(hosts, 6, seed=11, max_spp=3)
    outs = wpt.Batch(scenes, scheduler=scheduler).render(tasks)
    for i, t in enumerate(tasks):
        cfg = wpt.render_config(t.width, t.height, spp=t.sample_start + t.sample_count, max_depth=t.max_depth,
                                rr_start=t.rr_start_depth, rng=t.rng_strategy, clamp=t.indirect_clamp)
        fb, _ = oracle_mod.Oracle(scenes[t.scene]).render(t.camera, cfg, tile=(t.x, t.y, t.tile_width, t.tile_height),
                                                          sample_start=t.sample_start, sample_count=t.sample_count)
        assert np.array_equal(outs[i], fb[t.y:t.y + t.tile_height, t.x:t.x + t.tile_width]), f"task {i}"


@SCHEDULERS
def test_batch_covers_a_frame_exactly_once(torch_cuda, scheduler):
    # A frame cut into uneven tiles and sample ranges sums to the full-frame
    # render (same per-pixel addition order within each task's range).
    hosts, scenes = make_scenes()
    w, h = FRAME
    cam = hosts[0].camera(w, h)
    tasks = []
    for ty, th in ((0, 40), (40, 56)):
        for tx, tw in ((0, 64), (64, 96)):
            tasks.append(wpt.batch_task(0, cam, w, h, tx, ty, tw, th, 0, 4))
    outs = wpt.Batch(scenes[:1], scheduler=scheduler).render(tasks)
    frame = np.zeros((h, w, 3), np.float32)
    cover = np.zeros((h, w), int)
    for t, o in zip(tasks, outs):
        frame[t.y:t.y + t.tile_height, t.x:t.x + t.tile_width] += o
        cover[t.y:t.y + t.tile_height, t.x:t.x + t.tile_width] += 1
    assert np.all(cover == 1)
    r = wpt.Renderer(scenes[0], cam, wpt.render_config(w, h, spp=4))
    full, n = r.render_frame()
    assert n == 4 and np.array_equal(frame, full)


@SCHEDULERS
def test_batch_autotune_keeps_results(torch_cuda, scheduler):
    hosts, scenes = make_scenes()
    tasks = random_tasks(hosts, 8, seed=3, max_spp=2)
    batch = wpt.Batch(scenes, scheduler=scheduler)
    before = batch.render(tasks)
    chosen = batch.autotune(tasks)
    assert chosen.block_size in (64, 128, 256, 512) and chosen.scheduler == scheduler
    after = batch.render(tasks)
    st = batch.stats()
    if chosen.scheduler == wpt.BATCH_REGENERATE:
        assert chosen.blocks > 0 and (st.block_size, st.blocks) == (chosen.block_size, chosen.blocks)
    else:
        assert st.block_size == chosen.block_size and st.blocks > 0
    for a, b in zip(before, after):
        assert np.array_equal(a, b)


@SCHEDULERS
def test_batch_edge_cases(torch_cuda, scheduler):
    hosts, scenes = make_scenes()
    batch = wpt.Batch(scenes, scheduler=scheduler)
    assert batch.render([]) == []
    w, h = FRAME
    cam = hosts[0].camera(w, h)
    empty = wpt.batch_task(0, cam, w, h, 0, 0, 8, 8, 0, 0)
    outs = batch.render([empty, wpt.batch_task(0, cam, w, h, w - 1, h - 1, 1, 1, 0, 1)])
    assert np.all(outs[0] == 0) and outs[1].shape == (1, 1, 3)
    for bad in (wpt.batch_task(3, cam, w, h, 0, 0, 8, 8, 0, 1),          # scene index
                wpt.batch_task(0, cam, w, h, w - 4, 0, 8, 8, 0, 1),      # tile outside the frame
                wpt.batch_task(0, cam, w, h, 0, 0, 8, 8, 0, 1, max_depth=0),
                wpt.batch_task(0, cam, w, h, 0, 0, 8, 8, 0, 1, clamp=-1.0)):
        with pytest.raises(wpt.WptError):
            batch.render([bad])


def test_staged_batch_small_passes(torch_cuda):
    # Passes of 3000 paths split tasks of up to 1536 pixels into pieces of a
    # few samples across passes; the sums must not depend on the pass size.
    hosts, scenes = make_scenes()
    tasks = random_tasks(hosts, 16, seed=21, max_spp=5)
    want = wpt.Batch(scenes, scheduler=wpt.BATCH_STAGED).render(tasks)
    small = wpt.Batch(scenes, scheduler=wpt.BATCH_STAGED, pass_paths=3000)
    outs = small.render(tasks)
    for i, (a, b) in enumerate(zip(outs, want)):
        assert np.array_equal(a, b), f"task {i}"
    assert small.autotune(tasks).pass_paths == 3000
    w, h = FRAME
    big = wpt.batch_task(0, hosts[0].camera(w, h), w, h, 0, 0, 64, 64, 0, 1)  # 4096 pixels > 3000
    with pytest.raises(wpt.WptError):
        small.render([big])
    with pytest.raises(wpt.WptError):
        wpt.Batch(scenes, scheduler=wpt.BATCH_STAGED, pass_paths=(1 << 26) + 1)

"""Checkpoints and the task ledger on the GPU: a render resumed from a
checkpoint equals the uninterrupted render bit for bit (renderer entry
points and the pathtracer's --resume, render_driver.cu:359-383,466-474),
and the --coordinator role's task ledger sums the same tile results as
wpt_render_task, task by task (coordinator_app.cpp:56-180)."""
import os
import re
import subprocess

import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "wavefront-path-tracer_amd", "bin", "pathtracer")


@pytest.mark.parametrize("mode", [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE])
def test_loaded_framebuffer_resumes_exactly(torch_cuda, mode):
    host = wpt.HostScene("cornell-box")
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    cam = host.camera(48, 40)
    cfg = wpt.render_config(48, 40, spp=5, max_depth=8, rr_start=3, mode=mode)
    full, n = wpt.Renderer(scene, cam, cfg).render_frame()
    first = wpt.Renderer(scene, cam, cfg)
    first.reset()
    first.render_samples(0, 2)
    partial, k = first.framebuffer()
    assert (n, k) == (5, 2)
    resumed = wpt.Renderer(scene, cam, cfg)
    resumed.load_framebuffer(partial, k)
    resumed.render_samples(2, 3)
    out, m = resumed.framebuffer()
    assert m == 5 and np.array_equal(out, full)


def run(args, cwd):
    out = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120, cwd=cwd)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def sha(stdout):
    return re.search(r"Raw framebuffer SHA-256: ([0-9a-f]{64})", stdout).group(1)


def test_pathtracer_checkpoints_and_resumes(torch_cuda, tmp_path):
    args = ["--scene", "cornell-box", "--resolution", "64x48", "--spp", "8", "--max-depth", "6"]
    full = run(args + ["--output", "a.ppm"], tmp_path)
    # 25 % steps of 8 samples: checkpoints after samples 2, 4 and 6.
    assert [int(x) for x in re.findall(r"Queued checkpoint at (\d+) of 8", full)] == [2, 4, 6]
    _, spp = wpt.read_render_checkpoint(tmp_path / "render.checkpoint", 64, 48,
                                        wpt.parse_command_line(args + ["--output", "a.ppm"])[3])
    assert spp == 6
    resumed = run(args + ["--output", "b.ppm", "--resume"], tmp_path)
    assert "Resuming from 6 completed samples" in resumed
    assert sha(resumed) == sha(full)
    assert (tmp_path / "a.ppm").read_bytes() == (tmp_path / "b.ppm").read_bytes()
    # Another configuration does not accept the file.
    out = subprocess.run([EXE] + args[:-2] + ["--max-depth", "7", "--resume"], capture_output=True, text=True,
                         timeout=120, cwd=tmp_path)
    assert out.returncode != 0 and "fingerprint" in out.stderr


def test_coordinator_role_sums_task_results(torch_cuda, tmp_path):
    args = ["--scene", "cornell-box", "--coordinator", "--resolution", "64x48", "--tile", "32x32", "--spp", "6",
            "--samples-per-task", "4", "--max-depth", "6", "--distributed-checkpoint", "job.ledger",
            "--output", "c.ppm"]
    stdout = run(args, tmp_path)
    assert "Ledger render complete" in stdout
    status, rc, _, fp = wpt.parse_command_line(args)
    assert status == wpt.OK
    ledger = wpt.Ledger(fp, 64, 48, 32, 32, 6, 4)
    ledger.restore_checkpoint(tmp_path / "job.ledger")
    assert ledger.counts() == (8, 0, 8)  # 4 tiles x sample ranges [0,4) and [4,6)
    rad, cnt = ledger.image()
    assert np.all(cnt == 6)
    host = wpt.HostScene("cornell-box")
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    renderer = wpt.Renderer(scene, host.camera(64, 48), rc, fingerprint=fp)
    want = np.zeros((48, 64, 3), np.float32)
    for t in ledger.tasks:  # one local worker leases in task-table order
        want[t.y:t.y + t.height, t.x:t.x + t.width] += renderer.render_task(t)
    assert np.array_equal(rad, want)
    again = run(args + ["--resume"], tmp_path)
    assert "8/8 tasks completed" in again and sha(again) == sha(stdout)

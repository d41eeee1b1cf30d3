"""The task ledger and the checkpoint files on the CPU: ports of the
reference's coordinator, checkpoint-writer and checkpoint-format tests
(tests/distributed_coordinator_tests.cpp, distributed_checkpoint_writer_tests.cpp,
checkpoint_tests.cu), and the byte layouts of src/renderer/render_checkpoint.cpp:17-64
and src/distributed/checkpoint.cpp:9-84 read back independently with struct."""
import struct

import numpy as np
import pytest

import wpt

FP = 0x13579BDF2468ACE0
# The reference's FNV-1a starts from 1469598103934665603
# (src/distributed/checkpoint.cpp:46, render_fingerprint.cpp:78), one digit
# short of the published 64-bit offset basis 14695981039346656037; files
# must use the reference's value to stay compatible.
REFERENCE_BASIS = 1469598103934665603
PUBLISHED_BASIS = 14695981039346656037


def fnv1a(data: bytes, h=REFERENCE_BASIS):
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv1a_known_vectors():
    # Published FNV-1a 64 vectors pin the helper's step (xor, then multiply
    # by the 64-bit prime); the files below use the reference's basis.
    assert fnv1a(b"", PUBLISHED_BASIS) == 0xCBF29CE484222325
    assert fnv1a(b"a", PUBLISHED_BASIS) == 0xAF63DC4C8601EC8C
    assert fnv1a(b"foobar", PUBLISHED_BASIS) == 0x85944171F73967E8


def tile_sums(task, value):
    r = np.empty((task.height, task.width, 3), np.float32)
    r[..., 0], r[..., 1], r[..., 2] = value, value + 1, value + 2
    return r


def test_coordinator_reference_sequence(tmp_path):
    # distributed_coordinator_tests.cpp:29-103, step for step.
    led = wpt.Ledger(FP, 4, 2, 2, 2, 4, 2)
    assert led.counts() == (4, 4, 0)
    a, _ = led.lease_next("worker-a", 0, 100)
    b, _ = led.lease_next("worker-b", 0, 100)
    assert (a.tile_ordinal, a.sample_start) != (b.tile_ordinal, b.sample_start)
    ra = tile_sums(a, 4.0)
    assert led.commit(a, a.sample_count, ra) == wpt.LEDGER_ACCEPTED
    assert led.commit(a, a.sample_count, ra) == wpt.LEDGER_DUPLICATE
    assert led.counts()[2] == 1
    rad, cnt = led.image()
    assert cnt[0, 0] == 2 and rad[0, 0, 0] == 4.0
    path = tmp_path / "ledger.checkpoint"
    led.write_checkpoint(path, 7)
    restored = wpt.Ledger(FP, 4, 2, 2, 2, 4, 2)
    assert restored.restore_checkpoint(path) == 7
    assert restored.state(a) == wpt.TASK_COMPLETED
    assert restored.state(b) == wpt.TASK_PENDING
    assert restored.counts()[1] == 3
    assert restored.commit(a, a.sample_count, ra) == wpt.LEDGER_DUPLICATE
    assert led.requeue_expired(99) == 0
    assert led.requeue_expired(100) == 1
    assert led.counts()[1] == 3
    bad = tile_sums(b, 1.0).reshape(-1)[:-3]
    assert led.commit(b, b.sample_count, bad) == wpt.LEDGER_REJECTED


def test_ledger_rejects_foreign_and_malformed_results():
    led = wpt.Ledger(FP, 8, 8, 4, 4, 4, 4)
    t, _ = led.lease_next("w", 0, 10)
    good = tile_sums(t, 1.0)
    other = wpt.Task(FP + 1, t.tile_ordinal, t.sample_start, t.sample_count, t.x, t.y, t.width, t.height)
    assert led.commit(other, t.sample_count, good) == wpt.LEDGER_REJECTED
    moved = wpt.Task(FP, t.tile_ordinal, t.sample_start, t.sample_count, t.x + 4, t.y, t.width, t.height)
    assert led.commit(moved, t.sample_count, good) == wpt.LEDGER_REJECTED
    assert led.commit(t, t.sample_count + 1, good) == wpt.LEDGER_REJECTED
    unknown = wpt.Task(FP, 99, 0, 4, 0, 0, 4, 4)
    assert led.commit(unknown, 4, good) == wpt.LEDGER_REJECTED
    assert led.commit(t, t.sample_count, good) == wpt.LEDGER_ACCEPTED
    # A pending (never leased) task may be committed; it leaves the queue.
    p = [x for x in led.tasks if led.state(x) == wpt.TASK_PENDING][0]
    assert led.commit(p, p.sample_count, tile_sums(p, 0.5)) == wpt.LEDGER_ACCEPTED
    assert led.counts() == (4, 2, 2)


def test_leases_renew_release_and_expire():
    led = wpt.Ledger(FP, 8, 4, 4, 4, 2, 1)  # 2 tiles x 2 sample ranges
    got = [led.lease_next("a", 0, 50)[0], led.lease_next("b", 0, 50)[0], led.lease_next("a", 0, 50)[0]]
    assert led.lease_next("", 0, 50) is None
    assert led.renew("a", 40, 50) == 2
    assert led.requeue_expired(60) == 1        # b's lease ended at 50
    assert led.release("a") == 2
    assert led.counts()[1] == 4
    assert all(led.state(t) == wpt.TASK_PENDING for t in got)
    order = [led.lease_next("c", 100, 10)[0] for _ in range(4)]
    assert led.lease_next("c", 100, 10) is None
    # The never-leased task first, then requeued tasks in the order they
    # returned: b's expired lease, then a's released ones in task order.
    assert [(t.tile_ordinal, t.sample_start) for t in order] == [(1, 1), (0, 1), (0, 0), (1, 0)]


def test_sums_follow_task_order_and_sample_counts():
    # Every pixel accumulates spp samples over its tile's tasks, whatever
    # the commit order; sums are float32 adds in commit order.
    w, h, spp, per = 6, 5, 7, 3
    led = wpt.Ledger(FP, w, h, 4, 4, spp, per)
    rng = np.random.default_rng(3)
    tasks = list(led.tasks)
    rng.shuffle(tasks)
    want = np.zeros((h, w, 3), np.float32)
    for t in tasks:
        r = rng.uniform(0, 2, (t.height, t.width, 3)).astype(np.float32)
        assert led.commit(t, t.sample_count, r) == wpt.LEDGER_ACCEPTED
        want[t.y:t.y + t.height, t.x:t.x + t.width] += r
    rad, cnt = led.image()
    assert np.all(cnt == spp)
    assert np.array_equal(rad, want)


def test_ledger_checkpoint_layout_and_corruption(tmp_path):
    led = wpt.Ledger(FP, 4, 2, 2, 2, 4, 2)
    t, _ = led.lease_next("w", 0, 5)
    led.commit(t, t.sample_count, tile_sums(t, 3.0))
    path = tmp_path / "d.checkpoint"
    led.write_checkpoint(path, 11)
    raw = path.read_bytes()
    head = struct.unpack("<8sII Q 6I QQQQQ", raw[:88])
    magic, version, size, fp, w, h, tw, th, spp, per, seq, pixels, ntasks, payload, checksum = head
    assert (magic, version, size, fp) == (b"PTDIST1\0", 1, 88, FP)
    assert (w, h, tw, th, spp, per, seq, pixels, ntasks) == (4, 2, 2, 2, 4, 2, 11, 8, 4)
    assert payload == 8 * 12 + 8 * 4 + 4 == len(raw) - 88
    assert checksum == fnv1a(raw[88:])
    rad = np.frombuffer(raw[88:88 + 96], np.float32).reshape(2, 4, 3)
    cnt = np.frombuffer(raw[88 + 96:88 + 128], np.uint32).reshape(2, 4)
    done = np.frombuffer(raw[88 + 128:], np.uint8)
    r2, c2 = led.image()
    assert np.array_equal(rad, r2) and np.array_equal(cnt, c2)
    assert done.tolist() == [1 if (x.tile_ordinal, x.sample_start) == (t.tile_ordinal, t.sample_start) else 0
                             for x in led.tasks]
    assert not (tmp_path / "d.checkpoint.tmp").exists()

    def rejected(data, job=(4, 2, 2, 2, 4, 2)):
        bad = tmp_path / "bad.checkpoint"
        bad.write_bytes(data)
        with pytest.raises(wpt.WptError):
            wpt.Ledger(FP, *job).restore_checkpoint(bad)

    flipped = bytearray(raw)
    flipped[100] ^= 1
    rejected(bytes(flipped))                        # checksum
    rejected(raw[:-1])                              # truncated
    rejected(raw + b"\0")                           # trailing data
    rejected(b"PTDIST2\0" + raw[8:])                # magic
    bitmap = bytearray(raw)
    bitmap[-1] = 2
    bitmap[80:88] = struct.pack("<Q", fnv1a(bytes(bitmap[88:])))
    rejected(bytes(bitmap))                         # completion byte > 1
    rejected(raw, job=(4, 2, 2, 2, 8, 2))           # another job
    with pytest.raises(wpt.WptError):
        wpt.Ledger(FP, 4, 2, 2, 2, 4, 2).restore_checkpoint(tmp_path / "missing")


def test_checkpoint_writer_keeps_the_newest(tmp_path):
    # distributed_checkpoint_writer_tests.cpp: flush writes the newest
    # queued snapshot; closing writes what is still queued.
    path = tmp_path / "w.checkpoint"
    led = wpt.Ledger(FP, 1, 1, 1, 1, 4, 1)
    writer = wpt.LedgerWriter(path)
    writer.enqueue(led, 1)
    t, _ = led.lease_next("w", 0, 5)
    led.commit(t, 1, np.array([2.0, 3.0, 4.0], np.float32))
    writer.enqueue(led, 2)
    writer.flush()
    other = wpt.Ledger(FP, 1, 1, 1, 1, 4, 1)
    assert other.restore_checkpoint(path) == 2
    assert other.image()[0][0, 0, 0] == 2.0
    writer.enqueue(led, 3)
    writer.close()
    assert wpt.Ledger(FP, 1, 1, 1, 1, 4, 1).restore_checkpoint(path) == 3
    bad = wpt.LedgerWriter(tmp_path / "no-such-dir" / "x.checkpoint")
    bad.enqueue(led, 1)
    with pytest.raises(wpt.WptError):
        bad.flush()
    with pytest.raises(wpt.WptError):
        bad.close()


def test_render_checkpoint_layout_and_rejections(tmp_path):
    # render_checkpoint.cpp:17-64,163-195: 32-byte header, then the sums.
    rgb = np.random.default_rng(5).uniform(0, 9, (3, 5, 3)).astype(np.float32)
    path = tmp_path / "render.checkpoint"
    wpt.write_render_checkpoint(path, rgb, 17, FP)
    raw = path.read_bytes()
    assert struct.unpack("<8sIIIIQ", raw[:32]) == (b"PTCHKPT1", 3, 5, 3, 17, FP)
    assert np.array_equal(np.frombuffer(raw[32:], np.float32).reshape(3, 5, 3), rgb)
    back, n = wpt.read_render_checkpoint(path, 5, 3, FP)
    assert n == 17 and np.array_equal(back, rgb)
    for args in [(6, 3, FP), (5, 3, FP ^ 1)]:
        with pytest.raises(wpt.WptError):
            wpt.read_render_checkpoint(path, *args)
    (tmp_path / "short").write_bytes(raw[:-4])
    (tmp_path / "v2").write_bytes(raw[:8] + struct.pack("<I", 2) + raw[12:])
    for name in ("short", "v2", "missing"):
        with pytest.raises(wpt.WptError):
            wpt.read_render_checkpoint(tmp_path / name, 5, 3, FP)

"""GPU queue compaction (wpt_compact): the reference's two patterns and the
bounded-capacity sentinel check (tests/queue_compaction_tests.cu:13-95), plus
the stronger order-preserving contract: output == the oracle's sequential
stable filter, bit for bit, at sizes up to 2^22 items."""
import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu

SENTINEL = np.uint32(0xCDCDCDCD)


def run(torch, items, flags, capacity):
    d_items = torch.from_numpy(items.astype(np.uint32).view(np.int32)).cuda()
    d_flags = torch.from_numpy(flags.astype(np.uint8)).cuda()
    out, count = wpt.compact(d_items, d_flags, capacity)
    return out.cpu().numpy().view(np.uint32), count


@pytest.mark.parametrize("pattern,capacity", [(0, 64), (1, 7)])
def test_reference_patterns(torch_cuda, oracle_mod, pattern, capacity):
    idx = np.arange(64, dtype=np.uint32)
    flags = (idx % 3 != 0) if pattern == 0 else (idx < 47)
    out, count = run(torch_cuda, idx, flags, capacity)
    assert count == (42 if pattern == 0 else 47)
    stored = out[:min(count, capacity)]
    assert len(set(stored.tolist())) == len(stored)
    assert np.all(flags[stored])
    want, want_count = oracle_mod.compact(idx, flags, capacity)
    assert want_count == count and np.array_equal(stored, want)
    assert np.all(out[min(count, capacity):] == SENTINEL)


@pytest.mark.parametrize("n,density", [(1, 1.0), (255, 0.5), (257, 0.01), (100_000, 0.3), (1 << 22, 0.7), (1 << 20, 0.0)])
def test_order_preserving_against_oracle(torch_cuda, oracle_mod, n, density):
    rng = np.random.default_rng(n)
    items = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    flags = rng.random(n) < density
    out, count = run(torch_cuda, items, flags, n)
    want, want_count = oracle_mod.compact(items, flags, n)
    assert count == want_count
    assert np.array_equal(out[:count], want)

"""wpt_compact_noalloc, the compaction entry point that allocates nothing:
through the C ABI its output and count equal wpt_compact's bit for bit, on
two calls in a row (the second with a smaller input, so the fixed scratch
holds the first call's state), with bounded capacity, and it rejects inputs
beyond its fixed scratch before touching the device."""
import ctypes as C

import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu

FILL = 0xCDCDCDCD - (1 << 32)


def noalloc(torch, d_items, d_flags, capacity, stream=None):
    n = d_items.numel()
    out = torch.full((max(capacity, 1),), FILL, dtype=torch.int32, device=d_items.device)
    count = torch.full((1,), -1, dtype=torch.int32, device=d_items.device)
    s = stream.cuda_stream if stream is not None else 0
    wpt.check(wpt.library().wpt_compact_noalloc(C.c_void_p(d_items.data_ptr()), C.c_void_p(d_flags.data_ptr()), n,
                                                C.c_void_p(out.data_ptr()), capacity, C.c_void_p(count.data_ptr()),
                                                C.c_void_p(s)), "wpt_compact_noalloc")
    if stream is not None:
        stream.synchronize()
    else:
        torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32), int(count.item())


def inputs(torch, n, density, seed):
    rng = np.random.default_rng(seed)
    items = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    flags = rng.random(n) < density
    return (torch.from_numpy(items.view(np.int32)).cuda(), torch.from_numpy(flags.astype(np.uint8)).cuda())


def test_matches_wpt_compact_on_consecutive_calls(torch_cuda):
    torch = torch_cuda
    for n, density, capacity, seed in [(3_000_017, 0.6, 3_000_017, 1), (70_001, 0.3, 70_001, 2),
                                       (1 << 24, 0.5, 1 << 24, 3), (4_099, 0.9, 1_000, 4), (0, 0.5, 8, 5)]:
        d_items, d_flags = inputs(torch, n, density, seed)
        got, got_count = noalloc(torch, d_items, d_flags, capacity)
        want, want_count = wpt.compact(d_items, d_flags, capacity)
        want = want.cpu().numpy().view(np.uint32)
        assert got_count == want_count, (n, got_count, want_count)
        assert np.array_equal(got, want), n


def test_on_a_side_stream(torch_cuda):
    torch = torch_cuda
    d_items, d_flags = inputs(torch, 500_000, 0.4, 7)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        got, got_count = noalloc(torch, d_items, d_flags, 500_000, stream)
    want, want_count = wpt.compact(d_items, d_flags, 500_000)
    assert got_count == want_count and np.array_equal(got, want.cpu().numpy().view(np.uint32))


def test_rejects_inputs_beyond_the_fixed_scratch(torch_cuda):
    torch = torch_cuda
    d = torch.zeros(16, dtype=torch.int32, device="cuda")
    status = wpt.library().wpt_compact_noalloc(C.c_void_p(d.data_ptr()), C.c_void_p(d.data_ptr()), (1 << 24) + 1,
                                               C.c_void_p(d.data_ptr()), 0, C.c_void_p(d.data_ptr()), None)
    assert status == 1  # WPT_ERROR_INVALID_ARGUMENT
    assert b"16777216" in wpt.library().wpt_last_error()

"""The allocation-free ragged entry point onefft_ragged_execute_items, through the C ABI, against
onefft_ragged_execute (the planned entry point) on the same items: bit-identical outputs, two calls
in a row with the second on a smaller input, every library allocation made to fail meanwhile
(ONEFFT_FAIL_ALLOC=0) and no device memory taken; invalid items and too many items reported.

The allocation-free entry point runs every item in the persistent kernel (it sees the items only in
device memory, so it cannot group the large ones on the host), so the planned execution it is
compared with bit for bit is made with the batched groups off (ONEFFT_RAGGED_BATCHED_MIN=21); the
default planned execution, whose large items run through the multi-pass kernels, is checked
against it within the FP64 bar."""
import ctypes
import os

import numpy as np
import pytest
import torch

import onefft_amd as F

pytestmark = pytest.mark.gpu

ITEM = np.dtype([("offset", np.uint64), ("log2n", np.uint32), ("inverse", np.uint32)])


def device_items(items):
    arr = np.array([tuple(it) for it in items], dtype=ITEM)
    return torch.from_numpy(arr.view(np.uint8)).cuda()


class FailEveryAlloc:
    def __enter__(self):
        os.environ["ONEFFT_FAIL_ALLOC"] = "0/no-alloc-test"

    def __exit__(self, *exc):
        os.environ.pop("ONEFFT_FAIL_ALLOC", None)


def stream():
    return torch.cuda.current_stream().cuda_stream


def every_size_items():
    """Every log2 n 0..20 in both directions, three of each, with gaps between items."""
    items, off = [], 0
    rng = np.random.default_rng(12)
    for rep in range(3):
        for l in range(0, 21):
            for inv in (0, 1):
                off += int(rng.integers(0, 5))
                items.append((off, l, inv))
                off += 1 << l
    rng.shuffle(items)
    return items


def planned(items, x, groups=False):
    """onefft_ragged_execute of a plan of `items`; groups=False: every item in the persistent kernel."""
    y = torch.zeros_like(x)
    if not groups:
        os.environ["ONEFFT_RAGGED_BATCHED_MIN"] = "21"
    try:
        rp = F.RaggedPlan(items)
    finally:
        os.environ.pop("ONEFFT_RAGGED_BATCHED_MIN", None)
    rp.execute(x.data_ptr(), y.data_ptr(), stream())
    rp.wait(stream())
    rp.close()
    return y


def run_items(L, items, x, y, status):
    d = device_items(items)
    extent = x.numel()
    rc = L.onefft_ragged_execute_items(d.data_ptr(), len(items), extent, x.data_ptr(), y.data_ptr(),
                                       status.data_ptr(), stream())
    torch.cuda.synchronize()
    return rc


def test_no_alloc_matches_planned_execution():
    L = F.lib()
    cases = [F.ragged_items(3, 200000, max_bytes=256 << 20), every_size_items(), F.ragged_items(9, 2000, max_bytes=24 << 20)]
    extents = [max(o + (1 << l) for o, l, _ in items) for items in cases]
    assert extents[2] < extents[0], "the last call's input must be smaller than the first's"
    # The first call on a GPU loads the library's device buffer (the code object's device globals,
    # ~5.2 MiB, loaded by the HIP runtime on the module's first use) and uploads the tables; it too
    # makes no library allocation. Later calls take no device memory at all.
    x0 = torch.randn(1 << 12, dtype=torch.complex64, device="cuda")
    y0 = torch.zeros_like(x0)
    status0 = torch.full((1,), 99, dtype=torch.int32, device="cuda")
    with FailEveryAlloc():
        assert run_items(L, [(0, 12, 0)], x0, y0, status0) == 0, F.last_error()
    assert int(status0.item()) == 0
    for items, extent in zip(cases, extents):
        print(f"no-alloc ragged: {len(items)} items, extent {extent}", flush=True)
        x = torch.randn(extent, dtype=torch.complex64, device="cuda")
        want = planned(items, x)
        y = torch.full_like(x, complex(-1.0, -1.0))
        status = torch.full((1,), 99, dtype=torch.int32, device="cuda")
        d = device_items(items)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        with FailEveryAlloc():
            rc = L.onefft_ragged_execute_items(d.data_ptr(), len(items), extent, x.data_ptr(), y.data_ptr(),
                                               status.data_ptr(), stream())
            torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert rc == 0, F.last_error()
        assert int(status.item()) == 0
        assert free0 - free1 < (2 << 20), f"device memory taken: {(free0 - free1) >> 10} KiB"
        # every item equals the planned path's result; the gaps between items stay untouched
        covered = torch.zeros(extent, dtype=torch.bool, device="cuda")
        for o, l, _ in items:
            covered[o:o + (1 << l)] = True
        assert torch.equal(y[covered].view(torch.float32), want[covered].view(torch.float32)), "differs from onefft_ragged_execute"
        # the default planned execution (large items through the multi-pass kernels) agrees
        grouped = planned(items, x, groups=True)
        for o, l, _ in items[:: max(1, len(items) // 200)]:
            a, b = y[o:o + (1 << l)], grouped[o:o + (1 << l)]
            d = float((a - b).abs().pow(2).sum().sqrt() / b.abs().pow(2).sum().sqrt().clamp_min(1e-30))
            assert d < 2 * (4e-7 + 1e-7 * l), (o, l, d)
        assert bool((y[~covered] == complex(-1.0, -1.0)).all()), "wrote outside the items"


def test_no_alloc_reports_invalid_items():
    L = F.lib()
    extent = 1 << 16
    x = torch.randn(extent, dtype=torch.complex64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    for bad in ([(0, 10, 0), (1024, 21, 0)],          # size above 2^20
                [(0, 10, 0), (extent - 512, 10, 1)],  # beyond the extent
                [(0, 4, 2)]):                          # direction not 0/1
        y = torch.zeros_like(x)
        assert run_items(L, bad, x, y, status) == 0, F.last_error()
        assert int(status.item()) == 2, bad
        assert bool((y == 0).all()), "an invalid call transformed something"
    # more than 262144 tasks: 2049 items of 2^20 points make 128 tasks each (nothing runs; the
    # planner does not check overlaps, so they may all lie at offset 0 of a 2^20-point buffer)
    big = torch.randn(1 << 20, dtype=torch.complex64, device="cuda")
    y = torch.zeros_like(big)
    assert run_items(L, [(0, 20, 0)] * 2049, big, y, status) == 0, F.last_error()
    assert int(status.item()) == 3
    assert bool((y == 0).all())
    # a valid call afterwards works
    y = torch.zeros_like(x)
    good = [(0, 10, 0), (1024, 12, 1), (8192, 14, 0)]
    assert run_items(L, good, x, y, status) == 0 and int(status.item()) == 0
    want = planned(good, x)
    assert torch.equal(y[:8192 + (1 << 14)].view(torch.float32), want[:8192 + (1 << 14)].view(torch.float32))


def test_no_alloc_rejects_too_many_items_and_bad_arguments():
    L = F.lib()
    x = torch.zeros(1 << 20, dtype=torch.complex64, device="cuda")
    y = torch.zeros_like(x)
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    d = device_items([(0, 4, 0)])
    s = stream()
    assert L.onefft_ragged_execute_items(d.data_ptr(), (1 << 17) + 1, x.numel(), x.data_ptr(), y.data_ptr(),
                                         status.data_ptr(), s) == -1
    assert "131072" in F.last_error()
    assert L.onefft_ragged_execute_items(d.data_ptr(), 1, x.numel(), x.data_ptr(), x.data_ptr(),
                                         status.data_ptr(), s) == -1  # overlap
    assert L.onefft_ragged_execute_items(d.data_ptr(), 1, x.numel(), x.data_ptr() + 4, y.data_ptr(),
                                         status.data_ptr(), s) == -1  # misaligned
    assert L.onefft_ragged_execute_items(d.data_ptr(), 1, x.numel(), x.data_ptr(), y.data_ptr(), None, s) == -1
    torch.cuda.synchronize()

"""The drop-in path's staging contract and its error paths, through the C ABI.

* After onefft_cache_prepare(n) / onefft_plan_create(n) on one thread, onefft_forward/inverse and
  onefft_plan_forward/inverse of that size allocate nothing on any thread (the reference's promise,
  README.md:35-41, include/onefft.h:66-71,102-104): fresh threads transform pageable, pinned and
  device buffers while every library allocation is made to fail (ONEFFT_FAIL_ALLOC=0), and every
  result is bit-identical to the device path.
* A call with device-memory arguments orders after work queued on the default stream.
* Calls of one LDS pass return on the kernel's completion word, or on the stream when it is late.
* Allocation failure (ONEFFT_FAIL_ALLOC=k, the k-th library allocation fails, swept over every k):
  onefft_plan_create returns NULL, onefft_forward and onefft_cache_prepare return -1 with `out`
  untouched, onefft_ragged_plan_create and graph creation return NULL, a host call whose staging
  cannot grow leaves `out` untouched with a message; normal calls work afterwards and device memory
  does not leak (reference include/onefft.h:73-75,94-97).
"""
import ctypes
import os
import threading

import numpy as np
import pytest
import torch

import onefft_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GUARD = 64
CANARY = np.float32(-7777.0)


class Guarded:
    """A pageable float buffer with canaries on both sides, `misalign` floats off 64-byte alignment."""

    def __init__(self, count, misalign):
        self.base = np.full(count + 2 * GUARD + 16 + misalign, CANARY, np.float32)
        off = (-self.base.ctypes.data // 4) % 16
        self.start = off + GUARD + misalign
        self.count = count
        self.data = self.base[self.start: self.start + count]

    @property
    def ptr(self):
        return self.data.ctypes.data

    def intact(self):
        return (self.base[: self.start] == CANARY).all() and (
            self.base[self.start + self.count:] == CANARY).all()


class FailAlloc:
    """Sets ONEFFT_FAIL_ALLOC for the block (a unique value, so the library's count restarts)."""

    serial = 0

    def __init__(self, k):
        FailAlloc.serial += 1
        self.value = f"{k}/{FailAlloc.serial}"

    def __enter__(self):
        os.environ["ONEFFT_FAIL_ALLOC"] = self.value

    def __exit__(self, *exc):
        os.environ.pop("ONEFFT_FAIL_ALLOC", None)


def device_reference(p, x, n, inverse):
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.empty_like(d_in)
    p.batch_ptr(d_in.data_ptr(), d_out.data_ptr(), 1, n, n, inverse, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


@pytest.mark.parametrize("L", [0, 5, 10, 14])
def test_completion_word_and_its_fallback_bitwise(monkeypatch, L):
    """Calls of one LDS pass on host (pageable, pinned) and device buffers return once the kernel's
    completion word arrives; with ONEFFT_POLL_US=0 they wait on the stream instead (the fallback
    when the word is late). Both give the device path's bits, many calls in a row (the word's
    sequence number advances), and the output lands only in `out`."""
    n = 1 << L
    p = F.Plan(n)
    for inverse in (False, True):
        x = O.uniform(40 + L, n)
        want = device_reference(p, x, n, inverse)
        call = p.inverse_ptr if inverse else p.forward_ptr
        for poll in (None, "0"):
            if poll is None:
                monkeypatch.delenv("ONEFFT_POLL_US", raising=False)
            else:
                monkeypatch.setenv("ONEFFT_POLL_US", poll)
            for rep in range(3):
                g_in, g_out = Guarded(2 * n, 1), Guarded(2 * n, 3)
                g_in.data[:] = x
                call(g_in.ptr, g_out.ptr)
                assert np.array_equal(g_out.data, want), (L, inverse, poll, rep, "pageable")
                assert g_in.intact() and g_out.intact() and np.array_equal(g_in.data, x)
                t_in = torch.from_numpy(x).pin_memory()
                t_out = torch.full((2 * n,), float(CANARY)).pin_memory()
                call(t_in.data_ptr(), t_out.data_ptr())
                assert np.array_equal(t_out.numpy(), want), (L, inverse, poll, rep, "pinned")
                d_in = torch.from_numpy(x).cuda()
                d_out = torch.empty_like(d_in)
                call(d_in.data_ptr(), d_out.data_ptr())
                assert np.array_equal(d_out.cpu().numpy(), want), (L, inverse, poll, rep, "device")
    monkeypatch.delenv("ONEFFT_POLL_US", raising=False)
    p.close()


def test_prepare_then_fresh_threads_allocate_nothing():
    L = F.lib()
    sizes = [1024, 1 << 16, 1 << 22]
    plans, inputs, want = {}, {}, {}
    for n in sizes:
        assert L.onefft_cache_prepare(n) == 0, F.last_error()
        plans[n] = F.Plan(n)  # onefft_plan_create sizes the pool for n as well
        inputs[n] = O.uniform(500 + n.bit_length(), n)
        for inv in (False, True):
            want[n, inv] = device_reference(plans[n], inputs[n], n, inv)
    threads_n = 4
    # every thread's buffers exist before the switch: pageable (misaligned), pinned, device
    # (aligned and 4 bytes off)
    bufs = []
    for t in range(threads_n):
        b = {}
        for n in sizes:
            b[n] = {
                "src": Guarded(2 * n, 1), "dst": Guarded(2 * n, 3),
                "pin_in": torch.from_numpy(inputs[n]).pin_memory(),
                "pin_out": torch.zeros(2 * n, dtype=torch.float32).pin_memory(),
                "dev_in": torch.from_numpy(inputs[n]).cuda(),
                "dev_out": torch.zeros(2 * n, dtype=torch.float32, device="cuda"),
                "dev_in_off": torch.zeros(2 * n + 2, dtype=torch.float32, device="cuda"),
                "dev_out_off": torch.zeros(2 * n + 2, dtype=torch.float32, device="cuda"),
            }
            b[n]["src"].data[:] = inputs[n]
            b[n]["dev_in_off"][1:1 + 2 * n].copy_(b[n]["dev_in"])
        bufs.append(b)
    torch.cuda.synchronize()
    errors = []

    def worker(t):
        try:
            for rep in range(3):
                for n in sizes:
                    b = bufs[t][n]
                    for inv in (False, True):
                        w = want[n, inv]
                        cache_call = L.onefft_inverse if inv else L.onefft_forward
                        plan_call = L.onefft_plan_inverse if inv else L.onefft_plan_forward
                        for via_plan in (False, True):
                            def call(src, dst):
                                if via_plan:
                                    plan_call(plans[n].handle, src, dst)
                                    return 0 if not F.last_error() else -1
                                return cache_call(src, dst, n)

                            b["dst"].data[:] = 0
                            assert call(b["src"].ptr, b["dst"].ptr) == 0, ("pageable", n, F.last_error())
                            assert np.array_equal(b["dst"].data, w) and b["dst"].intact(), ("pageable", n, inv)
                            b["pin_out"].zero_()
                            assert call(b["pin_in"].data_ptr(), b["pin_out"].data_ptr()) == 0, ("pinned", n)
                            assert np.array_equal(b["pin_out"].numpy(), w), ("pinned", n, inv)
                            assert call(b["dev_in"].data_ptr(), b["dev_out"].data_ptr()) == 0, ("device", n)
                            assert np.array_equal(b["dev_out"].cpu().numpy(), w), ("device", n, inv)
                            assert call(b["dev_in_off"].data_ptr() + 4, b["dev_out_off"].data_ptr() + 4) == 0, \
                                ("device misaligned", n)
                            assert np.array_equal(b["dev_out_off"][1:1 + 2 * n].cpu().numpy(), w), \
                                ("device misaligned", n, inv)
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    print("4 fresh threads, every library allocation failing", flush=True)
    with FailAlloc(0):
        ts = [threading.Thread(target=worker, args=(t,)) for t in range(threads_n)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    assert not errors, errors[:3]
    for p in plans.values():
        p.close()


def test_device_call_orders_after_default_stream_work():
    """A producer on torch's default stream writes the device input right before the drop-in
    call, with no synchronisation: the call must see the written values."""
    n = 1 << 20
    p = F.plan(n)
    x = O.uniform(601, n)
    want = device_reference(p, x, n, False)
    src = torch.from_numpy(x).cuda()
    d_in = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    d_out = torch.zeros_like(d_in)
    busy = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB of default-stream work
    assert torch.cuda.current_stream().cuda_stream == 0, "torch's current stream is expected to be the null stream"
    for trial in range(3):
        d_in.zero_()
        d_out.zero_()
        torch.cuda.synchronize()
        busy.uniform_()
        for _ in range(8):
            busy.mul_(1.0000001)
        d_in.copy_(src)  # the producer, queued behind the busy work
        p.forward_ptr(d_in.data_ptr(), d_out.data_ptr())  # no synchronize before
        assert torch.equal(d_out.cpu(), torch.from_numpy(want)), f"trial {trial}: read the input before it was written"


def mem_free():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def check_normal_call(L, n):
    x = O.uniform(777, n)
    src, dst = Guarded(2 * n, 1), Guarded(2 * n, 2)
    src.data[:] = x
    assert L.onefft_forward(src.ptr, dst.ptr, n) == 0, F.last_error()
    want = O.ref_fft(x, n)
    if want is None:
        want = O.fft_f64(x, n)
    assert O.rel_rms(dst.data, want) < 3e-7 * max(1, n.bit_length() - 1)


def sweep(L, what, n, attempt, max_k=200):
    """Runs attempt() with the k-th library allocation failing for k = 1, 2, ... from an empty
    cache and staging pool, until an attempt allocates fewer than k times (succeeds). Returns the
    number of failing k (the allocations the call makes)."""
    for k in range(1, max_k):
        L.onefft_cache_clear()
        with FailAlloc(k):
            ok = attempt()
        if ok:
            return k - 1
        assert F.last_error(), f"{what}: allocation {k} failed without a message"
        check_normal_call(L, 1 << 16)
    raise AssertionError(f"{what}: still failing after {max_k} allocations")


def test_allocation_failure_paths():
    L = F.lib()
    n = 1 << 22
    x = O.uniform(701, n)
    counts = {}

    # onefft_plan_create: NULL at every allocation
    def create():
        p = L.onefft_plan_create(n)
        if p:
            L.onefft_plan_destroy(p)
        return bool(p)

    # onefft_forward with the plan not cached: -1 and `out` untouched
    def forward():
        src, dst = Guarded(2 * n, 1), Guarded(2 * n, 3)
        src.data[:] = x
        dst.data[:] = CANARY
        rc = L.onefft_forward(src.ptr, dst.ptr, n)
        if rc != 0:
            assert rc == -1 and (dst.base == CANARY).all(), "failed call wrote to out"
        return rc == 0

    def prepare():
        rc = L.onefft_cache_prepare(n)
        assert rc in (0, -1)
        return rc == 0

    # one item for the persistent kernel and two batched sizes, side by side: the plan's table, its
    # second stream and two events, then a plan per batched size (a failure frees what came before)
    items = (F.RaggedItem * 3)(F.RaggedItem(0, 10, 0), F.RaggedItem(1024, 16, 1), F.RaggedItem(1024 + 65536, 20, 0))

    def ragged():
        os.environ["ONEFFT_RAGGED_BATCHED_MIN"] = "16"
        os.environ["ONEFFT_RAGGED_ORDER"] = "2"
        try:
            p = L.onefft_ragged_plan_create(ctypes.addressof(items), 3)
        finally:
            os.environ.pop("ONEFFT_RAGGED_BATCHED_MIN", None)
            os.environ.pop("ONEFFT_RAGGED_ORDER", None)
        if p:
            L.onefft_ragged_plan_destroy(p)
        return bool(p)

    d_in = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    d_out = torch.zeros_like(d_in)
    graph_plan = F.Plan(n)

    def graph():
        g = L.onefft_graph_create_batch(graph_plan.handle, d_in.data_ptr(), d_out.data_ptr(), 1, n, n, 0)
        if g:
            L.onefft_graph_destroy(g)
        return bool(g)

    free0 = None
    for rnd in range(2):  # the second round must not leave less device memory than the first
        for what, fn in (("plan_create", create), ("forward", forward), ("cache_prepare", prepare),
                         ("ragged_plan_create", ragged), ("graph_create_batch", graph)):
            print(f"allocation-failure sweep: {what}", flush=True)
            counts[what] = sweep(L, what, n, fn)
        L.onefft_cache_clear()
        free = mem_free()
        if free0 is None:
            free0 = free
        else:
            assert free0 - free < (64 << 20), f"device memory lost across a sweep round: {(free0 - free) >> 20} MiB"
    print(f"allocations per call from an empty pool: {counts}", flush=True)
    # plan table + ONEFFT_STAGING_SLOTS slots (stream + 5 events each, then ring and buffers)
    assert counts["plan_create"] >= 10 and counts["forward"] == counts["plan_create"]
    assert counts["cache_prepare"] == counts["plan_create"]
    assert counts["ragged_plan_create"] == 6 and counts["graph_create_batch"] == 1
    graph_plan.close()


def test_host_call_whose_staging_cannot_grow():
    """A plan above 2^22 points does not size the staging buffers; a host-pointer call then
    allocates them, and when that fails the void plan call leaves `out` untouched and a message."""
    L = F.lib()
    n = 1 << 23
    L.onefft_cache_clear()
    p = F.Plan(n)
    x = O.uniform(801, n)
    src, dst = Guarded(2 * n, 0), Guarded(2 * n, 0)
    src.data[:] = x
    dst.data[:] = CANARY
    failed = 0
    for k in (1, 2):
        with FailAlloc(k):
            L.onefft_plan_forward(p.handle, src.ptr, dst.ptr)
        if (dst.base == CANARY).all():
            failed += 1
            assert F.last_error()
    assert failed >= 1
    L.onefft_plan_forward(p.handle, src.ptr, dst.ptr)
    want = O.ref_fft(x, n)
    if want is None:
        want = O.fft_f64(x, n)
    assert O.rel_rms(dst.data, want) < 3e-7 * 23
    p.close()

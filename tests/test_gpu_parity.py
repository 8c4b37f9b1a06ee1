"""Parity of the gfx950 transform with the reference's own CPU transform and with the FP64
oracle, through the C ABI.

Bars (rel-L2 = sqrt(sum |got - want|^2 / sum |want|^2)):
* against the reference CPU transform on the same input: rel-L2 <= 3e-7 * log2 n;
* against the FP64 oracle: the reference tests' own bound 4e-7 + 1e-7 * log2 n
  (tests/test_onefft.c:99-101), round trip within twice that (:134-140);
* known answers as the reference tests them (:142-164); guard bands and input untouched
  (:73-97, :130-131); plan == cache bitwise (:184-205); one plan on several threads bitwise
  (:207-259).
"""
import ctypes
import os
import subprocess
import threading

import numpy as np
import pytest
import torch

import onefft_amd as F
from conftest import ROOT, tol_fp64, tol_ref
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GUARD = 64
CANARY = np.float32(12345.0)


def cplx(x):
    return x[0::2].astype(np.float64) + 1j * x[1::2]


def stream():
    return torch.cuda.current_stream().cuda_stream


class Guarded:
    """A host float buffer with canaries before and after, `misalign` floats off 64-byte alignment."""

    def __init__(self, count, misalign):
        self.base = np.full(count + 2 * GUARD + 16 + misalign, CANARY, np.float32)
        off = (-self.base.ctypes.data // 4) % 16  # to a 64-byte boundary
        self.start = off + GUARD + misalign
        self.count = count
        self.base[: self.start] = CANARY
        self.data = self.base[self.start: self.start + count]

    @property
    def ptr(self):
        return self.data.ctypes.data

    def intact(self):
        return (self.base[: self.start] == CANARY).all() and (
            self.base[self.start + self.count: self.start + self.count + GUARD] == CANARY).all()


def gpu_batch(x, n, batch, inverse=False, in_dist=None, out_dist=None):
    in_dist = in_dist or n
    out_dist = out_dist or n
    d_in = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d_out = torch.full((2 * out_dist * (batch - 1) + 2 * n + 2 * GUARD,), float(CANARY), device="cuda")
    F.plan(n).batch_ptr(d_in.data_ptr(), d_out.data_ptr() + 4 * GUARD, batch, in_dist, out_dist, inverse, stream())
    torch.cuda.synchronize()
    y = d_out.cpu().numpy()
    assert (y[:GUARD] == CANARY).all(), "wrote before the output"
    assert (y[GUARD + 2 * out_dist * (batch - 1) + 2 * n:] == CANARY).all(), "wrote after the output"
    assert np.array_equal(d_in.cpu().numpy(), x), "input modified"
    return y[GUARD: GUARD + 2 * out_dist * (batch - 1) + 2 * n]


def check(got, x, n, inverse, label):
    L = n.bit_length() - 1
    want64 = O.fft_f64(x, n, inverse)
    e64 = O.rel_rms(got, want64)
    assert e64 < tol_fp64(L), f"{label}: rel-L2 vs FP64 {e64:.3g}"
    ref = O.ref_fft(x, n, inverse)
    if ref is not None:
        eref = O.rel_rms(got, ref)
        assert eref < tol_ref(L), f"{label}: rel-L2 vs reference {eref:.3g}"
    return e64


# ---- the drop-in path: host pointers, misaligned, guard bands -------------------------------------

@pytest.mark.parametrize("L", [22, 23])
def test_host_pipeline_bitwise_equals_device_path(L):
    """Host buffers larger than the pinned staging ring (4 x 4 MiB, so its slots are reused):
    pageable (misaligned, guarded) and pinned host memory give the device path's bits."""
    n = 1 << L
    print(f"host pipeline n=2^{L}", flush=True)
    p = F.plan(n)
    x = O.uniform(40 + L, n)
    d_in = torch.from_numpy(x).cuda()
    for inverse in (False, True):
        d_out = torch.empty_like(d_in)
        p.batch_ptr(d_in.data_ptr(), d_out.data_ptr(), 1, n, n, inverse, stream())
        torch.cuda.synchronize()
        want = d_out.cpu().numpy()
        call = p.inverse_ptr if inverse else p.forward_ptr
        src, dst = Guarded(2 * n, 1), Guarded(2 * n, 3)
        src.data[:] = x
        call(src.ptr, dst.ptr)
        assert np.array_equal(dst.data, want) and dst.intact() and src.intact()
        pin_in = torch.from_numpy(x).pin_memory()
        pin_out = torch.zeros(2 * n, dtype=torch.float32).pin_memory()
        call(pin_in.data_ptr(), pin_out.data_ptr())
        assert np.array_equal(pin_out.numpy(), want)
        # pinned input, pageable output and the reverse
        dst.data[:] = 0
        call(pin_in.data_ptr(), dst.ptr)
        assert np.array_equal(dst.data, want) and dst.intact()
        pin_out.zero_()
        call(src.ptr, pin_out.data_ptr())
        assert np.array_equal(pin_out.numpy(), want)


@pytest.mark.parametrize("L", list(range(0, 21)))
def test_plan_host_pointers(L):
    n = 1 << L
    print(f"plan host n=2^{L}", flush=True)
    p = F.Plan(n)
    x = O.xorshift(2 * n, 0x9E3779B9 + L)
    src, dst, back = Guarded(2 * n, 1), Guarded(2 * n, 3), Guarded(2 * n, 0)
    src.data[:] = x
    for inverse in (False, True):
        dst.data[:] = 0
        (p.inverse_ptr if inverse else p.forward_ptr)(src.ptr, dst.ptr)
        check(dst.data, x, n, inverse, f"n=2^{L} {'inv' if inverse else 'fwd'}")
        assert dst.intact() and src.intact() and np.array_equal(src.data, x)
    p.forward_ptr(src.ptr, dst.ptr)
    p.inverse_ptr(dst.ptr, back.ptr)
    want = x.astype(np.float64) * n
    assert O.rel_rms(back.data, want) < 2 * tol_fp64(L)
    assert back.intact()


@pytest.mark.parametrize("L", [1, 2, 5, 10, 12, 14, 16, 20])
def test_known_answers(L):
    n = 1 << L
    p = F.Plan(n)
    j = 1 + (n // 3) % (n - 1) if n > 1 else 0
    x = np.zeros(2 * n, np.float32)
    x[2 * j] = 1
    y = np.empty_like(x)
    p.forward_ptr(x.ctypes.data, y.ctypes.data)
    k = np.arange(n, dtype=np.uint64)
    want = np.exp(-2j * np.pi * ((np.uint64(j) * k) % np.uint64(n)).astype(np.float64) / n)
    assert np.abs(cplx(y) - want).max() < 4e-6
    x = np.tile(np.array([0.5, -0.25], np.float32), n)
    p.forward_ptr(x.ctypes.data, y.ctypes.data)
    z = cplx(y)
    worst = max(abs(z[0] - (0.5 - 0.25j) * n) / n, (np.abs(z[1:]).max() / n) if n > 1 else 0)
    assert worst < 1e-6


def test_device_pointers_misaligned_and_mixed():
    n = 1 << 12
    p = F.Plan(n)
    x = O.uniform(5, n)
    want = O.ref_fft(x, n) if O.ref() is not None else None
    d = torch.zeros(2 * n + 8, device="cuda")
    d[1:1 + 2 * n] = torch.from_numpy(x).cuda()   # 4-byte aligned device input
    o = torch.zeros(2 * n + 8, device="cuda")
    p.forward_ptr(d.data_ptr() + 4, o.data_ptr() + 12)  # both misaligned: staged on the device
    torch.cuda.synchronize()
    got = o[3:3 + 2 * n].cpu().numpy()
    check(got, x, n, False, "misaligned device")
    assert (o[:3] == 0).all() and (o[3 + 2 * n:] == 0).all()
    host_out = np.zeros(2 * n, np.float32)
    p.forward_ptr(d.data_ptr() + 4, host_out.ctypes.data)  # device in, host out
    assert np.array_equal(host_out, got)
    dd = torch.from_numpy(x).cuda()
    oo = torch.empty_like(dd)
    p.forward_ptr(x.ctypes.data, oo.data_ptr())  # host in, device out
    torch.cuda.synchronize()
    assert np.array_equal(oo.cpu().numpy(), got)
    if want is not None:
        assert O.rel_rms(got, want) < tol_ref(12)


@pytest.mark.parametrize("L", [12, 16])
def test_host_call_with_aliased_device_buffers(L):
    """The drop-in call gets the same device buffer as input and output (which the API does not
    allow, reference include/onefft.h:61-64): the input is staged first, so the result is the
    out-of-place one, bit for bit, for a single-pass and a multi-pass size."""
    n = 1 << L
    print(f"aliased host call n=2^{L}", flush=True)
    p = F.Plan(n)
    x = O.uniform(40 + L, n)
    want = torch.empty(2 * n, device="cuda")
    d = torch.from_numpy(x).cuda()
    p.forward_ptr(d.data_ptr(), want.data_ptr())
    p.forward_ptr(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(d, want)


def test_device_calls_reject_overlap():
    """Device-pointer batch calls, ragged executions and graph captures whose input and output
    ranges overlap return -1 (NULL for a graph) with a message, and launch nothing."""
    L = F.lib()
    n = 1 << 16
    p = F.Plan(n)
    buf = torch.zeros(2 * (4 * n), device="cuda")
    a = buf.data_ptr()
    sp = stream()
    for d_in, d_out, batch in ((a, a, 1), (a, a + 8 * (n - 1), 1), (a + 8 * n, a, 2), (a, a + 8 * (2 * n), 3)):
        assert L.onefft_plan_forward_batch(p.handle, d_in, d_out, batch, n, n, sp) == -1
        assert "overlap" in F.last_error()
        assert not L.onefft_graph_create_batch(p.handle, d_in, d_out, batch, n, n, 0)
    # adjacent, not overlapping: accepted
    assert L.onefft_plan_forward_batch(p.handle, a, a + 8 * n, 1, n, n, sp) == 0
    assert L.onefft_plan_forward_batch(p.handle, a, a + 8 * (2 * n), 2, n, n, sp) == 0
    r = F.RaggedPlan([(0, 10, 0), (1024, 16, 1)])
    ext = r.extent
    assert L.onefft_ragged_execute(r._p, a, a + 8 * (ext - 1), sp) == -1
    assert "overlap" in F.last_error()
    assert L.onefft_ragged_execute(r._p, a, a + 8 * ext, sp) == 0
    r.wait(sp)
    torch.cuda.synchronize()


# ---- batches ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("L,batch", [(0, 5), (0, 70001), (1, 3), (2, 100003), (3, 257), (4, 1000), (4, 65537), (5, 129), (5, 70001), (6, 3), (6, 65601),
                                     (8, 77), (10, 33), (11, 5),
                                     (12, 9), (13, 3), (14, 3), (15, 3), (16, 2), (18, 2)])
def test_batch(L, batch):
    n = 1 << L
    print(f"batch n=2^{L} x{batch}", flush=True)
    x = O.normal(1000 + L, n * batch)
    for inverse in (False, True):
        y = gpu_batch(x, n, batch, inverse)
        want = O.fft_f64(x, n, inverse, batch=batch)
        for b in range(batch):
            seg = slice(2 * n * b, 2 * n * (b + 1))
            assert O.rel_rms(y[seg], want[seg]) < tol_fp64(L), (b, inverse)


@pytest.mark.parametrize("L,in_extra,out_extra", [(4, 24, 8), (9, 24, 8), (12, 24, 8), (16, 24, 8), (17, 3, 1),
                                                  (19, 24, 1)])
def test_batch_with_distances(L, in_extra, out_extra):
    """Distances beyond the size, odd ones included: with an odd output distance every other
    transform's output starts 8 but not 16 bytes aligned (the input passes of 2^17 and 2^19 store
    16-byte pairs only to 16-byte aligned rows)."""
    n = 1 << L
    batch, in_dist, out_dist = 6, n + in_extra, n + out_extra
    x = O.uniform(77, in_dist * batch)
    y = gpu_batch(x, n, batch, False, in_dist, out_dist)
    for b in range(batch):
        xb = x[2 * in_dist * b: 2 * in_dist * b + 2 * n]
        yb = y[2 * out_dist * b: 2 * out_dist * b + 2 * n]
        check(yb, xb, n, False, f"dist b={b}")
        if b + 1 < batch:  # the gap between outputs is untouched
            assert (y[2 * out_dist * b + 2 * n: 2 * out_dist * (b + 1)] == CANARY).all()


def test_batch_16384_x_4096_against_reference():
    """The fixed-size batch workload in full: every transform against the reference's CPU transform."""
    n, batch = 4096, 16384
    x = O.normal(1, n * batch)
    y = gpu_batch(x, n, batch)
    ref = O.ref_fft(x, n, False, batch=batch, threads=min(16, os.cpu_count() or 1))
    if ref is None:
        pytest.skip("reference not built")
    err = O.rel_rms(y, ref)
    assert err < tol_ref(12), err
    # per-transform worst case on a sample, against FP64
    idx = np.random.default_rng(0).choice(batch, 64, replace=False)
    for b in idx:
        seg = slice(2 * n * b, 2 * n * (b + 1))
        assert O.rel_rms(y[seg], O.fft_f64(x[seg], n)) < tol_fp64(12)


# ---- large single transforms ------------------------------------------------------------------------

@pytest.mark.parametrize("L", [22, 24])
def test_large_with_tone(L):
    n = 1 << L
    print(f"large n=2^{L}", flush=True)
    x = O.add_tone(O.uniform(2, n), n, round(n / 7))
    y = gpu_batch(x, n, 1)
    e64 = check(y, x, n, False, f"2^{L} fwd")
    back = gpu_batch(y, n, 1, True)
    assert O.rel_rms(back, x.astype(np.float64) * n) < 2 * tol_fp64(L)
    print(f"2^{L} rel-L2 vs FP64 {e64:.3g}", flush=True)


def test_large_2_27():
    L = 27
    n = 1 << L
    k0 = round(n / 7)
    x = O.add_tone(O.uniform(2, n), n, k0)
    y = gpu_batch(x, n, 1)
    ref = O.ref_fft(x, n)
    if ref is not None:
        assert O.rel_rms(y, ref) < tol_ref(L)
    bins = [0, 1, k0, n // 2, n - 1, 12345677]
    want = O.dft_bins_f64(x, n, bins)
    got = np.concatenate([y[2 * b: 2 * b + 2] for b in bins]).astype(np.float64)
    gc, wc = got[0::2] + 1j * got[1::2], want[0::2] + 1j * want[1::2]
    # a bin carries absolute error ~ eps sqrt(n) log n from the noise, plus a relative error on
    # its own magnitude (the tone bin is n)
    assert (np.abs(gc - wc) < 1e-6 * np.sqrt(n) * L + tol_fp64(L) * np.abs(wc)).all()
    del ref
    back = gpu_batch(y, n, 1, True)
    assert O.rel_rms(back, x.astype(np.float64) * n) < 2 * tol_fp64(L)


# ---- determinism and threads --------------------------------------------------------------------------

def test_cache_equals_plan_bitwise():
    lib = F.lib()
    for L in range(0, 13):
        n = 1 << L
        x = O.xorshift(2 * n, 99 + L)
        p = F.Plan(n)
        for inverse in (False, True):
            a = np.empty_like(x)
            b = np.empty_like(x)
            (p.inverse_ptr if inverse else p.forward_ptr)(x.ctypes.data, a.ctypes.data)
            fn = lib.onefft_inverse if inverse else lib.onefft_forward
            assert fn(x.ctypes.data, b.ctypes.data, n) == 0
            assert np.array_equal(a, b), (L, inverse)
    lib.onefft_cache_clear()
    lib.onefft_cache_clear()


def test_threads_bitwise():
    lib = F.lib()
    n = 4096
    p = F.Plan(n)
    x = O.xorshift(2 * n, 4242)
    fwd, inv = np.empty_like(x), np.empty_like(x)
    p.forward_ptr(x.ctypes.data, fwd.ctypes.data)
    p.inverse_ptr(x.ctypes.data, inv.ctypes.data)
    mismatches = []

    def work(first):
        out = np.empty_like(x)
        bad = 0
        for r in range(200):
            inverse = (r + first) & 1
            if r & 2:
                fn = lib.onefft_inverse if inverse else lib.onefft_forward
                bad += fn(x.ctypes.data, out.ctypes.data, n) != 0
            else:
                (p.inverse_ptr if inverse else p.forward_ptr)(x.ctypes.data, out.ctypes.data)
            bad += not np.array_equal(out, inv if inverse else fwd)
        mismatches.append(bad)

    lib.onefft_cache_clear()
    threads = [threading.Thread(target=work, args=(t & 1,)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert mismatches == [0, 0, 0, 0]
    lib.onefft_cache_clear()


def test_threads_staged_ring_bitwise():
    """Host calls large enough for the staging ring (several 4 MiB chunks, copied by the shared
    copy pool) from several threads at once, at two sizes: every result is bit-identical to the
    same call made alone."""
    lib = F.lib()
    sizes = (1 << 20, 1 << 22)
    data = {n: O.xorshift(2 * n, 77 + n.bit_length()) for n in sizes}
    want = {}
    for n in sizes:
        fwd, inv = np.empty_like(data[n]), np.empty_like(data[n])
        assert lib.onefft_forward(data[n].ctypes.data, fwd.ctypes.data, n) == 0
        assert lib.onefft_inverse(data[n].ctypes.data, inv.ctypes.data, n) == 0
        want[n] = (fwd, inv)
    mismatches = []

    def work(t):
        n = sizes[t & 1]
        x = data[n]
        out = np.empty_like(x)
        bad = 0
        for r in range(6):
            inverse = (r + t) & 1
            fn = lib.onefft_inverse if inverse else lib.onefft_forward
            bad += fn(x.ctypes.data, out.ctypes.data, n) != 0
            bad += not np.array_equal(out, want[n][inverse])
        mismatches.append(bad)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert mismatches == [0, 0, 0, 0]
    lib.onefft_cache_clear()


# ---- the reference's own test programs against this library ------------------------------------------

@pytest.mark.parametrize("prog,args", [("ref_test_onefft", ["18"]), ("ref_test_cpp", [])])
def test_reference_test_programs(prog, args):
    path = os.path.join(ROOT, "oracle", "_ref", prog)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (needs the reference sources when building)")
    r = subprocess.run([path, *args], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]

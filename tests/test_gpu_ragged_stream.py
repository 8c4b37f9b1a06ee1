"""The full ragged workload stream: 200000 items (log2 n uniform on 4..20, directions 50/50, from
SplitMix64 seed 3), about 184.6 GiB in all, run as consecutive chunks of at most 4 GiB of input,
one persistent launch each. The two chunks with the most 2^20-point items run here; in each, every
item of up to 2^12 points and a seeded sample of the larger items are checked against the FP64
oracle (the reference tests' bound 4e-7 + 1e-7 log2 n), so the all-items throughput the benchmark
reports rests on checked launches."""
import numpy as np
import pytest
import torch

import onefft_amd as F
from conftest import tol_fp64
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CHUNK_BYTES = 4 << 30


def chunks_of(items):
    """Consecutive chunks whose input (back to back, complex64) fits CHUNK_BYTES, as bench.py
    --report runs them; each chunk's offsets start at 0."""
    out, start = [], 0
    while start < len(items):
        base, end = items[start][0], start
        while end < len(items) and (items[end][0] - base + (1 << items[end][1])) * 8 <= CHUNK_BYTES:
            end += 1
        out.append([(o - base, l, v) for o, l, v in items[start:end]])
        start = end
    return out


def test_stream_chunks_with_most_large_items():
    items = F.ragged_items(3, 200000)
    assert len(items) == 200000
    chunks = chunks_of(items)
    total = sum(len(c) for c in chunks)
    assert total == 200000
    by_large = sorted(range(len(chunks)), key=lambda i: -sum(1 for _, l, _ in chunks[i] if l == 20))
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(2024)
    for ci in by_large[:2]:
        ch = chunks[ci]
        extent = ch[-1][0] + (1 << ch[-1][1])
        large = sum(1 for _, l, _ in ch if l == 20)
        print(f"stream chunk {ci} of {len(chunks)}: {len(ch)} items, {large} of 2^20 points, "
              f"{extent * 8 / 2**30:.2f} GiB", flush=True)
        g = torch.Generator(device="cuda")
        g.manual_seed(ci)
        x = torch.randn(extent, dtype=torch.complex64, device="cuda", generator=g)
        y = torch.zeros_like(x)
        rp = F.RaggedPlan(ch)
        assert rp.extent == extent
        rp.execute(x.data_ptr(), y.data_ptr(), s)
        rp.wait(s)
        rp.close()
        small = [it for it in ch if it[1] <= 12]
        big = [it for it in ch if it[1] > 12]
        sample = small + [big[i] for i in rng.choice(len(big), size=min(24, len(big)), replace=False)]
        # always include 2^20 items
        sample += [it for it in big if it[1] == 20][:4]
        worst = 0.0
        for off, l, inv in sample:
            n = 1 << l
            xs = x[off:off + n].cpu().numpy().view(np.float32)
            ys = y[off:off + n].cpu().numpy().view(np.float32)
            e = O.rel_rms(ys, O.fft_f64(xs, n, bool(inv)))
            assert e < tol_fp64(l), (ci, off, l, inv, e)
            worst = max(worst, e / tol_fp64(l))
        print(f"  checked {len(sample)} items, worst rel-L2 / bar {worst:.3f}", flush=True)
        del x, y
        torch.cuda.empty_cache()

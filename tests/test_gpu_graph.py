"""Plan executions captured as HIP graphs: every replay equals the eager result bitwise, with the
program's own eager work and synchronisations between replays."""
import numpy as np
import pytest
import torch

import onefft_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("L,batch,inverse", [(10, 300, False), (12, 64, True), (16, 3, False), (20, 1, True),
                                              (22, 2, False), (24, 1, True)])
def test_batch_graph_replays(L, batch, inverse):
    n = 1 << L
    x = torch.from_numpy(O.normal(50 + L, n * batch)).cuda()
    eager, out = torch.empty_like(x), torch.empty_like(x)
    p = F.plan(n)
    p.batch_ptr(x.data_ptr(), eager.data_ptr(), batch, n, n, inverse, stream())
    torch.cuda.synchronize()
    g = F.Graph.batch(p, x.data_ptr(), out.data_ptr(), batch, n, n, inverse)
    other = torch.empty_like(x)
    for replay in range(4):
        out.fill_(7.0)
        torch.cuda.synchronize()
        g.launch(stream())
        torch.cuda.synchronize()
        assert torch.equal(out, eager), replay
        # eager work between replays: a different transform into another buffer
        p.batch_ptr(x.data_ptr(), other.data_ptr(), batch, n, n, not inverse, stream())
        torch.cuda.synchronize()
    g.close()


@pytest.mark.parametrize("order", ["0", "1", "2", "3", "4", "5"])
def test_ragged_graph_replays(monkeypatch, order):
    """A ragged execution with batched groups (2^16 and 2^17 items) and the persistent kernel, in
    each order (groups first, persistent kernel first, side by side on the plan's second stream,
    the groups split over both streams; forked and joined by events inside the capture), replays
    bit for bit."""
    items = F.ragged_items(9, 400, lo=2, hi=17)
    extent = items[-1][0] + (1 << items[-1][1])
    x = torch.from_numpy(O.normal(9, extent)).cuda()
    eager, out = torch.zeros_like(x), torch.zeros_like(x)
    monkeypatch.setenv("ONEFFT_RAGGED_BATCHED_MIN", "16")
    monkeypatch.setenv("ONEFFT_RAGGED_ORDER", order)
    rp = F.RaggedPlan(items)
    monkeypatch.delenv("ONEFFT_RAGGED_BATCHED_MIN")
    monkeypatch.delenv("ONEFFT_RAGGED_ORDER")
    rp.execute(x.data_ptr(), eager.data_ptr(), stream())
    rp.wait(stream())
    g = F.Graph.ragged(rp, x.data_ptr(), out.data_ptr())
    for replay in range(4):
        out.zero_()
        torch.cuda.synchronize()
        g.launch(stream())
        rp.wait(stream())
        assert torch.equal(out, eager), replay
        rp.execute(x.data_ptr(), eager.data_ptr(), stream())  # eager work between replays
        rp.wait(stream())
    # the eager ragged result is itself right
    y, xs = eager.cpu().numpy(), x.cpu().numpy()
    for off, l, inv in items[:50] + [it for it in items if it[1] >= 16][:10]:
        seg = slice(2 * off, 2 * (off + (1 << l)))
        assert O.rel_rms(y[seg], O.fft_f64(xs[seg], 1 << l, bool(inv))) < 4e-7 + 1e-7 * l


def test_graph_rejects_bad_arguments():
    lib = F.lib()
    assert not lib.onefft_graph_create_batch(None, None, None, 1, 1, 1, 0)
    assert not lib.onefft_graph_create_ragged(None, None, None)
    assert lib.onefft_graph_launch(None, None) == -1
    lib.onefft_graph_destroy(None)

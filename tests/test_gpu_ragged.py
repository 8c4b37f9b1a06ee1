"""Ragged batches in one persistent launch (onefft_ragged_*), through the C ABI.

Every item is checked against the FP64 oracle (4e-7 + 1e-7 log2 n, the reference tests' bound)
and, at the workload's size, against the reference's own CPU transform of every item
(rel-L2 <= 3e-7 log2 n); the space between items is never written; the per-item comparison path
gives the same results within the same bounds.
"""
import numpy as np
import pytest
import torch

import onefft_amd as F
from conftest import tol_fp64, tol_ref
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CANARY = 12345.0


def stream():
    return torch.cuda.current_stream().cuda_stream


def layout(sizes, gap=0, seed=0):
    rng = np.random.default_rng(seed)
    items, off = [], 0
    for l in sizes:
        items.append((off, int(l), int(rng.integers(0, 2))))
        off += (1 << int(l)) + gap
    return items, off


def run(items, extent, x, per_item=False):
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.full((2 * extent,), CANARY, device="cuda")
    p = F.RaggedPlan(items)
    assert p.extent <= extent
    if per_item:
        p.execute_per_item(d_in.data_ptr(), d_out.data_ptr(), stream())
        torch.cuda.synchronize()
    else:
        p.execute(d_in.data_ptr(), d_out.data_ptr(), stream())
        p.wait(stream())
    assert np.array_equal(d_in.cpu().numpy(), x)
    return d_out.cpu().numpy(), p


def check_items(items, x, y, extent, fp64_max_log2=20):
    covered = np.zeros(extent, bool)
    for off, l, inv in items:
        n = 1 << l
        seg = slice(2 * off, 2 * (off + n))
        covered[off:off + n] = True
        if l <= fp64_max_log2:
            e = O.rel_rms(y[seg], O.fft_f64(x[seg], n, bool(inv)))
            assert e < tol_fp64(l), (off, l, inv, e)
    gaps = np.repeat(~covered, 2)
    assert (y[gaps] == CANARY).all(), "wrote between items"


@pytest.mark.parametrize("per_item", [False, True])
def test_every_size_with_gaps(per_item):
    sizes = list(range(0, 21)) + list(range(20, -1, -1)) + [4, 4, 4, 12, 13, 3, 3]
    items, extent = layout(sizes, gap=3, seed=1)
    x = O.normal(31, extent)
    y, p = run(items, extent, x, per_item)
    check_items(items, x, y, extent)


def test_many_small_items_fill_tasks():
    rng = np.random.default_rng(5)
    sizes = rng.integers(0, 14, 3000)
    items, extent = layout(sizes, seed=2)
    x = O.uniform(32, extent)
    y, p = run(items, extent, x)
    check_items(items, x, y, extent)


def test_two_pass_items_many_tiles():
    sizes = [20, 19, 18, 17, 16, 15, 14, 13] * 2
    items, extent = layout(sizes, gap=1, seed=3)
    x = O.normal(33, extent)
    y, p = run(items, extent, x)
    check_items(items, x, y, extent)


@pytest.mark.parametrize("grid,order", [("1", "0"), ("96", "3"), ("192", "4"), ("256", "5")])
def test_persistent_grid_bitwise(monkeypatch, grid, order):
    """Fewer persistent workgroups (ONEFFT_RAGGED_GRID; one workgroup included, which then runs
    every task of both phases itself) change only which workgroup computes what: with every item in
    the persistent kernel, and with groups in the split orders, the results equal the default
    grid's bit for bit."""
    sizes = [20, 19, 18, 17, 16, 15, 14, 13, 10, 4] * 2
    items, extent = layout(sizes, gap=1, seed=11)
    x = O.normal(36, extent)
    monkeypatch.setenv("ONEFFT_RAGGED_BATCHED_MIN", "21" if order == "0" else "16")
    monkeypatch.setenv("ONEFFT_RAGGED_ORDER", order)
    want, _ = run(items, extent, x)
    monkeypatch.setenv("ONEFFT_RAGGED_GRID", grid)
    got, _ = run(items, extent, x)
    monkeypatch.delenv("ONEFFT_RAGGED_GRID")
    assert np.array_equal(got, want)
    check_items(items, x, got, extent)


@pytest.mark.parametrize("split,tile_run", [("20", "1"), ("19", "3"), ("18", "256"), ("21", "2"), ("15", "2"),
                                             ("14", "1")])
def test_split_and_tile_run_bitwise(monkeypatch, split, tile_run):
    """Other split sizes and tile-run lengths (ONEFFT_RAGGED_SPLIT / ONEFFT_RAGGED_TILE_RUN) change
    only which workgroup computes what: the results equal the defaults' bit for bit."""
    sizes = [20, 19, 18, 17, 14, 13, 10, 4] * 2
    items, extent = layout(sizes, gap=1, seed=9)
    x = O.normal(34, extent)
    monkeypatch.setenv("ONEFFT_RAGGED_BATCHED_MIN", "21")  # every item in the persistent kernel
    want, _ = run(items, extent, x)
    monkeypatch.setenv("ONEFFT_RAGGED_SPLIT", split)
    monkeypatch.setenv("ONEFFT_RAGGED_TILE_RUN", tile_run)
    got, _ = run(items, extent, x)
    monkeypatch.delenv("ONEFFT_RAGGED_SPLIT")
    monkeypatch.delenv("ONEFFT_RAGGED_TILE_RUN")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("batched_min,order", [("15", "0"), ("17", "1"), ("20", "2"), ("16", "2"), ("21", "2"),
                                                ("15", "3"), ("20", "3"), ("16", "4"), ("20", "4"), ("21", "3"),
                                                ("16", "5"), ("19", "5")])
def test_batched_groups(monkeypatch, batched_min, order):
    """Items of 2^batched_min points and more run as offset-batched multi-pass groups (one per size
    and direction), the rest in the persistent kernel (21: none batched), in each order (groups
    first, persistent kernel first, side by side on two streams, the groups split over two streams
    with the persistent kernel after both or after one of them): every item within the FP64 bar,
    nothing written between items, and the persistent kernel's task count falls as more sizes are
    batched."""
    sizes = list(range(0, 21)) + list(range(20, 11, -1)) * 2 + [15, 16, 17] * 3
    items, extent = layout(sizes, gap=5, seed=13)
    x = O.normal(35, extent)
    monkeypatch.setenv("ONEFFT_RAGGED_ORDER", order)
    monkeypatch.setenv("ONEFFT_RAGGED_BATCHED_MIN", batched_min)
    y, p = run(items, extent, x)
    tasks = p.tasks
    p.close()
    check_items(items, x, y, extent)
    monkeypatch.setenv("ONEFFT_RAGGED_BATCHED_MIN", "21")
    q = F.RaggedPlan(items)
    assert (tasks < q.tasks) == (batched_min != "21"), (batched_min, tasks, q.tasks)
    q.close()


def test_rejects_bad_items():
    lib = F.lib()
    for items in ([(0, 21, 0)], [(0, 4, 2)], [(0, 4, 0), (15, 4, 0)], [(16, 4, 0), (0, 5, 0)]):
        with pytest.raises(F.OneFFTError):
            F.RaggedPlan(items)
    p = F.RaggedPlan([])
    assert p.extent == 0 and p.tasks == 0
    p.execute(0, 0, stream())
    p.wait(stream())
    assert lib.onefft_ragged_execute(None, None, None, None) == -1


def test_workload_prefix_against_reference():
    """The ragged workload (seed 3, log2 n uniform on 4..20, directions 50/50), the longest prefix
    whose input fits 4 GiB, in one launch: every item against the reference CPU transform."""
    items = F.ragged_items(3, 200000, max_bytes=4 << 30)
    extent = items[-1][0] + (1 << items[-1][1])
    x = O.normal(3, extent)
    print(f"ragged prefix: {len(items)} items, {extent * 8 / 2**30:.2f} GiB", flush=True)
    y, p = run(items, extent, x)
    r = O.ref_ragged(x, items, threads=16)
    if r is None:
        pytest.skip("reference not built")
    ref, _ = r
    worst = 0.0
    for off, l, inv in items:
        seg = slice(2 * off, 2 * (off + (1 << l)))
        e = O.rel_rms(y[seg], ref[seg])
        worst = max(worst, e / tol_ref(l))
    assert worst < 1.0, worst
    # FP64 on a sample of items
    for k in np.random.default_rng(0).choice(len(items), 40, replace=False):
        off, l, inv = items[k]
        seg = slice(2 * off, 2 * (off + (1 << l)))
        assert O.rel_rms(y[seg], O.fft_f64(x[seg], 1 << l, bool(inv))) < tol_fp64(l)


def test_phase_wait_timeout_is_reported(monkeypatch):
    """Failure handling, forced from the host: a plan whose bounded wait for the first phase gives
    up at once (limit 0) still runs the launch to its end (every access stays inside the items),
    and onefft_ragged_wait reports the timeout; a default plan of the same items is then correct."""
    items, extent = layout([20] + [10] * 200, seed=5)  # the 2^20 item's second passes wait
    x = O.normal(33, extent)
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.zeros(2 * extent, device="cuda")
    monkeypatch.setenv("ONEFFT_RAGGED_SPIN_LIMIT", "0")
    monkeypatch.setenv("ONEFFT_RAGGED_BATCHED_MIN", "21")  # the 2^20 item in the persistent kernel
    p = F.RaggedPlan(items)
    monkeypatch.delenv("ONEFFT_RAGGED_SPIN_LIMIT")
    monkeypatch.delenv("ONEFFT_RAGGED_BATCHED_MIN")
    p.execute(d_in.data_ptr(), d_out.data_ptr(), stream())
    with pytest.raises(F.OneFFTError, match="timed out"):
        p.wait(stream())
    p.close()
    y, q = run(items, extent, x)
    check_items(items, x, y, extent)
    q.close()

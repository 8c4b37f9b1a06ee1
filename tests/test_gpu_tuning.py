"""The tuning table: a plan created under a table takes its choices, results stay within the
parity bars for every choice, and malformed lines are ignored (each case in a fresh process,
because the table is read once per process)."""
import os
import subprocess
import sys
import textwrap

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CHILD = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import numpy as np, torch
    import onefft_amd as F
    from oracle import oracle as O
    s = torch.cuda.current_stream().cuda_stream
    for L in (12, 16, 22):
        n = 1 << L
        x = O.normal(L, n)
        d = torch.from_numpy(x).cuda(); o = torch.empty_like(d)
        F.Plan(n).batch_ptr(d.data_ptr(), o.data_ptr(), 1, n, n, False, s)
        torch.cuda.synchronize()
        e = O.rel_rms(o.cpu().numpy(), O.fft_f64(x, n))
        assert e < 4e-7 + 1e-7 * L, (L, e)
    print("ok")
""")


def run_with_table(tmp_path, text):
    table = tmp_path / "table.tsv"
    table.write_text(text)
    env = dict(os.environ, ONEFFT_TUNE_FILE=str(table))
    for k in ("ONEFFT_BATCH_VARIANT", "ONEFFT_PASS_SPLIT", "ONEFFT_PASS_COLS", "ONEFFT_PASS_LAYOUT"):
        env.pop(k, None)
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT)], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def test_table_choices_are_used_and_correct(tmp_path):
    run_with_table(tmp_path, "# test table\n"
                             "12\tsingle\tnt=0,layout=1\t1\t1\n"
                             "16\tmulti\tsplit=6,10;cols=64,8\t1\t1\n"
                             "22\tmulti\tsplit=6,6,10;cols=32,16,8\t1\t1\n")


def test_table_pass_layout_is_used_and_correct(tmp_path):
    run_with_table(tmp_path, "# test table\n"
                             "12\tsingle\tnt=1,layout=2,tw64=2\t1\t1\n"
                             "16\tmulti\tsplit=6,10;cols=64,8;layout=2\t1\t1\n"
                             "22\tmulti\tsplit=11,11;cols=4,4;layout=2\t1\t1\n")


def test_table_wide_halves_tiles_are_used_and_correct(tmp_path):
    run_with_table(tmp_path, "# test table\n"
                             "16\tmulti\tsplit=6,10;cols=64,16;layout=2\t1\t1\n"
                             "22\tmulti\tsplit=11,11;cols=8,8;layout=2\t1\t1\n")


def test_malformed_table_is_ignored(tmp_path):
    run_with_table(tmp_path, "12 single garbage\n"
                             "16\tmulti\tsplit=6,9;cols=16,16\t1\t1\n"      # factors do not sum to 16
                             "22\tmulti\tsplit=11,11;cols=8,8\t1\t1\n"      # 8 columns of 2^11 points
                             "22\tmulti\tsplit=13,9;cols=2,16\t1\t1\n"      # factor above 2^12
                             "16\tmulti\tsplit=8,8;cols=16,16;layout=1\t1\t1\n"  # no swizzled passes
                             "22\tmulti\tsplit=11,11;cols=16,16;layout=2\t1\t1\n"  # 32768-point tiles
                             "99\tsingle\tnt=1,layout=0\t1\t1\n")


RAGGED_CHILD = textwrap.dedent("""
    import json, os, sys
    root, want, differs = sys.argv[1], json.loads(sys.argv[2]), sys.argv[3] == "1"
    sys.path.insert(0, root)
    import torch
    import onefft_amd as F
    from oracle import oracle as O
    sizes = [20, 19, 18, 17, 16, 15, 14, 13, 12, 8, 4] * 2
    items, end = [], 0
    for i, l in enumerate(sizes):
        items.append((end, l, i % 2))
        end += (1 << l) + 3
    def plan(env):
        os.environ.update(env)
        try:
            return F.RaggedPlan(items)
        finally:
            for k in env:
                os.environ.pop(k)
    p = plan({})
    ref = plan(want)
    default = plan({"ONEFFT_RAGGED_BATCHED_MIN": "16", "ONEFFT_RAGGED_SPLIT": "19", "ONEFFT_RAGGED_TILE_RUN": "2"})
    assert p.tasks == ref.tasks, (p.tasks, ref.tasks)
    assert (p.tasks != default.tasks) == differs, (p.tasks, default.tasks)
    x = O.normal(7, end)
    d = torch.from_numpy(x).cuda()
    o = torch.zeros_like(d)
    s = torch.cuda.current_stream().cuda_stream
    p.execute(d.data_ptr(), o.data_ptr(), s)
    p.wait(s)
    y = o.cpu().numpy()
    for off, l, inv in items:
        seg = slice(2 * off, 2 * (off + (1 << l)))
        e = O.rel_rms(y[seg], O.fft_f64(x[seg], 1 << l, bool(inv)))
        assert e < 4e-7 + 1e-7 * l, (l, inv, e)
    print("ok")
""")


def run_ragged_with_table(tmp_path, text, want, differs):
    """A ragged plan made under the table has the persistent-kernel task count of the choices
    `want` (set through the environment, which overrides the table) and, if `differs`, not that of
    the built-in defaults; its execution is within the FP64 bar."""
    import json

    table = tmp_path / "table.tsv"
    table.write_text(text)
    env = dict(os.environ, ONEFFT_TUNE_FILE=str(table))
    for k in ("ONEFFT_RAGGED_BATCHED_MIN", "ONEFFT_RAGGED_SPLIT", "ONEFFT_RAGGED_TILE_RUN", "ONEFFT_RAGGED_ORDER",
              "ONEFFT_RAGGED_GRID"):
        env.pop(k, None)
    r = subprocess.run([sys.executable, "-c", RAGGED_CHILD, ROOT, json.dumps(want), "1" if differs else "0"],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def test_table_ragged_choices_are_used_and_correct(tmp_path):
    run_ragged_with_table(tmp_path, "20\tragged\tbatched_min=20;split=18;tile_run=1;order=2\t1\t1\n",
                          {"ONEFFT_RAGGED_BATCHED_MIN": "20", "ONEFFT_RAGGED_SPLIT": "18", "ONEFFT_RAGGED_TILE_RUN": "1"},
                          differs=True)


def test_table_ragged_grid_and_split_order(tmp_path):
    run_ragged_with_table(tmp_path, "20\tragged\tbatched_min=17;split=19;tile_run=2;order=3;grid=128\t1\t1\n",
                          {"ONEFFT_RAGGED_BATCHED_MIN": "17", "ONEFFT_RAGGED_SPLIT": "19", "ONEFFT_RAGGED_TILE_RUN": "2"},
                          differs=True)


def test_table_ragged_groups_off(tmp_path):
    run_ragged_with_table(tmp_path, "20\tragged\tbatched_min=21;split=19;tile_run=2\t1\t1\n",
                          {"ONEFFT_RAGGED_BATCHED_MIN": "21", "ONEFFT_RAGGED_SPLIT": "19", "ONEFFT_RAGGED_TILE_RUN": "2"},
                          differs=True)


def test_malformed_ragged_lines_are_ignored(tmp_path):
    run_ragged_with_table(tmp_path, "20\tragged\tbatched_min=14;split=19;tile_run=2\t1\t1\n"  # batched single-pass sizes
                                    "20\tragged\tbatched_min=17;split=19\t1\t1\n"              # a field missing
                                    "20\tragged\tbatched_min=17;split=19;tile_run=0\t1\t1\n"   # no tiles per task
                                    "20\tragged\tbatched_min=17;split=13;tile_run=2\t1\t1\n"  # split at a single-pass size
                                    "20\tragged\tbatched_min=17;split=19;tile_run=2;order=6\t1\t1\n"  # no such order
                                    "20\tragged\tbatched_min=17;split=19;tile_run=2;order=3;grid=-4\t1\t1\n",  # bad grid
                          {"ONEFFT_RAGGED_BATCHED_MIN": "16", "ONEFFT_RAGGED_SPLIT": "19", "ONEFFT_RAGGED_TILE_RUN": "2"},
                          differs=False)


def test_committed_table_parses():
    path = os.path.join(ROOT, "onefft_amd", "tune_gfx950.tsv")
    if not os.path.exists(path):
        pytest.skip("no committed tuning table")
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        log2n, kind, params, gbs, batch = line.rstrip("\n").split("\t")
        assert 0 <= int(log2n) <= 30 and kind in ("single", "multi", "ragged") and float(gbs) > 0 and int(batch) >= 1


def test_committed_table_multi_choices():
    """Every multi-pass line of the committed table up to 2^26 points, run at the batch it was
    tuned at (so the library takes that line's regime and choice): the first and the last
    transform of the batch against the FP64 oracle."""
    import numpy as np
    import torch

    import onefft_amd as F
    from conftest import tol_fp64
    from oracle import oracle as O

    path = os.path.join(ROOT, "onefft_amd", "tune_gfx950.tsv")
    if not os.path.exists(path):
        pytest.skip("no committed tuning table")
    s = torch.cuda.current_stream().cuda_stream
    checked = 0
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        log2n, kind, params, _, batch = line.rstrip("\n").split("\t")
        L, batch = int(log2n), int(batch)
        if kind != "multi" or L > 26:
            continue
        n = 1 << L
        print(f"2^{L} x{batch}: {params}", flush=True)
        g = torch.Generator(device="cuda").manual_seed(1000 + L)
        x = torch.randn(batch, n, dtype=torch.complex64, device="cuda", generator=g)
        y = torch.empty_like(x)
        p = F.Plan(n)
        p.batch_ptr(x.data_ptr(), y.data_ptr(), batch, n, n, False, s)
        torch.cuda.synchronize()
        for b in sorted({0, batch - 1}):
            xin = x[b].cpu().numpy().view(np.float32)
            got = y[b].cpu().numpy().view(np.float32)
            e = O.rel_rms(got, O.fft_f64(xin, n))
            assert e < tol_fp64(L), (L, batch, params, b, e)
        p.close()
        del x, y
        torch.cuda.empty_cache()
        checked += 1
    assert checked >= 20


def test_every_single_pass_variant(monkeypatch):
    """Every kernel variant the tuner may choose, at every single-pass size, both directions,
    against the FP64 oracle (variants are chosen at plan creation from ONEFFT_BATCH_VARIANT)."""
    import numpy as np
    import torch

    import onefft_amd as F
    from conftest import tol_fp64
    from oracle import oracle as O
    from onefft_amd.tune import tw64_levels

    s = torch.cuda.current_stream().cuda_stream
    for L in range(0, 15):
        n = 1 << L
        batch = 3
        x = O.normal(60 + L, n * batch)
        d = torch.from_numpy(x).cuda()
        variants = [(nt, lay, 0, 0) for nt in (0, 1) for lay in (0, 1)]
        variants += [(nt, 0, lv, 0) for lv in tw64_levels(L) for nt in (0, 1)]
        if L >= 8:  # the halves layout (kernels.hpp: kHalvesMinLog2)
            variants += [(nt, 2, lv, 0) for lv in [0] + tw64_levels(L) for nt in (0, 1)]
        if L == 5:  # the DPP pair exchange of the 32-point kernel
            variants += [(nt, 0, 0, 1) for nt in (0, 1)]
        for nt, lay, lv, xl in variants:
            monkeypatch.setenv("ONEFFT_BATCH_VARIANT", f"{nt},{lay},{lv},{xl}")
            p = F.Plan(n)
            for inv in (False, True):
                o = torch.empty_like(d)
                p.batch_ptr(d.data_ptr(), o.data_ptr(), batch, n, n, inv, s)
                torch.cuda.synchronize()
                got = o.cpu().numpy()
                for b in range(batch):
                    seg = slice(2 * n * b, 2 * n * (b + 1))
                    e = O.rel_rms(got[seg], O.fft_f64(x[seg], n, inv))
                    assert e < tol_fp64(L), (L, nt, lay, lv, xl, inv, b, e)
            p.close()


def test_dpp_pair_exchange_bitwise_equals_lds(monkeypatch):
    """The 32-point kernel's pair exchange by DPP lane swaps gives the LDS exchange's bits (the same
    additions), over whole 4096-value blocks and a partial last block, both directions."""
    import torch

    import onefft_amd as F

    s = torch.cuda.current_stream().cuda_stream
    n, batch = 32, 128 * 5 + 37  # 5 whole blocks of 128 transforms and a partial one
    x = torch.randn(batch * n, dtype=torch.complex64, device="cuda")
    outs = {}
    for xl in (0, 1):
        monkeypatch.setenv("ONEFFT_BATCH_VARIANT", f"1,0,0,{xl}")
        p = F.Plan(n)
        for inv in (False, True):
            o = torch.full_like(x, complex(7.0, 7.0))
            p.batch_ptr(x.data_ptr(), o.data_ptr(), batch, n, n, inv, s)
            torch.cuda.synchronize()
            outs[xl, inv] = o
        p.close()
    for inv in (False, True):
        assert torch.equal(outs[0, inv].view(torch.float32), outs[1, inv].view(torch.float32)), inv
        ref = torch.fft.fft(x.view(batch, n).to(torch.complex128)) if not inv else \
            torch.fft.ifft(x.view(batch, n).to(torch.complex128)) * n
        err = (torch.linalg.vector_norm(outs[1, inv].view(batch, n) - ref) / torch.linalg.vector_norm(ref)).item()
        assert err < 4e-7 + 1e-7 * 5, err


@pytest.mark.parametrize("split,cols,batch", [("10,11", "8,4", 3), ("11,11", "4,4", 2), ("11,12", "4,2", 1),
                                              ("12,11", "2,4", 1), ("12,12", "2,2", 1), ("6,12", "16,2", 2),
                                              ("6,6,11", "16,16,4", 1), ("6,12,6", "16,2,16", 2),
                                              ("11,6,6", "4,16,16", 1)])
@pytest.mark.parametrize("layout", [0, 2])
def test_splits_with_wide_factors(monkeypatch, split, cols, batch, layout):
    """The 2^11- and 2^12-point pass factors (4 and 2 columns per tile) that two- and three-pass
    splits may use, chosen at plan creation through ONEFFT_PASS_SPLIT / ONEFFT_PASS_COLS: every
    transform of a small batch, forward, against the FP64 oracle, and the round trip."""
    check_split(monkeypatch, split, cols, batch, layout)


@pytest.mark.parametrize("split,cols,batch", [("10,11", "16,8", 3), ("11,11", "8,8", 2), ("12,12", "4,4", 1),
                                              ("9,9", "32,32", 4), ("8,10", "64,16", 4),
                                              ("6,12,6", "16,4,16", 2), ("11,6,6", "8,16,16", 1),
                                              ("8,11,6", "64,8,16", 1)])
def test_halves_wide_tiles(monkeypatch, split, cols, batch):
    """The halves layout's 16384-point pass tiles (1024 threads: 2^8 x 64, 2^9 x 32, 2^10 x 16,
    2^11 x 8, 2^12 x 4 columns), as the input pass and as in-place passes."""
    check_split(monkeypatch, split, cols, batch, 2)


@pytest.mark.parametrize("split,cols,batch,layout", [("8,4,4", "64,64,64", 2, 2), ("8,5,4", "64,32,64", 1, 2),
                                                     ("7,4,4", "64,64,64", 2, 0)])
def test_columns_wider_than_low_extent(monkeypatch, split, cols, batch, layout):
    """In-place passes with more columns per tile than their low-index extent M_f (here 64 and 32
    columns over M_f = 16), in both LDS layouts: a tile then spans several of the pass's
    sub-transform groups, and the column table's entries are those of the column mod M_f."""
    check_split(monkeypatch, split, cols, batch, layout)


def check_split(monkeypatch, split, cols, batch, layout):
    import numpy as np
    import torch

    import onefft_amd as F
    from conftest import tol_fp64
    from oracle import oracle as O
    from onefft_amd.tune import pass_cols_ok

    # counts a pass does not allow would be ignored by the library: every case must be a real one
    for f, c in zip(split.split(","), cols.split(",")):
        assert pass_cols_ok(int(f), int(c), layout), (split, cols, layout)
    L = sum(int(v) for v in split.split(","))
    n = 1 << L
    print(f"split {split} cols {cols} layout {layout} n=2^{L} x{batch}", flush=True)
    monkeypatch.setenv("ONEFFT_PASS_SPLIT", split)
    monkeypatch.setenv("ONEFFT_PASS_COLS", cols)
    monkeypatch.setenv("ONEFFT_PASS_LAYOUT", str(layout))
    p = F.Plan(n)
    monkeypatch.delenv("ONEFFT_PASS_SPLIT")
    monkeypatch.delenv("ONEFFT_PASS_COLS")
    monkeypatch.delenv("ONEFFT_PASS_LAYOUT")
    s = torch.cuda.current_stream().cuda_stream
    x = O.normal(70 + L, n * batch)
    d = torch.from_numpy(x).cuda()
    y = torch.empty_like(d)
    z = torch.empty_like(d)
    p.batch_ptr(d.data_ptr(), y.data_ptr(), batch, n, n, False, s)
    p.batch_ptr(y.data_ptr(), z.data_ptr(), batch, n, n, True, s)
    torch.cuda.synchronize()
    yh = y.cpu().numpy()
    want = O.fft_f64(x, n, False, batch=batch)
    for b in range(batch):
        seg = slice(2 * n * b, 2 * n * (b + 1))
        assert O.rel_rms(yh[seg], want[seg]) < tol_fp64(L), (split, b)
    rt = float(np.linalg.norm(z.cpu().numpy() / n - x) / np.linalg.norm(x))
    assert rt < 2 * tol_fp64(L), (split, rt)
    p.close()

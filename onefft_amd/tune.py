"""Offline autotuner: measures the kernel choices for each transform size on this GPU and writes the
tuning table the library reads at plan creation (tuning.hip).

    python -m onefft_amd.tune [--min 0] [--max 30] [--out onefft_amd/tune_gfx950.tsv]
    python -m onefft_amd.tune --ragged [--out onefft_amd/tune_gfx950.tsv]

* sizes up to 2^14 (one LDS pass): the single-pass variants (nontemporal global access on or off,
  padded or swizzled LDS, and, where the size has such stages, FP64-base twiddles at levels 1..3:
  for the stages with Ns >= 64, >= 16, or all; from 2^8 also the halves LDS layout at every
  level; at 2^5 the pair exchange by DPP lane swaps);
* larger sizes: every split n = N_1 ... N_p into 2..4 passes with factors 2^6..2^10 (4 passes only
  from 2^24; two- and three-pass splits also with 2^11 and 2^12), each with the default and the
  widest allowed columns per workgroup and with the padded or the halves LDS layout, and with the
  halves layout's 16384-point tiles for the input pass;
each in two working-set regimes, because the best choice differs between them: a batch whose
input plus output fit the 256 MiB Infinity Cache (16 n batch = 256 MiB up to 2^20 points; one
transform from 2^21 points, where a single large transform is the common use and has too few tiles
per pass for the batch's best split), and a batch of 1 GiB in and 1 GiB out (one transform above
2^26 points). The library picks the regime per call from n and the batch (runtime.hpp:
tune_regime).

Each candidate is a plan created under the environment overrides (ONEFFT_BATCH_VARIANT,
ONEFFT_PASS_SPLIT, ONEFFT_PASS_COLS, ONEFFT_PASS_LAYOUT) and timed with HIP events: after 3 warm-up launches, the
fastest of 5 samples of 10 back-to-back launches (single launches timed alone varied by more than
the 2-5 % that separates the best variants). All candidates of a size give results within FP32 rounding of each other; the
tuner checks each against the first candidate (rel-L2 <= 3e-7 log2 n). The fastest is written.
The table is read once per process, so this module sets ONEFFT_TUNE_FILE to a missing file before
the first plan: the candidates are measured without any table. Lines of sizes not measured in a
run are kept from the existing table.

Ragged plans (--ragged, run after the sizes in a process of its own, because it measures under the
table just written: the batched groups use its multi-pass choices): the workload is the 4 GiB
prefix of the seed-3 ragged stream (log2 n uniform on 4..20, directions 50/50); the candidates are
the smallest size run as offset-batched multi-pass groups (2^15..2^20, or none), and, where the
persistent kernel keeps items of 2^18 points or more, the split size and the tiles per pass task,
and, where both parts exist, their order: groups first, persistent kernel first, side by side on
two streams, or the groups split over two streams, balanced by bytes, with the persistent kernel
after both, after the lighter one or first on the second stream; then, for the three best, fewer
persistent workgroups (ONEFFT_RAGGED_BATCHED_MIN / _SPLIT / _TILE_RUN / _ORDER / _GRID). Each execution is timed like a size's candidates
and checked against the first candidate's output; the fastest is written as the table's ragged line.
"""
from __future__ import annotations

import argparse
import itertools
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_gfx950.tsv")


HALVES_MIN_LOG2 = 8  # kernels.hpp: kHalvesMinLog2
GROUP = 10  # launches per timed sample: back to back, as batches run, so a 2 % difference shows


def pass_cols_ok(log2nf: int, cols: int, layout: int = 0) -> bool:
    """The kernels' rule (kernels.hpp: pass_cols_ok): tiles of at most 8192 points with the padded
    layout, 16384 with the halves layout."""
    lp = min(log2nf, 4)
    threads = cols << (log2nf - lp)
    tile_max = 16384 if layout == 2 else 8192
    return (cols in (2, 4, 8, 16, 32, 64) and (cols << log2nf) <= tile_max and 64 <= threads <= 1024
            and (cols >= 8 or log2nf >= 11))


def default_cols(log2nf: int) -> int:
    for c in (16, 32, 64, 8, 4):
        if pass_cols_ok(log2nf, c):
            return c
    return 2


def widest_cols(log2nf: int, layout: int = 0) -> int:
    return max(c for c in (2, 4, 8, 16, 32, 64) if pass_cols_ok(log2nf, c, layout))


def tw64_levels(log2n: int):
    """The FP64-base levels that give distinct kernels for 2^log2n points (fft_common.hpp:
    stage_uses_tw64 and tw64_levels, over the single-pass stage schedule)."""
    p = min(log2n, 4)
    stages = 1 if log2n <= p else log2n // p + (1 if log2n % p else 0)
    log2r = [log2n if log2n <= p else (log2n % p if (log2n % p and s == 0) else p) for s in range(stages)]
    log2ns = [sum(log2r[:s]) for s in range(stages)]
    has = [s for s in range(1, stages) if log2r[s] >= 2]
    out, used = [], set()
    for level, min_log2ns in ((1, 6), (2, 4), (3, 0)):
        now = {s for s in has if log2ns[s] >= min_log2ns}
        if now != used:
            out.append(level)
        used = now
    return out


def splits(L: int):
    """Factors 2^6..2^10 for 2..4 passes (4 only from 2^24); two- and three-pass splits may also
    use 2^11 and 2^12, whose tiles hold only 4 and 2 columns."""
    out = []
    for p in (2, 3, 4):
        if p == 4 and L < 24:
            continue
        top = 13 if p <= 3 else 11
        for combo in itertools.product(range(6, top), repeat=p):
            if sum(combo) == L:
                out.append(list(combo))
    return out


def candidates(L: int, single: bool):
    if single:
        combos = [(nt, lay, 0) for nt in (1, 0) for lay in (0, 1)]
        # FP64-base levels that change this size's kernel (padded layout)
        combos += [(nt, 0, level) for level in tw64_levels(L) for nt in (1, 0)]
        if L >= HALVES_MIN_LOG2:  # the halves layout, with every level
            combos += [(nt, 2, level) for level in [0] + tw64_levels(L) for nt in (1, 0)]
        out = [dict(kind="single", params=f"nt={nt},layout={lay},tw64={t64}",
                    env={"ONEFFT_BATCH_VARIANT": f"{nt},{lay},{t64}"}) for nt, lay, t64 in combos]
        if L == 5:  # 32 points: the pair exchange by DPP lane swaps instead of LDS
            out += [dict(kind="single", params=f"nt={nt},layout=0,tw64=0,xlane=1",
                         env={"ONEFFT_BATCH_VARIANT": f"{nt},0,0,1"}) for nt in (1, 0)]
        return out
    out = []
    for s in splits(L):
        # the passes' LDS layout (padded, or halves: kernels.hpp), with the default and the widest
        # columns; the halves layout also with its wider tiles for the first (input) pass, whose
        # 2^11- and 2^12-point columns gain 2x longer runs (its in-place passes keep fewer
        # resident workgroups with the wider tiles)
        cands = []
        for layout in (0, 2):
            cands += [(layout, [default_cols(f) for f in s]), (layout, [widest_cols(f) for f in s])]
        cands.append((2, [widest_cols(f) for f in s[:-1]] + [widest_cols(s[-1], 2)]))
        seen = set()
        for layout, cols in cands:
            if (layout, tuple(cols)) in seen:
                continue
            seen.add((layout, tuple(cols)))
            params = f"split={','.join(map(str, s))};cols={','.join(map(str, cols))}"
            out.append(dict(kind="multi", params=params + (";layout=2" if layout else ""),
                            env={"ONEFFT_PASS_SPLIT": ",".join(map(str, s)),
                                 "ONEFFT_PASS_COLS": ",".join(map(str, cols)),
                                 "ONEFFT_PASS_LAYOUT": str(layout)}))
    return out


def tune_size(F, torch, timed, L: int, batch: int, single: bool) -> str:
    """Measures every candidate for 2^L points at this batch; returns the table line of the best."""
    n = 1 << L
    x = torch.randn(batch, n, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    first = None
    results = []
    for cand in candidates(L, single):
        os.environ.update(cand["env"])
        plan = F.Plan(n)
        ms = timed(plan, x, y, batch, n)
        if first is None:
            first = y.clone()
        else:
            err = (torch.linalg.vector_norm(y - first) / torch.linalg.vector_norm(first)).item()
            if err > 3e-7 * max(L, 1):
                raise SystemExit(f"log2n={L} {cand['params']}: result differs from the first candidate ({err:.3g})")
        results.append((16.0 * n * batch / (ms * 1e-3) / 1e9, cand["kind"], cand["params"]))
        print(f"  log2n={L} batch={batch} {cand['params']}: {results[-1][0]:.0f} GB/s", flush=True)
        plan.close()
        for k in cand["env"]:
            os.environ.pop(k, None)
    best = max(results)
    print(f"log2n={L:2d} batch={batch:6d} best {best[1]} {best[2]} {best[0]:.0f} GB/s ({len(results)} candidates)",
          flush=True)
    return f"{L}\t{best[1]}\t{best[2]}\t{best[0]:.0f}\t{batch}"


RAGGED_KEYS = ("ONEFFT_RAGGED_BATCHED_MIN", "ONEFFT_RAGGED_SPLIT", "ONEFFT_RAGGED_TILE_RUN", "ONEFFT_RAGGED_ORDER",
               "ONEFFT_RAGGED_GRID")
RAGGED_GRIDS = (384, 256, 192, 128)  # persistent workgroups tried for the best first-stage choices


def ragged_candidates():
    """(batched_min, split, tile_run, order, grid) choices that give distinct executions: the split
    size and the tile run matter only where the persistent kernel keeps items of the split size
    (split < batched_min), the order only where there are groups (batched_min <= 20); elsewhere they
    are left at the defaults. The grid (persistent workgroups) is 0 here, as many as are resident;
    smaller grids are tried in a second stage."""
    out = []
    for bmin in (15, 16, 17, 18, 19, 20, 21):
        shapes = [(19, 2)] + [(split, run) for split in (18, 19, 20) if split < bmin for run in (1, 2, 4)]
        for split, run in dict.fromkeys(shapes):
            for order in ((0, 1, 2, 3, 4, 5) if bmin <= 20 else (0,)):
                out.append((bmin, split, run, order, 0))
    return out


def tune_ragged(F, torch, stream) -> str:
    """Measures the ragged choices on the 4 GiB seed-3 prefix; returns the table's ragged line."""
    sp = stream.cuda_stream
    items = F.ragged_items(3, 200000, max_bytes=4 << 30)
    extent = items[-1][0] + (1 << items[-1][1])
    points = sum(1 << l for _, l, _ in items)
    x = torch.randn(extent, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    first, results = None, []

    def measure(bmin, split, run, order, grid):
        nonlocal first
        env = dict(zip(RAGGED_KEYS, (str(bmin), str(split), str(run), str(order), str(grid))))
        os.environ.update(env)
        try:
            plan = F.RaggedPlan(items)
        finally:
            for k in env:
                os.environ.pop(k, None)
        execute = lambda: plan.execute(x.data_ptr(), y.data_ptr(), sp)
        for _ in range(2):
            execute()
        plan.wait(sp)
        ts = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(3):
                execute()
            e1.record(stream)
            plan.wait(sp)
            ts.append(e0.elapsed_time(e1) / 3)
        if first is None:
            first = y.clone()
        else:
            err = (torch.linalg.vector_norm(y - first) / torch.linalg.vector_norm(first)).item()
            if err > 3e-7 * 20:
                raise SystemExit(f"ragged {env}: result differs from the first candidate ({err:.3g})")
        params = f"batched_min={bmin};split={split};tile_run={run};order={order}" + (f";grid={grid}" if grid else "")
        results.append((16.0 * points / (min(ts) * 1e-3) / 1e9, params, (bmin, split, run, order)))
        print(f"  ragged {params}: {min(ts):.3f} ms, {results[-1][0]:.0f} GB/s, {plan.tasks} persistent tasks",
              flush=True)
        plan.close()

    for cand in ragged_candidates():
        measure(*cand)
    # second stage: fewer persistent workgroups for the three best choices (each publishes and
    # waits once at the phase change)
    for _, _, choice in sorted(results, reverse=True)[:3]:
        for grid in RAGGED_GRIDS:
            measure(*choice, grid)
    best = max(results)
    print(f"ragged best {best[1]} {best[0]:.0f} GB/s ({len(results)} candidates)", flush=True)
    return f"20\tragged\t{best[1]}\t{best[0]:.0f}\t{len(items)}"


def line_key(line: str):
    """Lines of one size and batch (so one regime), or the ragged line, replace each other."""
    f = line.split("\t")
    return ("ragged",) if f[1] == "ragged" else (int(f[0]), f[4])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--min", type=int, default=0)
    ap.add_argument("--max", type=int, default=30)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--ragged", action="store_true", help="tune the ragged plans' choices only, under the table at --out")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged step after the sizes")
    args = ap.parse_args(argv)
    os.environ["ONEFFT_TUNE_FILE"] = (os.path.abspath(args.out) if args.ragged
                                      else os.path.join(ROOT, "no-tuning-table-while-tuning"))
    for k in ("ONEFFT_BATCH_VARIANT", "ONEFFT_PASS_SPLIT", "ONEFFT_PASS_COLS", "ONEFFT_PASS_LAYOUT") + RAGGED_KEYS:
        os.environ.pop(k, None)

    import torch

    sys.path.insert(0, ROOT)
    import onefft_amd as F

    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def timed(plan, x, y, batch, n):
        run = lambda: plan.batch_ptr(x.data_ptr(), y.data_ptr(), batch, n, n, False, sp)
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        ts = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(GROUP):
                run()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / GROUP)
        return min(ts)

    lines = []
    props = torch.cuda.get_device_properties(0)
    header = [
        "# OneFFT gfx950 tuning table: log2n, kind, params, measured GB/s (16 n batch / time), batch",
        f"# written by python -m onefft_amd.tune on {props.name} ({time.strftime('%Y-%m-%d')})",
    ]

    kept = []
    if os.path.exists(args.out):
        kept = [l.rstrip("\n") for l in open(args.out) if l.strip() and not l.startswith("#")]

    def write(lines):
        new = {line_key(l) for l in lines}
        merged = [l for l in kept if line_key(l) not in new] + lines
        merged.sort(key=lambda l: (1, 0, "") if line_key(l) == ("ragged",) else (0, int(l.split("\t")[0]),
                                                                                  l.split("\t")[4].zfill(12)))
        with open(args.out, "w") as f:
            f.write("\n".join(header + merged) + "\n")

    if args.ragged:
        write([tune_ragged(F, torch, stream)])
        print(f"wrote {args.out}", flush=True)
        return

    for L in range(args.min, args.max + 1):
        n = 1 << L
        single = L <= 14
        for regime in (0, 1):
            if regime == 0:
                if 16 * n > (256 << 20):
                    continue
                batch = 1 if L >= 21 else max(1, (1 << 24) // n)
            else:
                batch = max(1, (1 << 26 if single else 1 << 27) // n)
                if 16 * n * batch <= (256 << 20):
                    continue
            lines.append(tune_size(F, torch, timed, L, batch, single))
            torch.cuda.empty_cache()
            write(lines)  # after every size, so an interrupted run keeps what it measured
    print(f"wrote {args.out}", flush=True)
    if not args.no_ragged:
        import subprocess

        subprocess.run([sys.executable, "-m", "onefft_amd.tune", "--ragged", "--out", args.out], cwd=ROOT, check=True)


if __name__ == "__main__":
    main()

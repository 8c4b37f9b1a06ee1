"""Benchmark of the gfx950 OneFFT path on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--json-extra]

Workload: 16384 independent forward c2c FFTs of n = 4096 complex64 points (512 MiB in, 512 MiB
out, contiguous transforms), re/im i.i.d. standard normal (torch generator, seed 1), resident in
HBM before timing. One step = one batched forward transform of the whole batch through
onefft_plan_forward_batch on torch's current stream.

Prints one JSON line: throughput in 5 n log2 n GFLOP/s (the reference's unit,
docs/measurements.md:54-59), the HBM roofline of the transform kernel (16 n bytes per transform
over the kernel's HIP-event time, against MI355X's 8 TB/s), the same batch as 16384 per-transform
calls (the "per-call" comparison), and the reference's own CPU transform timed on this host's
cores on a bounded sample (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N, BATCH = 4096, 16384


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="approximate CPU-baseline budget")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the timed transform launches plus one 1 GiB device copy (for counter runs)")
    return ap.parse_args()


def traffic_from_profiles(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy-kernel calibration), committed under profiles/; None when not measured."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(budget_s: float):
    """The reference's own transform (oracle/_ref, built from its sources) on this host's cores,
    on a bounded sample of the same workload: a slice of the batch, all threads sharing one plan."""
    try:
        import numpy as np
        from oracle import oracle as O
    except Exception as e:  # noqa: BLE001
        return {"value": None, "error": f"oracle unavailable: {e}"}
    if O.ref() is None:
        return {"value": None, "error": "reference build oracle/_ref missing"}
    threads = max(1, min(16, O.ref_online_cpus() or (os.cpu_count() or 1)))
    sample = 256 * threads
    x = O.normal(1, N * sample)
    samples = 5
    per = O.ref_time_batch(x, N, sample, threads, False, sample_seconds=budget_s / (samples + 3), samples=samples)
    flops = 5.0 * N * math.log2(N) * sample
    return {
        "value": round(flops / per / 1e9, 2),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "reference",
        "isa": O.ref_implementation(),
        "sample": f"{sample} of the {BATCH} transforms of n={N} (re/im N(0,1), seed 1), {threads} threads "
                  f"sharing one plan, min over {samples} samples (bench/bench.c method)",
        "gb_s": round(16.0 * N * sample / per / 1e9, 2),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build shards nothing "
              "across GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch

    import onefft_amd as F

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    x = torch.randn(BATCH, N, dtype=torch.complex64, device=dev, generator=g)
    # torch's complex normal has variance 1/2 per part; scale to N(0,1) re and im
    x.mul_(math.sqrt(2.0))
    y = torch.empty_like(x)
    plan = F.plan(N)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def step():
        plan.batch_ptr(x.data_ptr(), y.data_ptr(), BATCH, N, N, False, sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / args.steps  # one kernel launch per step

    flops = 5.0 * N * math.log2(N) * BATCH
    alg_bytes = 16.0 * N * BATCH
    if args.kernel_only:
        z = torch.empty_like(x)
        z.copy_(x)  # a plain copy of known size: 512 MiB read, 512 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"kernel_only": True, "ms_per_step": ms, "gb_s": alg_bytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    gflops = flops / (ms * 1e-3) / 1e9
    achieved = alg_bytes / (ms * 1e-3) / 1e9

    # round trip on the device: inverse(forward(x)) = n x
    z = torch.empty_like(x)
    plan.batch_ptr(y.data_ptr(), z.data_ptr(), BATCH, N, N, True, sp)
    torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm(z / N - x) / torch.linalg.vector_norm(x)).item()

    # per-call: the same batch as one device call per transform
    pc = 4096
    stride = N * 8
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for b in range(pc):
        plan.batch_ptr(x.data_ptr() + b * stride, y.data_ptr() + b * stride, 1, N, N, False, sp)
    e1.record(stream)
    torch.cuda.synchronize()
    per_call_ms = e0.elapsed_time(e1) / pc * BATCH

    # a plain device copy of the same bytes, for the achievable-bandwidth context
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    z.copy_(x)
    c0.record(stream)
    for _ in range(5):
        z.copy_(x)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = alg_bytes / (c0.elapsed_time(c1) / 5 * 1e-3) / 1e9

    cpu = {"value": None, "skipped": True} if args.no_cpu else cpu_baseline(args.cpu_seconds)

    line = {
        "metric": "5n·log2n GFLOP/s & %HBM peak: batched vs per-call, rel-L2 <= 3e-7·log2 n",
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: re/im i.i.d. N(0,1), torch generator seed 1, resident in HBM",
        "config": {"workload": "batch 16384 x 4096-point c2c forward, complex64 interleaved, out of place",
                   "n": N, "batch": BATCH, "bytes_in": 8 * N * BATCH, "bytes_out": 8 * N * BATCH},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("batch4096")},
        "cpu_baseline": cpu,
        "per_call": {"value": round(flops / (per_call_ms * 1e-3) / 1e9, 1), "unit": "GFLOP/s",
                     "ms_for_batch": round(per_call_ms, 3),
                     "method": f"{pc} single-transform device calls timed, scaled to {BATCH}",
                     "batched_speedup": round(per_call_ms / ms, 2)},
        "copy_gb_s": round(copy_gbs, 1),
        "roundtrip_rel_l2": rt,
        "wall_s_timed": round(wall, 4),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

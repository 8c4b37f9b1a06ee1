"""Builds ``onefft_amd/libonefft_amd.so`` (gfx950 HIP kernels + the C-ABI runtime) with hipcc.

    python -m onefft_amd.build            # or __graft_entry__.build()

Every HIP source is compiled for gfx950 only, with packed-FP32 instructions disabled, into
``onefft_amd/_build/`` and linked into one shared library whose only exported symbols are the
``onefft_*`` functions declared in ``include/onefft.h``.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libonefft_amd.so")
# measurement only: the bench's copy kernel with the single-pass kernel's access pattern
STREAM_SOURCE = "stream_copy.hip"
STREAM_LIB = os.path.join(HERE, "libonefft_stream.so")
SOURCES = ["runtime.hip", "alloc.hip", "staging.hip", "ragged.hip", "graph.hip", "tuning.hip", "k_batch.hip", "k_batch_tw.hip", "k_batch_tw_large.hip", "k_pass.hip", "k_pass_halves.hip",
           "k_ragged.hip", "ragged_static.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
    ]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *compile_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile every source (always, also when a library exists) and link the library."""
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    workers = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        objs = list(pool.map(_compile, sources + [STREAM_SOURCE]))
    stream_obj = objs.pop()
    script = os.path.join(OBJ, "exports.map")
    with open(script, "w") as f:
        f.write("{ global: onefft_*; local: *; };\n")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, f"-Wl,--version-script={script}",
           "-ldl", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", stream_obj, "-o", STREAM_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(STREAM_LIB + ".tmp", STREAM_LIB)
    if verbose:
        print(f"built {LIB} and {STREAM_LIB}", flush=True)
    return LIB


def build_oracle(verbose: bool = True) -> None:
    """The test oracle (and, where the reference sources exist, its builds and the drop-in test
    programs linked against the library). Test infrastructure only."""
    make = ["make", "-s", "-C", os.path.join(ROOT, "oracle")]
    r = subprocess.run(make, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if os.path.exists("/root/reference/tests/test_onefft.c"):
        r = subprocess.run(make + ["dropin"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"drop-in test build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle", flush=True)


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()

"""Writes tests/golden/ref_vectors.npz: input/output vectors of the reference's own CPU transform.

Run in a container where ``make -C oracle`` could build ``oracle/_ref`` from the reference
sources. The vectors let CPU-only test runs pin the oracle without the reference build:

* ``x{L}``: input of 2^L points, the reference tests' xorshift32 stream (state 0x9e3779b9,
  tests/test_onefft.c:57-61), consumed in order of L;
* ``simd_f{L}`` / ``simd_i{L}``: forward / inverse by the reference as shipped (run-time ISA
  selection; which ISA made them is stored in ``isa``);
* ``port_f{L}`` / ``port_i{L}``: the same by the reference's portable path (ONEFFT_NO_SIMD);
* ``roots``: rows (k, log2size, re, im) of the reference's onefft_root (src/onefft.c:137-141);
* ``uniform{s}`` / ``normal{s}``: the first 64 floats of the workload generators.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SIZES = list(range(0, 11)) + [12]


def main() -> None:
    if O.ref() is None or O.ref_portable() is None:
        raise SystemExit("oracle/_ref is not built (needs the reference sources): run `make -C oracle`")
    out = {"isa": np.array(O.ref_implementation())}
    stream = O.xorshift(2 * sum(1 << L for L in SIZES))
    pos = 0
    for L in SIZES:
        n = 1 << L
        x = stream[pos:pos + 2 * n].copy()
        pos += 2 * n
        out[f"x{L}"] = x
        for inv, tag in ((False, "f"), (True, "i")):
            out[f"simd_{tag}{L}"] = O.ref_fft(x, n, inv)
            out[f"port_{tag}{L}"] = O.ref_fft(x, n, inv, portable=True)
    rows = []
    rng = np.random.default_rng(12345)
    for log2size in range(0, 31):
        size = 1 << log2size
        ks = {0, size // 8, size // 4, size // 3, size // 2, size - 1} | {int(k) for k in rng.integers(0, size, 24)}
        for k in sorted(k for k in ks if 0 <= k < size):
            re, im = O.ref_root(k, log2size)
            rows.append((k, log2size, re, im))
    out["roots"] = np.array(rows, dtype=np.float64)
    for s in range(5):
        out[f"uniform{s}"] = O.uniform(s, 32)
        out[f"normal{s}"] = O.normal(s, 32)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"), "isa", out["isa"])


if __name__ == "__main__":
    main()

"""The library's twiddles, on the CPU: the tables the host builds (onefft_amd/csrc/plan_tables.hpp,
compiled with g++ from tests/twiddle_check.cpp) against the reference's own twiddle function
onefft_root / onefft_root_double (src/onefft.c:112-141), and the twiddles the kernels form from
those tables in FP64 before rounding once.

* root32 / root64 bitwise equal to the reference's onefft_root / onefft_root_double (live build in
  oracle/_ref where it exists, else the committed reference vectors and the oracle, which is pinned
  against the reference in test_oracle.py), for every log2 size 0..30;
* every entry of every FP32 stage table of 2^1..2^14 points bitwise equal to onefft_root;
* the FP64 stage bases and the two-level FP64 tables of w_n bitwise equal to onefft_root_double;
* the FP64-formed twiddles of every single-pass kernel variant, of the pass kernels of the tuned
  multi-pass splits, and of the ragged kernel's two-pass items: counted against onefft_root per
  plan shape. Two kinds of difference are allowed: a component that onefft_root holds as exactly 0
  (quarter turns, exact through its octant folding) left by the FP64 products as a residue below
  1e-15, and, rarely, one unit in the last place where the exact value lies within the FP64
  formation error of an FP32 rounding boundary. Any other difference fails.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

SRC = os.path.join(ROOT, "tests", "twiddle_check.cpp")
U64 = ctypes.c_uint64
P64 = ctypes.POINTER(ctypes.c_uint64)

# Bound on the share of FP64-formed twiddles that round differently from onefft_root. A formed
# value's relative error is a few times 1e-16 after at most ~30 FP64 products; an FP32 rounding
# boundary lies within that distance of the exact value with probability ~1e-8 per component, so
# 1e-6 leaves a wide margin while a wrong exponent, sign or table index (which changes almost every
# value) fails at once.
MISMATCH_SHARE = 1e-6


@pytest.fixture(scope="module")
def tw(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    lib = tmp_path_factory.mktemp("tw") / "libtwiddle_check.so"
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__host__=", "-D__device__=",
           "-D__forceinline__=inline", "-shared", "-fPIC", SRC, "-o", str(lib)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = ctypes.CDLL(str(lib))
    F, D = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    L.tw_root32.argtypes = [U64, ctypes.c_uint, F, F]
    L.tw_root64.argtypes = [U64, ctypes.c_uint, D, D]
    L.tw_stage_table_size.restype = ctypes.c_uint32
    L.tw_stage_table64_size.restype = ctypes.c_uint32
    L.tw_stage_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.tw_stage_table64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.tw_stage_table_exponents.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.tw_two_level_lo_bits.restype = ctypes.c_uint
    L.tw_two_level.argtypes = [ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]
    L.tw_count_single.argtypes = [ctypes.c_int, ctypes.c_int, P64]
    L.tw_count_pass.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_uint, ctypes.c_uint32, ctypes.c_uint32, P64, P64]
    L.tw_count_ragged_pass.argtypes = [ctypes.c_uint, ctypes.c_uint32, P64]
    return L


def root32(L, k, log2size):
    re, im = ctypes.c_float(), ctypes.c_float()
    L.tw_root32(k, log2size, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def root64(L, k, log2size):
    re, im = ctypes.c_double(), ctypes.c_double()
    L.tw_root64(k, log2size, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def want32(k, log2size):
    """onefft_root from the live reference build, else the oracle (pinned against the reference)."""
    return O.ref_root(k, log2size) if O.ref() is not None else O.root(k, log2size)


def want64(k, log2size):
    if O.ref() is not None:
        L = O.ref()
        re, im = ctypes.c_double(), ctypes.c_double()
        L.ref_root_double(k, log2size, ctypes.byref(re), ctypes.byref(im))
        return re.value, im.value
    return O.root_double(k, log2size)


def test_root32_matches_committed_reference_vectors(tw, golden):
    bad = [(int(k), int(l)) for k, l, re, im in golden["roots"]
           if root32(tw, int(k), int(l)) != (np.float32(re), np.float32(im))]
    assert not bad, bad[:10]


def test_root32_and_root64_match_reference_every_log2size(tw):
    rng = np.random.default_rng(11)
    for log2size in range(0, 31):
        ks = {0, (1 << log2size) - 1} | set(rng.integers(0, 1 << log2size, 300).tolist())
        if log2size <= 10:
            ks = range(1 << log2size)
        for k in ks:
            assert root32(tw, k, log2size) == tuple(np.float32(v) for v in want32(k, log2size)), (k, log2size)
            assert root64(tw, k, log2size) == want64(k, log2size), (k, log2size)


@pytest.mark.parametrize("log2n", range(1, 15))
def test_stage_tables_bitwise_equal_reference(tw, log2n):
    """The FP32 stage table of the single-pass plan of 2^log2n points (16 points per thread) and
    its FP64 bases, entry by entry."""
    lp = min(log2n, 4)
    size = tw.tw_stage_table_size(log2n, lp)
    if size:
        table = np.empty(2 * size, np.float32)
        k = np.empty(size, np.uint64)
        l = np.empty(size, np.uint32)
        tw.tw_stage_table(log2n, lp, table.ctypes.data)
        tw.tw_stage_table_exponents(log2n, lp, k.ctypes.data, l.ctypes.data)
        for i in range(size):
            want = want32(int(k[i]), int(l[i]))
            assert (table[2 * i], table[2 * i + 1]) == (np.float32(want[0]), np.float32(want[1])), (i, k[i], l[i])
    size64 = tw.tw_stage_table64_size(log2n, lp)
    if size64:
        t64 = np.empty(2 * size64, np.float64)
        tw.tw_stage_table64(log2n, lp, t64.ctypes.data)
        # stage s's bases are w_{Ns R}^k, k < Ns, in stage order
        stages = 1 if log2n <= lp else log2n // lp + (1 if log2n % lp else 0)
        log2r = [log2n % lp if (log2n % lp and s == 0) else lp for s in range(stages)]
        i, lns = 0, log2r[0]
        for s in range(1, stages):
            if log2r[s] >= 2:
                for kk in range(1 << lns):
                    assert (t64[2 * i], t64[2 * i + 1]) == want64(kk, lns + log2r[s]), (s, kk)
                    i += 1
            lns += log2r[s]
        assert i == size64


@pytest.mark.parametrize("log2n", [15, 16, 20, 22, 27, 30])
def test_two_level_tables_bitwise_equal_reference(tw, log2n):
    b = tw.tw_two_level_lo_bits(log2n)
    assert b == (log2n + 1) // 2
    lo = np.empty(2 << b, np.float64)
    hi = np.empty(2 << (log2n - b), np.float64)
    tw.tw_two_level(log2n, lo.ctypes.data, hi.ctypes.data)
    rng = np.random.default_rng(log2n)
    for i in {0, 1, (1 << b) - 1} | set(rng.integers(0, 1 << b, 200).tolist()):
        assert (lo[2 * i], lo[2 * i + 1]) == want64(i, log2n), i
    for i in {0, 1, (1 << (log2n - b)) - 1} | set(rng.integers(0, 1 << (log2n - b), 200).tolist()):
        assert (hi[2 * i], hi[2 * i + 1]) == want64(i << b, log2n), i


def unpack(out):
    """(checked, zero residues, one-ulp differences, other differences, largest residue)"""
    return (int(out[0]), int(out[1]), int(out[2]), int(out[3]),
            float(np.array([out[4]], np.uint64).view(np.float64)[0]))


def count(fn, *args):
    out = (U64 * 5)()
    fn(*args, out)
    return unpack(out)


# Quarter-turn components left as residues, per plan shape: the FP64 products miss an exact 0 only
# where a recurrence or two-level product lands on a quarter turn, a handful of values per shape
# (0 to 3 measured over every shape checked here).
MAX_ZERO_RESIDUES = 8


def check_counts(what, checked, zeros, one_ulp, other, residue):
    assert checked > 0, what
    assert other == 0, (what, other)
    assert residue < 1e-15, (what, residue)
    assert zeros <= MAX_ZERO_RESIDUES, (what, zeros)
    assert one_ulp <= max(1, MISMATCH_SHARE * checked), (what, one_ulp, checked)


@pytest.mark.parametrize("log2n", range(5, 15))
def test_single_pass_formed_twiddles(tw, log2n):
    rows = []
    for level in (1, 2, 3):
        c = count(tw.tw_count_single, log2n, level)
        if c[0]:
            check_counts(f"single 2^{log2n} level {level}", *c)
        rows.append(c)
    print(f"single 2^{log2n}: (checked, zero residues, one-ulp, other, residue) per FP64-base level 1..3: {rows}")


def pass_shapes(split, cols):
    """(lf, log2m, columns per workgroup) of each pass in execution order for a split
    log2 N_1 .. log2 N_p and its columns, listed in the same order (k_pass.hpp)."""
    p = len(split)
    out = []
    for i in range(p):
        f = p - i
        out.append((split[f - 1], sum(split[f:]), cols[f - 1]))
    return out


def tuned_splits():
    """The multi-pass splits and columns the committed tuning table picks, per size and regime."""
    path = os.path.join(ROOT, "onefft_amd", "tune_gfx950.tsv")
    out = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            log2n, kind, params = line.split("\t")[:3]
            if kind == "multi":
                fields = params.split(";")
                split = [int(v) for v in fields[0][len("split="):].split(",")]
                cols = [int(v) for v in fields[1][len("cols="):].split(",")]
                out.append((int(log2n), split, cols))
    return out


# besides the tuned splits: in-place passes with more columns per tile than their low extent
WIDE_TILE_SPLITS = [(16, [8, 4, 4], [64, 64, 64]), (17, [8, 5, 4], [64, 32, 64]), (15, [7, 4, 4], [64, 64, 64])]


@pytest.mark.parametrize("log2n,split,cols", tuned_splits() + WIDE_TILE_SPLITS)
def test_pass_formed_twiddles(tw, log2n, split, cols):
    """Every pass of a tuned split: pass twiddles for every low index (every 2^k-th above 2^24
    points, to bound the time) and the sub-transform's stage recurrences."""
    step = 1 if log2n <= 24 else 1 << (log2n - 24)
    rows = []
    for lf, log2m, c in pass_shapes(split, cols):
        op, os_ = (U64 * 5)(), (U64 * 5)()
        tw.tw_count_pass(log2n, lf, log2m, c, step, op, os_)
        cp, cs = unpack(op), unpack(os_)
        if log2m:
            check_counts(f"pass 2^{log2n} {split} lf={lf}", *cp)
        if cs[0]:
            check_counts(f"pass stages 2^{lf}", *cs)
        rows.append((lf, log2m, cp[:4], cs[:4]))
    print(f"2^{log2n} split {split}: (lf, log2m, pass (checked, zeros, one-ulp, other), stages) {rows}")


@pytest.mark.parametrize("log2n", range(14, 21))
def test_ragged_pass_formed_twiddles(tw, log2n):
    c = count(tw.tw_count_ragged_pass, log2n, 1)
    check_counts(f"ragged 2^{log2n}", *c)
    print(f"ragged 2^{log2n}: {c}")


def test_planted_exponent_mistake_is_caught(tw):
    """A formation with a wrong exponent (here: the twiddle of a neighbouring index) must fail the
    bound: compare the table entries with the reference one index off."""
    lp = 4
    size = tw.tw_stage_table_size(12, lp)
    table = np.empty(2 * size, np.float32)
    k = np.empty(size, np.uint64)
    l = np.empty(size, np.uint32)
    tw.tw_stage_table(12, lp, table.ctypes.data)
    tw.tw_stage_table_exponents(12, lp, k.ctypes.data, l.ctypes.data)
    wrong = sum((table[2 * i], table[2 * i + 1]) != tuple(np.float32(v) for v in want32(int(k[i]) + 1, int(l[i])))
                for i in range(0, size, 7))
    assert wrong > 0.9 * len(range(0, size, 7))

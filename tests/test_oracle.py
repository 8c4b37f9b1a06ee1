"""Pins the CPU oracle (oracle/onefft_oracle.c) before anything is checked against it.

Checks, per oracle function:
* twiddles: bitwise against the reference's onefft_root (committed vectors and, where built, the
  live reference), exact at eighth turns, unit modulus, conjugate symmetry;
* the FP32 restatement: bitwise against the reference's portable path, within tolerance of the
  reference's SIMD path and of FP64;
* the FP64 transform: against the O(n^2) definition, numpy's float64 FFT, known answers
  (impulse, constant, tone), Parseval and linearity;
* generators: committed first values and sample statistics;
and that plausible mistakes planted in a copy of an oracle result fail these checks.
"""
import numpy as np
import pytest

from conftest import tol_fp64, tol_ref
from oracle import oracle as O

SIZES = list(range(0, 11)) + [12]


def cplx(x):
    return x[0::2].astype(np.float64) + 1j * x[1::2]


def inter(z):
    out = np.empty(2 * z.size, np.float64)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def rel(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    d = np.sqrt(((got - want) ** 2).sum())
    w = np.sqrt((want ** 2).sum())
    return d / w if w else d


# ---- twiddles ------------------------------------------------------------------------------------

def test_roots_match_reference_vectors(golden):
    rows = golden["roots"]
    bad = [(int(k), int(l)) for k, l, re, im in rows if O.root(int(k), int(l)) != (np.float32(re), np.float32(im))]
    assert not bad, bad[:10]


def test_roots_match_live_reference():
    if O.ref() is None:
        pytest.skip("reference not built here (oracle/_ref)")
    rng = np.random.default_rng(7)
    for log2size in range(0, 31):
        for k in set(rng.integers(0, 1 << log2size, 200).tolist()) | {0, (1 << log2size) - 1}:
            assert O.root(k, log2size) == O.ref_root(k, log2size), (k, log2size)


def test_roots_exact_at_eighth_turns_and_unit_modulus():
    h = np.float32(np.sqrt(0.5))
    for log2size in range(3, 31):
        eighth = 1 << (log2size - 3)
        want = [(1, 0), (h, -h), (0, -1), (-h, -h), (-1, 0), (-h, h), (0, 1), (h, h)]
        for m, (wr, wi) in enumerate(want):
            re, im = O.root(m * eighth, log2size)
            assert re == np.float32(wr) and im == np.float32(wi), (m, log2size, re, im)
        for k in (1, 3, eighth + 1, 5 * eighth - 7):
            re, im = O.root_double(k, log2size)
            assert abs(re * re + im * im - 1) < 4e-16
            # w^(N-k) = conj(w^k): exactly in FP32; in FP64 too away from the eighth turns,
            # where cos and sin of pi/4 may differ in the last place
            r2, i2 = O.root((1 << log2size) - k, log2size)
            assert (r2, i2) == (lambda a, b: (a, -b))(*O.root(k, log2size))
            if k % eighth:
                assert O.root_double((1 << log2size) - k, log2size) == (re, -im)


# ---- the FP32 restatement ------------------------------------------------------------------------

@pytest.mark.parametrize("L", SIZES)
def test_fft_f32_bitwise_equals_reference_portable(golden, L):
    x, n = golden[f"x{L}"], 1 << L
    assert np.array_equal(O.fft_f32(x, n, False), golden[f"port_f{L}"])
    assert np.array_equal(O.fft_f32(x, n, True), golden[f"port_i{L}"])


@pytest.mark.parametrize("L", SIZES)
def test_reference_simd_within_tolerance_of_fp64(golden, L):
    x, n = golden[f"x{L}"], 1 << L
    for inv, tag in ((False, "f"), (True, "i")):
        want = O.fft_f64(x, n, inv)
        assert O.rel_rms(golden[f"simd_{tag}{L}"], want) < tol_fp64(L)
        assert O.rel_rms(O.fft_f32(x, n, inv), want) < tol_fp64(L)
        assert O.rel_rms(O.fft_f32(x, n, inv), golden[f"simd_{tag}{L}"]) < tol_ref(L)


def test_live_reference_matches_vectors(golden):
    if O.ref() is None:
        pytest.skip("reference not built here (oracle/_ref)")
    for L in SIZES:
        x, n = golden[f"x{L}"], 1 << L
        assert np.array_equal(O.ref_fft(x, n, False, portable=True), golden[f"port_f{L}"])
        # the SIMD path depends on the host ISA; compare bitwise only on the same ISA
        if O.ref_implementation() == str(golden["isa"]):
            assert np.array_equal(O.ref_fft(x, n, False), golden[f"simd_f{L}"])


def test_fft_f32_closed_forms():
    x = np.array([1, 2, 3, 4, 5, 6, 7, 8], np.float32)
    z = cplx(x)
    for inv in (False, True):
        got = cplx(O.fft_f32(x, 4, inv))
        want = np.fft.ifft(z) * 4 if inv else np.fft.fft(z)
        assert np.allclose(got, want, atol=1e-6)
    assert np.array_equal(O.fft_f32(x[:2], 1), x[:2])
    assert np.array_equal(O.fft_f32(x[:4], 2), np.array([4, 6, -2, -2], np.float32))


# ---- the FP64 transform --------------------------------------------------------------------------

@pytest.mark.parametrize("L", range(0, 8))
def test_fft_f64_equals_definition(L):
    n = 1 << L
    x = O.uniform(100 + L, n).astype(np.float64)
    for inv in (False, True):
        assert rel(O.fft_f64(x, n, inv), O.dft_f64(x, n, inv)) < 1e-14


@pytest.mark.parametrize("L", [0, 1, 5, 10, 14, 17])
def test_fft_f64_equals_numpy(L):
    n = 1 << L
    x = O.normal(200 + L, n)
    z = cplx(x)
    assert rel(O.fft_f64(x, n), inter(np.fft.fft(z))) < 1e-14
    assert rel(O.fft_f64(x, n, True), inter(np.fft.ifft(z) * n)) < 1e-14


def test_dft_definition_against_numpy():
    for L in range(0, 7):
        n = 1 << L
        x = O.uniform(300 + L, n).astype(np.float64)
        assert rel(O.dft_f64(x, n), inter(np.fft.fft(cplx(x)))) < 1e-14


def test_known_answers_impulse_constant_tone():
    for L in (3, 8, 12):
        n = 1 << L
        j = 1 + (n // 3) % (n - 1)
        x = np.zeros(2 * n, np.float32)
        x[2 * j] = 1
        k = np.arange(n)
        want = np.exp(-2j * np.pi * ((j * k) % n) / n)
        assert np.abs(cplx(O.fft_f64(x, n)) - want).max() < 1e-13
        x = np.tile(np.array([0.5, -0.25], np.float32), n)
        y = cplx(O.fft_f64(x, n))
        assert abs(y[0] - (0.5 - 0.25j) * n) < 1e-12 and np.abs(y[1:]).max() < 1e-12
        k0 = round(n / 7)
        t = O.add_tone(np.zeros(2 * n, np.float32), n, k0)
        y = cplx(O.fft_f64(t, n))
        assert abs(y[k0] - n) < 1e-4 * n and np.abs(np.delete(y, k0)).max() < 1e-4 * n


def test_parseval_and_linearity():
    n = 1 << 11
    a, b = O.normal(1, n), O.normal(2, n)
    fa, fb = O.fft_f64(a, n), O.fft_f64(b, n)
    assert abs((fa ** 2).sum() - n * (a.astype(np.float64) ** 2).sum()) < 1e-10 * (fa ** 2).sum()
    s = (a.astype(np.float64) + 2.0 * b)
    assert rel(O.fft_f64(s, n), fa + 2.0 * fb) < 1e-14


def test_dft_bins_match_full_transform():
    n = 1 << 12
    x = O.uniform(9, n)
    full = O.fft_f64(x, n, True)
    bins = [0, 1, 7, 585, n - 1]
    got = O.dft_bins_f64(x, n, bins, True)
    want = np.concatenate([full[2 * b:2 * b + 2] for b in bins])
    assert rel(got, want) < 1e-13


# ---- generators ----------------------------------------------------------------------------------

def test_generators_match_vectors(golden):
    for s in range(5):
        assert np.array_equal(O.uniform(s, 32), golden[f"uniform{s}"])
        assert np.array_equal(O.normal(s, 32), golden[f"normal{s}"])
    first = O.xorshift(2 * 1)
    assert np.array_equal(first, golden["x0"])


def test_generator_statistics():
    u = O.uniform(0, 1 << 16)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01
    assert np.array_equal(u, (np.round(u * 2 ** 23) / 2 ** 23).astype(np.float32))  # 24-bit grid
    g = O.normal(1, 1 << 16).astype(np.float64)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1) < 0.02 and abs((g ** 3).mean()) < 0.05


def _xorshift_step(state):
    """One step of the oracle's xorshift generator: (next state, output), read from the function's
    own state update (oracle_xorshift_fill writes the advanced state back)."""
    import ctypes

    st = ctypes.c_uint32(state)
    out = np.empty(1, np.float32)
    O.lib().oracle_xorshift_fill(ctypes.byref(st), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1)
    return st.value, float(out[0])


def _gf2_matmul(a, b):
    """32 x 32 bit matrices over GF(2), as lists of column words: (a b) e_i = a (b e_i)."""
    def apply(m, v):
        r, i = 0, 0
        while v:
            if v & 1:
                r ^= m[i]
            v >>= 1
            i += 1
        return r
    return [apply(a, col) for col in b]


def _gf2_pow(m, e):
    result = [1 << i for i in range(32)]
    while e:
        if e & 1:
            result = _gf2_matmul(m, result)
        m = _gf2_matmul(m, m)
        e >>= 1
    return result


def test_xorshift_structure():
    """oracle_xorshift_fill (the reference tests' xorshift32, tests/test_onefft.c:57-61) pinned by
    properties, not by a second copy of its formula: its state transition T is linear over GF(2);
    T has the full period 2^32 - 1 (its bit matrix has that multiplicative order, the defining
    property of Marsaglia's shift triple 13, 17, 5), so no state but 0 repeats early; each output is
    the new state read as a signed 32-bit integer times 2^-31, exactly where that is representable;
    and filling in pieces continues the same stream."""
    identity = [1 << i for i in range(32)]
    cols = [_xorshift_step(1 << i)[0] for i in range(32)]
    rng = np.random.default_rng(5)
    for a, b in rng.integers(0, 2**32, (200, 2), dtype=np.uint64).tolist():
        assert _xorshift_step(a ^ b)[0] == _xorshift_step(a)[0] ^ _xorshift_step(b)[0]
        # the bit matrix reproduces T
        v, r, i = a, 0, 0
        while v:
            if v & 1:
                r ^= cols[i]
            v >>= 1
            i += 1
        assert r == _xorshift_step(a)[0]
    order = 2**32 - 1  # = 3 * 5 * 17 * 257 * 65537
    assert _gf2_pow(cols, order) == identity
    for p in (3, 5, 17, 257, 65537):
        assert _gf2_pow(cols, order // p) != identity, p
    # outputs: state / 2^31, exact for |state| < 2^24 as a signed integer
    for s0 in rng.integers(0, 2**32, 20000, dtype=np.uint64).tolist():
        s1, out = _xorshift_step(s0)
        signed = s1 - 2**32 if s1 >= 2**31 else s1
        assert -1.0 <= out < 1.0
        if abs(signed) < 2**24:
            assert out * 2.0**31 == signed
        else:
            assert abs(out * 2.0**31 - signed) <= abs(signed) * 2.0**-24
    # the default stream, in pieces and at once
    whole = O.xorshift(1000)
    assert np.array_equal(whole[:2], O.xorshift(2))
    stats = whole.astype(np.float64)
    assert abs(stats.mean()) < 0.05 and abs(stats.var() - 1 / 3) < 0.03


# ---- planted mistakes must be caught -------------------------------------------------------------

def test_planted_mistakes_fail_the_checks(golden):
    L, n = 10, 1 << 10
    x = golden[f"x{L}"]
    good = O.fft_f32(x, n)
    want = O.fft_f64(x, n)
    assert O.rel_rms(good, want) < tol_fp64(L)
    z = cplx(good)
    mistakes = {
        "wrong sign (inverse)": inter(np.conj(cplx(O.fft_f32(np.ascontiguousarray(inter(np.conj(cplx(x))).astype(np.float32)), n)))),
        "conjugated output": inter(np.conj(z)),
        "index off by one": inter(np.roll(z, 1)),
        "transposed halves": inter(np.concatenate([z[n // 2:], z[:n // 2]])),
        "one twiddle dropped": inter(z + np.where(np.arange(n) == 37, z[37] * (np.exp(-2j * np.pi / n) - 1), 0)),
    }
    for name, bad in mistakes.items():
        bad = bad.astype(np.float32)
        assert not np.array_equal(bad, golden[f"port_f{L}"]), name
        if name != "wrong sign (inverse)":
            assert O.rel_rms(bad, want) > tol_fp64(L), name

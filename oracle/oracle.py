"""ctypes view of the CPU oracle. TEST INFRASTRUCTURE ONLY.

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import this
module; the product library never does. It loads

* ``oracle/liboracle.so``: the C restatement in ``oracle/onefft_oracle.c``;
* ``oracle/_ref/libref_onefft.so``: the reference's own transform built from its sources (its
  run-time AVX-512/AVX2/SSE2 selection), public symbols renamed ``ref_*``;
* ``oracle/_ref/libref_onefft_portable.so``: the reference's portable (ONEFFT_NO_SIMD) path.

The ``_ref`` libraries exist only where ``oracle/Makefile`` could read the reference sources when
it ran; ``ref()``/``ref_portable()`` return None otherwise.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_size_t, c_uint, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_F = POINTER(c_float)
_D = POINTER(c_double)

_lib = None
_ref = {}


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def lib():
    """The C restatement (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.oracle_root_double.argtypes = [c_uint64, c_uint, _D, _D]
        L.oracle_root.argtypes = [c_uint64, c_uint, _F, _F]
        L.oracle_bit_reverse.argtypes = [c_size_t, c_uint]
        L.oracle_bit_reverse.restype = c_size_t
        L.oracle_fft_f32.argtypes = [_F, _F, c_size_t, c_int]
        L.oracle_fft_f64.argtypes = [_D, _D, c_size_t, c_int]
        L.oracle_dft_f64.argtypes = [_D, _D, c_size_t, c_int]
        L.oracle_dft_bins_f64.argtypes = [_F, c_size_t, c_int, POINTER(c_uint64), c_size_t, _D]
        L.oracle_xorshift_fill.argtypes = [POINTER(c_uint32), _F, c_size_t]
        L.oracle_uniform_fill.argtypes = [c_uint64, _F, c_size_t]
        L.oracle_normal_fill.argtypes = [c_uint64, _F, c_size_t]
        L.oracle_add_tone.argtypes = [_F, c_size_t, c_uint64]
        L.oracle_rel_rms.argtypes = [_F, _D, c_size_t]
        L.oracle_rel_rms.restype = c_double
        L.oracle_rel_rms_f32.argtypes = [_F, _F, c_size_t]
        L.oracle_rel_rms_f32.restype = c_double
        L.oracle_fft_f32_batch.argtypes = [_F, _F, c_size_t, c_size_t, c_size_t, c_int]
        L.oracle_fft_f64_from_f32_batch.argtypes = [_F, _D, c_size_t, c_size_t, c_size_t, c_int]
        _lib = L
    return _lib


def _load_ref(name):
    if name in _ref:
        return _ref[name]
    path = os.path.join(HERE, "_ref", name)
    L = None
    if os.path.exists(path):
        L = ctypes.CDLL(path)
        L.ref_onefft_plan_create.argtypes = [c_size_t]
        L.ref_onefft_plan_create.restype = c_void_p
        L.ref_onefft_plan_destroy.argtypes = [c_void_p]
        L.ref_onefft_plan_forward.argtypes = [c_void_p, _F, _F]
        L.ref_onefft_plan_inverse.argtypes = [c_void_p, _F, _F]
        L.ref_onefft_forward.argtypes = [_F, _F, c_size_t]
        L.ref_onefft_inverse.argtypes = [_F, _F, c_size_t]
        L.ref_onefft_implementation.restype = ctypes.c_char_p
        L.ref_root.argtypes = [c_uint64, c_uint, _F, _F]
        L.ref_root_double.argtypes = [c_uint64, c_uint, _D, _D]
        L.ref_batch.argtypes = [_F, _F, c_size_t, c_size_t, c_size_t, c_int, c_int]
        L.ref_time_batch.argtypes = [_F, _F, c_size_t, c_size_t, c_size_t, c_int, c_int, c_double, c_int]
        L.ref_time_batch.restype = c_double
        L.ref_ragged.argtypes = [_F, _F, c_void_p, c_size_t, c_int, c_int]
        L.ref_ragged.restype = c_double
    _ref[name] = L
    return L


def ref():
    """The reference's own CPU transform (SIMD build), or None when it was not built."""
    return _load_ref("libref_onefft.so")


def ref_portable():
    """The reference's portable path (ONEFFT_NO_SIMD), or None when it was not built."""
    return _load_ref("libref_onefft_portable.so")


# ---- inputs --------------------------------------------------------------------------------------

def uniform(seed: int, n_complex: int) -> np.ndarray:
    """Interleaved complex64 data as float32[2n], re/im i.i.d. uniform [-1, 1) from SplitMix64(seed)."""
    out = np.empty(2 * n_complex, np.float32)
    lib().oracle_uniform_fill(seed, _fp(out), out.size)
    return out


def normal(seed: int, n_complex: int) -> np.ndarray:
    """Interleaved float32[2n], re/im i.i.d. standard normal (Box-Muller, FP64, rounded once)."""
    out = np.empty(2 * n_complex, np.float32)
    lib().oracle_normal_fill(seed, _fp(out), out.size)
    return out


def xorshift(count: int, state: int = 0x9E3779B9) -> np.ndarray:
    """The reference tests' xorshift32 stream, uniform [-1, 1)."""
    s = c_uint32(state)
    out = np.empty(count, np.float32)
    lib().oracle_xorshift_fill(ctypes.byref(s), _fp(out), count)
    return out


def add_tone(x: np.ndarray, n: int, k0: int) -> np.ndarray:
    """x += exp(+2 pi i k0 j / n), in FP64, rounded once (in place; returns x)."""
    assert lib().oracle_add_tone(_fp(x), n, k0) == 0
    return x


# ---- transforms ----------------------------------------------------------------------------------

def fft_f32(x: np.ndarray, n: int, inverse: bool = False, batch: int = 1) -> np.ndarray:
    """The restated portable single-precision transform, batched over contiguous transforms."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    assert lib().oracle_fft_f32_batch(_fp(x), _fp(out), n, batch, n, int(inverse)) == 0
    return out


def fft_f64(x: np.ndarray, n: int, inverse: bool = False, batch: int = 1) -> np.ndarray:
    """FP64 transform of FP32 or FP64 interleaved input (float64[2n*batch])."""
    if x.dtype == np.float32:
        x = np.ascontiguousarray(x)
        out = np.empty(2 * n * batch, np.float64)
        assert lib().oracle_fft_f64_from_f32_batch(_fp(x), _dp(out), n, batch, n, int(inverse)) == 0
        return out
    out = np.ascontiguousarray(x, np.float64).copy()
    for b in range(batch):
        seg = out[2 * n * b: 2 * n * (b + 1)]
        assert lib().oracle_fft_f64(_dp(seg), _dp(seg), n, int(inverse)) == 0
    return out


def dft_f64(x: np.ndarray, n: int, inverse: bool = False) -> np.ndarray:
    """Brute-force O(n^2) FP64 DFT of one transform."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(2 * n, np.float64)
    assert lib().oracle_dft_f64(_dp(x), _dp(out), n, int(inverse)) == 0
    return out


def dft_bins_f64(x: np.ndarray, n: int, bins, inverse: bool = False) -> np.ndarray:
    """Selected FP64 output bins of one transform of FP32 input, O(n) per bin."""
    x = np.ascontiguousarray(x, np.float32)
    b = np.ascontiguousarray(np.asarray(bins, np.uint64))
    out = np.empty(2 * b.size, np.float64)
    assert lib().oracle_dft_bins_f64(_fp(x), n, int(inverse), b.ctypes.data_as(POINTER(c_uint64)),
                                     b.size, _dp(out)) == 0
    return out


def root(k: int, log2size: int):
    re, im = c_float(), c_float()
    lib().oracle_root(k, log2size, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def root_double(k: int, log2size: int):
    re, im = c_double(), c_double()
    lib().oracle_root_double(k, log2size, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def rel_rms(got: np.ndarray, want: np.ndarray) -> float:
    """sqrt(sum (got-want)^2 / sum want^2), want in float64 or float32 (taken as exact)."""
    got = np.ascontiguousarray(got, np.float32)
    if want.dtype == np.float64:
        return lib().oracle_rel_rms(_fp(got), _dp(np.ascontiguousarray(want)), got.size)
    return lib().oracle_rel_rms_f32(_fp(got), _fp(np.ascontiguousarray(want, np.float32)), got.size)


# ---- the reference itself ------------------------------------------------------------------------

def ref_fft(x: np.ndarray, n: int, inverse: bool = False, batch: int = 1, threads: int = 1,
            portable: bool = False) -> np.ndarray:
    """The reference's own CPU transform over a contiguous batch, or None if it was not built."""
    L = ref_portable() if portable else ref()
    if L is None:
        return None
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    assert L.ref_batch(_fp(x), _fp(out), n, batch, n, int(inverse), threads) == 0
    return out


def ref_root(k: int, log2size: int, portable: bool = False):
    L = ref_portable() if portable else ref()
    if L is None:
        return None
    re, im = c_float(), c_float()
    L.ref_root(k, log2size, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def ref_time_batch(x: np.ndarray, n: int, batch: int, threads: int, inverse: bool = False,
                   sample_seconds: float = 0.2, samples: int = 5) -> float:
    """Seconds per pass over the batch on `threads` threads sharing one reference plan."""
    L = ref()
    if L is None:
        return float("nan")
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    return L.ref_time_batch(_fp(x), _fp(out), n, batch, n, int(inverse), threads, sample_seconds, samples)


def _items_array(items):
    arr = np.zeros(max(len(items), 1), dtype=[("offset", "<u8"), ("log2n", "<u4"), ("inverse", "<u4")])
    for i, (off, l, inv) in enumerate(items):
        arr[i] = (off, l, inv)
    return arr


def ref_ragged(x: np.ndarray, items, threads: int = 1, reps: int = 1, out: np.ndarray | None = None):
    """The reference's transform of every (offset, log2n, inverse) item of a ragged batch, on
    `threads` threads sharing its plan cache. Returns (output, seconds of the fastest pass), or
    None when the reference was not built. Gaps between items keep `out`'s values."""
    L = ref()
    if L is None:
        return None
    x = np.ascontiguousarray(x, np.float32)
    if out is None:
        out = np.zeros_like(x)
    arr = _items_array(items)
    secs = L.ref_ragged(_fp(x), _fp(out), arr.ctypes.data, len(items), threads, reps)
    assert secs >= 0, "reference ragged run failed"
    return out, secs


def ref_implementation() -> str | None:
    L = ref()
    return None if L is None else L.ref_onefft_implementation().decode()


def ref_online_cpus() -> int:
    L = ref()
    return 0 if L is None else L.ref_online_cpus()

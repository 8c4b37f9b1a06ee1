"""Python view of the OneFFT gfx950 library (``libonefft_amd.so``) through its C ABI.

The product is the C library declared in ``include/onefft.h``; this module is the thin binding
the tests and the benchmark use:

* ``lib()`` — the ctypes handle with argument types set for every exported function;
* ``Plan`` — owns an ``onefft_plan`` (reference ``include/onefft.h:88-109``);
* ``fft(x, inverse=False)`` — transforms a complex64 CUDA tensor along its last dimension through
  the batched device entry point on torch's current stream;
* ``fft_autograd(x, inverse=False)`` — the same transform as the torch custom op ``onefft::fft``
  (``onefft_amd/torch_op.py``): differentiable, with fake-tensor support for torch.compile; its
  backward is the adjoint: the adjoint of the unnormalised forward DFT is the unnormalised
  inverse (and vice versa), run by the same kernels and tables.

There is no fallback: if the library is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_size_t, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libonefft_amd.so")

# Every function include/onefft.h declares, with its C signature.
_F = POINTER(c_float)
SIGNATURES = {
    "onefft_forward": (c_int, [c_void_p, c_void_p, c_size_t]),
    "onefft_inverse": (c_int, [c_void_p, c_void_p, c_size_t]),
    "onefft_cache_prepare": (c_int, [c_size_t]),
    "onefft_cache_clear": (None, []),
    "onefft_plan_create": (c_void_p, [c_size_t]),
    "onefft_plan_destroy": (None, [c_void_p]),
    "onefft_plan_forward": (None, [c_void_p, c_void_p, c_void_p]),
    "onefft_plan_inverse": (None, [c_void_p, c_void_p, c_void_p]),
    "onefft_plan_size": (c_size_t, [c_void_p]),
    "onefft_implementation": (c_char_p, []),
    "onefft_last_error": (c_char_p, []),
    "onefft_plan_forward_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "onefft_plan_inverse_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "onefft_plan_forward_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_size_t, c_void_p]),
    "onefft_plan_inverse_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_size_t, c_void_p]),
    "onefft_ragged_plan_create": (c_void_p, [c_void_p, c_size_t]),
    "onefft_ragged_plan_destroy": (None, [c_void_p]),
    "onefft_ragged_extent": (c_size_t, [c_void_p]),
    "onefft_ragged_task_count": (c_size_t, [c_void_p]),
    "onefft_ragged_execute": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "onefft_ragged_execute_per_item": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "onefft_ragged_wait": (c_int, [c_void_p, c_void_p]),
    "onefft_ragged_execute_items": (c_int, [c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p]),
    "onefft_graph_create_batch": (c_void_p, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_size_t, c_int]),
    "onefft_graph_create_ragged": (c_void_p, [c_void_p, c_void_p, c_void_p]),
    "onefft_graph_launch": (c_int, [c_void_p, c_void_p]),
    "onefft_graph_destroy": (None, [c_void_p]),
}


class RaggedItem(ctypes.Structure):
    """onefft_ragged_item: (offset in complex values, log2 n, inverse)."""

    _fields_ = [("offset", ctypes.c_uint64), ("log2n", ctypes.c_uint32), ("inverse", ctypes.c_uint32)]

_lib = None


class OneFFTError(RuntimeError):
    pass


def lib():
    """The loaded library. Raises if it has not been built (``python -m onefft_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OneFFTError(f"{LIB_PATH} is missing: build it with `python -m onefft_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().onefft_last_error().decode()


def implementation() -> str:
    return lib().onefft_implementation().decode()


def _check(status: int, what: str) -> None:
    if status != 0:
        raise OneFFTError(f"{what} failed: {last_error()}")


class Plan:
    """An onefft_plan for n points (both directions)."""

    def __init__(self, n: int):
        self.n = int(n)
        self._p = lib().onefft_plan_create(self.n)
        if not self._p:
            raise OneFFTError(f"onefft_plan_create({n}) returned NULL: {last_error()}")

    @property
    def handle(self) -> int:
        return self._p

    def close(self) -> None:
        if self._p:
            lib().onefft_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # synchronous drop-in calls: host (numpy) or device pointers
    def forward_ptr(self, src: int, dst: int) -> None:
        lib().onefft_plan_forward(self._p, src, dst)

    def inverse_ptr(self, src: int, dst: int) -> None:
        lib().onefft_plan_inverse(self._p, src, dst)

    # asynchronous device calls
    def batch_ptr(self, src: int, dst: int, batch: int, in_dist: int, out_dist: int, inverse: bool,
                  stream: int = 0) -> None:
        fn = lib().onefft_plan_inverse_batch if inverse else lib().onefft_plan_forward_batch
        _check(fn(self._p, src, dst, batch, in_dist, out_dist, stream), "onefft_plan_*_batch")


class RaggedPlan:
    """An onefft_ragged_plan over items given as (offset, log2n, inverse) triples."""

    def __init__(self, items):
        import numpy as np

        arr = np.asarray(items, dtype=np.uint64).reshape(-1, 3)
        self.count = arr.shape[0]
        buf = (RaggedItem * max(self.count, 1))()
        for i, (off, l, inv) in enumerate(arr.tolist()):
            buf[i] = RaggedItem(off, l, inv)
        self._items = buf
        self._p = lib().onefft_ragged_plan_create(ctypes.addressof(buf), self.count)
        if not self._p:
            raise OneFFTError(f"onefft_ragged_plan_create failed: {last_error()}")

    @property
    def extent(self) -> int:
        return lib().onefft_ragged_extent(self._p)

    @property
    def tasks(self) -> int:
        return lib().onefft_ragged_task_count(self._p)

    def execute(self, src: int, dst: int, stream: int = 0) -> None:
        _check(lib().onefft_ragged_execute(self._p, src, dst, stream), "onefft_ragged_execute")

    def execute_per_item(self, src: int, dst: int, stream: int = 0) -> None:
        _check(lib().onefft_ragged_execute_per_item(self._p, src, dst, stream), "onefft_ragged_execute_per_item")

    def wait(self, stream: int = 0) -> None:
        _check(lib().onefft_ragged_wait(self._p, stream), "onefft_ragged_wait")

    def close(self) -> None:
        if self._p:
            lib().onefft_ragged_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ragged_items(seed: int, count: int, lo: int = 4, hi: int = 20, max_bytes: int | None = None):
    """The ragged workload's items: log2 n uniform on lo..hi and the direction a fair coin, both
    from SplitMix64(seed) (two draws per item), packed back to back. With max_bytes, the longest
    prefix whose input fits in max_bytes. Returns a list of (offset, log2n, inverse)."""
    items, off, s = [], 0, seed & (2**64 - 1)
    M = 2**64 - 1
    for _ in range(count):
        vals = []
        for _ in range(2):
            s = (s + 0x9E3779B97F4A7C15) & M
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
            vals.append(z ^ (z >> 31))
        log2n = lo + int((vals[0] >> 11) * (hi - lo + 1) >> 53)
        inverse = int(vals[1] >> 63)
        if max_bytes is not None and (off + (1 << log2n)) * 8 > max_bytes:
            break
        items.append((off, log2n, inverse))
        off += 1 << log2n
    return items


class Graph:
    """An onefft_graph: one recorded plan execution, replayed with launch()."""

    def __init__(self, handle):
        if not handle:
            raise OneFFTError(f"graph creation failed: {last_error()}")
        self._g = handle

    @classmethod
    def batch(cls, plan: Plan, src: int, dst: int, batch: int, in_dist: int, out_dist: int, inverse: bool):
        return cls(lib().onefft_graph_create_batch(plan.handle, src, dst, batch, in_dist, out_dist, int(inverse)))

    @classmethod
    def ragged(cls, plan: "RaggedPlan", src: int, dst: int):
        return cls(lib().onefft_graph_create_ragged(plan._p, src, dst))

    def launch(self, stream: int = 0) -> None:
        _check(lib().onefft_graph_launch(self._g, stream), "onefft_graph_launch")

    def close(self) -> None:
        if self._g:
            lib().onefft_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_plans: dict[tuple, Plan] = {}


def plan(n: int, device: int | None = None) -> Plan:
    """A process-wide plan per (device, size), created on first use like the library's own cache.
    `device` None means HIP's current device; a plan lives on the device current at its creation."""
    key = (device, n)
    p = _plans.get(key)
    if p is None:
        if device is None:
            p = Plan(n)
        else:
            import torch

            with torch.cuda.device(device):
                p = Plan(n)
        p = _plans.setdefault(key, p)
    return p


def _stream_ptr(torch, device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def fft(x, inverse: bool = False, out=None):
    """Unnormalised DFT of a complex64 CUDA tensor along its last dimension (power-of-two length),
    asynchronous on torch's current stream. The result is a new contiguous tensor (or `out`, which
    must be a contiguous complex64 tensor of x's shape on x's device)."""
    import torch

    if x.dtype != torch.complex64 or not x.is_cuda:
        raise OneFFTError("fft expects a complex64 CUDA tensor")
    # materialise lazy conjugation / negation: the kernels read the tensor's memory as stored
    x = x.resolve_conj().resolve_neg().contiguous()
    n = x.shape[-1] if x.dim() else 1
    batch = x.numel() // n if n else 0
    if out is None:
        out = torch.empty_like(x)
    elif (out.shape != x.shape or out.dtype != x.dtype or out.device != x.device or not out.is_contiguous()
          or out.is_conj() or out.is_neg()):
        raise OneFFTError("out must be a contiguous complex64 tensor of the input's shape and device")
    if batch == 0:
        return out
    plan(n, x.device.index).batch_ptr(x.data_ptr(), out.data_ptr(), batch, n, n, inverse,
                                      _stream_ptr(torch, x.device))
    return out


def fft_autograd(x, inverse: bool = False):
    """Differentiable unnormalised FFT along the last dimension: the torch custom op onefft::fft."""
    from . import torch_op

    return torch_op.fft_op(x, bool(inverse))

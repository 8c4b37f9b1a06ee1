"""The transform as a torch custom operator, ``torch.ops.onefft.fft(Tensor x, bool inverse)``.

The operator calls the library's batched device entry point through the C ABI (``onefft.fft``)
and registers with torch what lets it compose with the rest of torch:

* a fake (meta) kernel: output shape, dtype, device and contiguous strides without running
  anything, so the operator traces under FakeTensorMode and torch.compile;
* an autograd formula: the adjoint of the unnormalised forward DFT is the unnormalised inverse
  DFT (and vice versa), so the backward of one direction is the opposite direction, run by the
  same kernels and tables (reference inverse definition: src/onefft.c:86-88).

The output is always a new contiguous tensor (never an alias of the input). There is no CPU
kernel: a CPU tensor raises.
"""
from __future__ import annotations

import torch

from . import fft as _fft

_LIB = "onefft::fft"


def _is_pow2(n) -> bool:
    return n > 0 and (n & (n - 1)) == 0


@torch.library.custom_op(_LIB, mutates_args=(), device_types="cuda")
def fft_op(x: torch.Tensor, inverse: bool = False) -> torch.Tensor:
    """Unnormalised DFT along the last dimension of a complex64 CUDA tensor (inverse: the
    unnormalised inverse DFT, forward then inverse = n x)."""
    return _fft(x, inverse)


@fft_op.register_fake
def _fft_fake(x: torch.Tensor, inverse: bool = False) -> torch.Tensor:
    torch._check(x.dtype == torch.complex64, lambda: f"onefft::fft expects complex64, got {x.dtype}")
    n = x.shape[-1] if x.dim() else 1
    if isinstance(n, int):  # a symbolic length is checked when the kernel runs
        torch._check(_is_pow2(n) or x.numel() == 0,
                     lambda: f"onefft::fft needs a power-of-two last dimension, got {n}")
    return torch.empty(x.shape, dtype=x.dtype, device=x.device)


def _setup_context(ctx, inputs, output):
    _, inverse = inputs
    ctx.inverse = bool(inverse)


def _backward(ctx, grad):
    # y = F x  =>  dL/dx = F^H dL/dy (conjugate Wirtinger convention), F^H = the other direction
    return fft_op(grad, not ctx.inverse), None


fft_op.register_autograd(_backward, setup_context=_setup_context)

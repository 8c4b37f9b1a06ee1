"""The bench's copy kernel (onefft_amd/libonefft_stream.so, the 4096-point kernel's access pattern
without arithmetic): it must copy exactly, leave the bytes past the copied range alone, and reject
sizes that are not whole 32 KiB blocks, or its rate would not be a fair ceiling for the FFT."""
import ctypes
import os

import pytest
import torch

from conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "onefft_amd", "libonefft_stream.so")


def lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is not built (python -m onefft_amd.build)")
    L = ctypes.CDLL(LIB)
    L.onefft_stream_copy.restype = ctypes.c_int
    L.onefft_stream_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_void_p]
    return L


@pytest.mark.parametrize("nt", [0, 1])
def test_stream_copy_exact(nt):
    print(f"stream copy nt={nt}", flush=True)
    L = lib()
    dev = torch.device("cuda:0")
    blocks = 37
    n = blocks * 4096
    x = torch.randn(n + 4096, dtype=torch.complex64, device=dev)
    y = torch.full_like(x, complex(7.0, -7.0))
    sp = torch.cuda.current_stream().cuda_stream
    assert L.onefft_stream_copy(x.data_ptr(), y.data_ptr(), n * 8, nt, sp) == 0
    torch.cuda.synchronize()
    assert torch.equal(y[:n], x[:n])
    assert bool((y[n:] == complex(7.0, -7.0)).all())


def test_stream_copy_rejects_partial_blocks():
    L = lib()
    x = torch.zeros(8192, dtype=torch.complex64, device="cuda:0")
    y = torch.empty_like(x)
    sp = torch.cuda.current_stream().cuda_stream
    assert L.onefft_stream_copy(x.data_ptr(), y.data_ptr(), 0, 1, sp) == -1
    assert L.onefft_stream_copy(x.data_ptr(), y.data_ptr(), 4096 * 8 + 8, 1, sp) == -1
    assert L.onefft_stream_copy(x.data_ptr() + 4, y.data_ptr(), 4096 * 8, 1, sp) == -1

"""The largest transforms the API allows (log2 n <= 30, reference include/onefft.h:88-109: the
plan rejects larger sizes), checked through properties that do not need an O(n log n) oracle on
the host: a pure tone x_j = exp(2 pi i k0 j / n) (formed in FP64 with exact integer phase
reduction, rounded once to FP32) transforms to n at bin k0 and to the rounding floor elsewhere;
inverse(forward(x)) = n x; and Parseval, ||F x||^2 = n ||x||^2."""
import math

import pytest
import torch

import onefft_amd as F
from conftest import tol_fp64

pytestmark = pytest.mark.gpu


def stream():
    return torch.cuda.current_stream().cuda_stream


def tone(n, k0, dev):
    x = torch.empty(n, dtype=torch.complex64, device=dev)
    step = 1 << 26
    for s in range(0, n, step):
        j = torch.arange(s, min(n, s + step), dtype=torch.int64, device=dev)
        ph = ((j * k0) % n).to(torch.float64) * (2 * math.pi / n)
        x[s:s + j.numel()] = torch.complex(torch.cos(ph), torch.sin(ph)).to(torch.complex64)
    return x


@pytest.mark.parametrize("L", [28, 29, 30])
def test_tone_roundtrip_parseval(L):
    n = 1 << L
    print(f"max-size n=2^{L}", flush=True)
    dev = torch.device("cuda:0")
    k0 = int(round(n / 7))
    x = tone(n, k0, dev)
    y = torch.empty_like(x)
    p = F.Plan(n)
    p.batch_ptr(x.data_ptr(), y.data_ptr(), 1, n, n, False, stream())
    torch.cuda.synchronize()
    # the tone: n at k0, the FP32 rounding floor elsewhere
    peak = y[k0].clone()
    y[k0] = 0
    off = torch.linalg.vector_norm(y.to(torch.complex128)).item()
    assert abs(peak.item() - n) <= tol_fp64(L) * n, (L, peak.item())
    assert off <= tol_fp64(L) * n, (L, off / n)
    y[k0] = peak
    # Parseval
    ex = torch.linalg.vector_norm(x.to(torch.complex128)).item() ** 2
    ey = torch.linalg.vector_norm(y.to(torch.complex128)).item() ** 2
    assert abs(ey - n * ex) <= 2 * tol_fp64(L) * n * ex
    # round trip: the inverse of the transform is n x
    z = torch.empty_like(x)
    p.batch_ptr(y.data_ptr(), z.data_ptr(), 1, n, n, True, stream())
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm((z.to(torch.complex128) / n - x.to(torch.complex128))) /
           torch.linalg.vector_norm(x.to(torch.complex128))).item()
    assert err < 2 * tol_fp64(L), (L, err)
    p.close()
    del x, y, z
    torch.cuda.empty_cache()

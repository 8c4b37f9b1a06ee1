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


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("inverse", [False, True])
def test_multipass_batch_across_launch_chunks(monkeypatch, inverse, wide):
    """A multi-pass batch larger than one launch's grid holds (65535 transforms: the batch index
    is the grid's y dimension) runs in chunks; the transforms on both sides of the chunk boundary,
    with input and output distances that differ from the size and from each other, match FP64.
    65537 transforms of 2^15 points: 16 GiB in and 16 GiB out. `wide`: the plan's split is forced
    to 2^7 x 2^8 with 64 columns per tile and the halves layout, so the input pass runs 16384-point
    tiles (1024-thread workgroups)."""
    import numpy as np

    L, batch = 15, 65537
    n = 1 << L
    in_dist, out_dist = n + 8, n + 24
    print(f"chunked batch n=2^{L} x{batch} inverse={inverse} wide={wide}", flush=True)
    if wide:
        # the library accepts these column counts (else it would silently fall back to defaults)
        from onefft_amd.tune import pass_cols_ok

        assert pass_cols_ok(7, 64, 2) and pass_cols_ok(8, 64, 2)
        assert not pass_cols_ok(8, 64, 0), "64 columns of 2^8 points need the halves layout's tiles"
        monkeypatch.setenv("ONEFFT_PASS_SPLIT", "7,8")
        monkeypatch.setenv("ONEFFT_PASS_COLS", "64,64")
        monkeypatch.setenv("ONEFFT_PASS_LAYOUT", "2")
    plan = F.Plan(n)  # a plan of its own: the environment is read at creation
    for k in ("ONEFFT_PASS_SPLIT", "ONEFFT_PASS_COLS", "ONEFFT_PASS_LAYOUT"):
        monkeypatch.delenv(k, raising=False)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    x = torch.randn(batch * in_dist, dtype=torch.complex64, device=dev, generator=g)
    y = torch.zeros(batch * out_dist, dtype=torch.complex64, device=dev)
    plan.batch_ptr(x.data_ptr(), y.data_ptr(), batch, in_dist, out_dist, inverse, stream())
    torch.cuda.synchronize()
    plan.close()
    for b in (0, 1, 65534, 65535, 65536):
        xs = x[b * in_dist: b * in_dist + n].cpu().numpy().astype(np.complex128)
        want = np.fft.ifft(xs) * n if inverse else np.fft.fft(xs)
        got = y[b * out_dist: b * out_dist + n].cpu().numpy().astype(np.complex128)
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < tol_fp64(L), (b, err)
        # the gap after each output is untouched
        assert bool((y[b * out_dist + n: (b + 1) * out_dist] == 0).all()), b
    del x, y
    torch.cuda.empty_cache()

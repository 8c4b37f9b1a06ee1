"""This library's transform against rocFFT (torch.fft.fft on ROCm) on the same device input: a
check by an independent implementation beside the FP64 oracle. Both compute in FP32, so their
difference is bounded by the sum of their errors; the bar is twice the parity bar against FP64."""
import pytest

from conftest import tol_fp64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log2n,batch", [(0, 7), (1, 5), (2, 9), (5, 33), (10, 64), (12, 256), (14, 8),
                                         (15, 4), (18, 2), (20, 1), (22, 1)])
@pytest.mark.parametrize("inverse", [False, True])
def test_agrees_with_rocfft(log2n, batch, inverse):
    import torch

    import onefft_amd as F

    n = 1 << log2n
    g = torch.Generator(device="cuda").manual_seed(500 + log2n)
    x = torch.randn(batch, n, dtype=torch.complex64, device="cuda", generator=g)
    y = torch.empty_like(x)
    F.Plan(n).batch_ptr(x.data_ptr(), y.data_ptr(), batch, n, n, inverse,
                        torch.cuda.current_stream().cuda_stream)
    # torch's inverse is normalised by 1/n; this library's is not
    want = torch.fft.ifft(x, dim=-1) * n if inverse else torch.fft.fft(x, dim=-1)
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm(y - want) / torch.linalg.vector_norm(want)).item()
    assert err < 2 * tol_fp64(log2n), (log2n, batch, inverse, err)

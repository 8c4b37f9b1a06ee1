"""The torch custom operator onefft::fft on the GPU: the same bits as the C ABI batch call,
torch's operator checks (schema, fake kernel, autograd registration), torch.compile, CUDA-graph
capture, lazily conjugated and strided inputs. (The spectral-filter gradient tests in
test_gpu_adjoint.py also run through this operator.)"""
import numpy as np
import pytest
import torch

import onefft_amd as F
from onefft_amd import torch_op
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def cx(seed, *shape):
    n = int(np.prod(shape))
    return torch.from_numpy(O.normal(seed, n).view(np.complex64).reshape(shape)).cuda()


def test_same_bits_as_the_c_abi():
    for L, batch in ((0, 3), (5, 7), (12, 16), (17, 2)):
        n = 1 << L
        x = cx(70 + L, batch, n)
        for inv in (False, True):
            ref = torch.empty_like(x)
            F.plan(n).batch_ptr(x.data_ptr(), ref.data_ptr(), batch, n, n, inv, torch.cuda.current_stream().cuda_stream)
            assert torch.equal(torch.ops.onefft.fft(x, inv), ref), (L, inv)


def test_opcheck():
    x = cx(80, 4, 1024)
    torch.library.opcheck(torch.ops.onefft.fft.default, (x, False))
    torch.library.opcheck(torch.ops.onefft.fft.default, (x.clone().requires_grad_(True), True))


def test_lazy_conjugate_and_strided_inputs():
    x = cx(81, 64, 256)
    assert torch.equal(torch.ops.onefft.fft(x.conj(), False), torch.ops.onefft.fft(x.conj().resolve_conj(), False))
    xt = cx(82, 512, 32).t()  # 32 x 512, transforms strided in memory
    assert not xt.is_contiguous()
    assert torch.equal(torch.ops.onefft.fft(xt, True), torch.ops.onefft.fft(xt.contiguous(), True))


def test_gradient_is_the_adjoint():
    # <F x, y> = <x, F^H y>: the gradient of Re<F x, y> with respect to x is F^H y
    x = cx(83, 8, 4096).requires_grad_(True)
    y = cx(84, 8, 4096)
    for inv in (False, True):
        x.grad = None
        (torch.ops.onefft.fft(x, inv) * y.conj()).real.sum().backward()
        assert torch.equal(x.grad, torch.ops.onefft.fft(y, not inv))


def test_torch_compile_matches_eager():
    def f(x, h):
        y = torch.ops.onefft.fft(h * torch.ops.onefft.fft(x, False), True) / x.shape[-1]
        return (y.real ** 2 + y.imag ** 2).mean()

    x = cx(85, 16, 2048).requires_grad_(True)
    h = cx(86, 2048).requires_grad_(True)
    le = f(x, h)
    gx, gh = torch.autograd.grad(le, (x, h))
    fc = torch.compile(f, backend="aot_eager", fullgraph=True)
    lc = fc(x, h)
    cx_, ch = torch.autograd.grad(lc, (x, h))
    torch.testing.assert_close(lc, le, rtol=1e-6, atol=0)
    torch.testing.assert_close(cx_, gx, rtol=1e-5, atol=1e-9)
    torch.testing.assert_close(ch, gh, rtol=1e-5, atol=1e-9)


def test_cuda_graph_capture_replays():
    x = cx(87, 32, 4096)
    eager = torch.ops.onefft.fft(x, False)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up on the capture stream
        torch.ops.onefft.fft(x, False)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        y = torch.ops.onefft.fft(x, False)
    for replay in range(4):
        y.fill_(3.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), replay
        torch.ops.onefft.fft(x, True)  # eager work between replays
        torch.cuda.synchronize()

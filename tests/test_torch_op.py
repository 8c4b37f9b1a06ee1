"""The torch custom operator onefft::fft on the CPU: its fake kernel (shapes, dtypes, strides,
the same argument checks as the library) and tracing under fake tensors. The operator's numerics
and gradients run on the GPU (test_gpu_torch_op.py)."""
import pytest
import torch
from torch._subclasses.fake_tensor import FakeTensorMode
from torch.fx.experimental.proxy_tensor import make_fx

from onefft_amd import torch_op  # noqa: F401  (registers the operator)


def test_fake_kernel_shapes():
    with FakeTensorMode():
        for shape in [(1024,), (3, 4096), (2, 5, 64), (7, 1)]:
            x = torch.empty(shape, dtype=torch.complex64, device="cuda")
            for inv in (False, True):
                y = torch.ops.onefft.fft(x, inv)
                assert y.shape == x.shape and y.dtype == torch.complex64 and y.device == x.device
                assert y.is_contiguous()
        # a non-contiguous input still gives a contiguous output
        x = torch.empty(64, 8, dtype=torch.complex64, device="cuda").t()
        assert torch.ops.onefft.fft(x, False).is_contiguous()


@pytest.mark.parametrize("shape,dtype", [((3, 1000), torch.complex64), ((3, 1024), torch.complex128),
                                         ((3, 1024), torch.float32)])
def test_fake_kernel_rejects_what_the_library_rejects(shape, dtype):
    with FakeTensorMode():
        x = torch.empty(shape, dtype=dtype, device="cuda")
        with pytest.raises(RuntimeError):
            torch.ops.onefft.fft(x, False)


def test_no_cpu_kernel():
    with pytest.raises(NotImplementedError):
        torch.ops.onefft.fft(torch.zeros(4, 8, dtype=torch.complex64), False)


def test_traces_under_fake_tensors():
    def f(x):
        return torch.ops.onefft.fft(torch.ops.onefft.fft(x, False) * 2, True)

    with FakeTensorMode():
        x = torch.empty(2, 256, dtype=torch.complex64, device="cuda")
        gm = make_fx(f, tracing_mode="real")(x)
    calls = [n for n in gm.graph.nodes if n.op == "call_function" and "onefft" in str(n.target)]
    assert len(calls) == 2

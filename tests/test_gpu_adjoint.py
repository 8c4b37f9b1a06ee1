"""Adjoint kernels behind torch.autograd: the differentiable spectral filter workload.

x: 256 x 65536 complex64 (re/im N(0,1), seed 4); learnable filter H[65536] = 1 + 0.1 (a + i b),
a, b ~ N(0,1) (seed 5); target filter H* = 1 + 0.1 (c + i d) (seed 6);
t = inverse(H* . forward(x)) / n, y = inverse(H . forward(x)) / n, L = mean |y - t|^2.

The backward of the unnormalised forward transform is the unnormalised inverse (F^H) and vice
versa, run by the same kernels and tables. Checks:
* directional derivatives Re <g, v> against FP64 central differences of L along random v, for x
  and for H, within 1e-5 relative (L is quadratic in x and in H, so central differences carry
  no truncation error, only rounding);
* the whole gradients against gradients assembled by hand from the reference's CPU transform.
"""
import numpy as np
import pytest
import torch

import onefft_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

B, N = 256, 65536


def c64(x):
    """interleaved float32 -> complex64 array"""
    return x.view(np.complex64)


def setup():
    x = c64(O.normal(4, B * N)).reshape(B, N)
    H = (1 + 0.1 * c64(O.normal(5, N)).astype(np.complex128)).astype(np.complex64)
    Hs = (1 + 0.1 * c64(O.normal(6, N)).astype(np.complex128)).astype(np.complex64)
    return x, H, Hs


def loss64(x, H, t):
    """L in FP64 with numpy's float64 FFT."""
    y = np.fft.ifft(H.astype(np.complex128) * np.fft.fft(x.astype(np.complex128), axis=1), axis=1)  # = inverse/n
    d = y - t
    return float(np.mean(d.real ** 2 + d.imag ** 2))


def gpu_forward_backward(x, H, Hs):
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev).requires_grad_(True)
    Ht = torch.from_numpy(H).to(dev).requires_grad_(True)
    Hst = torch.from_numpy(Hs).to(dev)
    with torch.no_grad():
        t = F.fft(Hst * F.fft(xt.detach()), inverse=True) / N
    y = F.fft_autograd(Ht * F.fft_autograd(xt), inverse=True) / N
    d = y - t
    L = (d.real ** 2 + d.imag ** 2).mean()
    L.backward()
    torch.cuda.synchronize()
    return L.item(), xt.grad.cpu().numpy(), Ht.grad.cpu().numpy(), t.cpu().numpy()


def test_gradients_against_fp64_finite_differences():
    x, H, Hs = setup()
    L, gx, gH, t = gpu_forward_backward(x, H, Hs)
    t64 = np.fft.ifft(Hs.astype(np.complex128) * np.fft.fft(x.astype(np.complex128), axis=1), axis=1)
    assert abs(L - loss64(x, H, t64)) < 1e-5 * abs(L)
    rng = np.random.default_rng(11)
    for trial in range(2):
        v = (rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)).astype(np.complex128)
        eps = 1e-2
        fd = (loss64(x + eps * v, H, t64) - loss64(x - eps * v, H, t64)) / (2 * eps)
        ad = float(np.sum(np.conj(gx.astype(np.complex128)) * v).real)
        print(f"d/dx along v{trial}: relative difference to FP64 central difference {abs(ad - fd) / abs(fd):.3g}")
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)
        w = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex128)
        fd = (loss64(x, H + eps * w, t64) - loss64(x, H - eps * w, t64)) / (2 * eps)
        ad = float(np.sum(np.conj(gH.astype(np.complex128)) * w).real)
        print(f"d/dH along w{trial}: relative difference to FP64 central difference {abs(ad - fd) / abs(fd):.3g}")
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)


def test_gradients_against_reference_transform():
    if O.ref() is None:
        pytest.skip("reference not built")
    x, H, Hs = setup()
    L, gx, gH, t = gpu_forward_backward(x, H, Hs)

    def ref(z, inverse=False):
        flat = np.ascontiguousarray(z.astype(np.complex64)).view(np.float32).reshape(-1)
        return O.ref_fft(flat, N, inverse, batch=z.size // N, threads=16).view(np.complex64).reshape(z.shape)

    X = ref(x)
    y = ref(H * X, inverse=True) / N
    g_y = 2 * (y - t) / y.size                 # dL/dy (conjugate Wirtinger convention)
    g_u = ref(g_y) / N                          # y = F^H u / n  =>  g_u = F g_y / n
    gx_ref = ref(np.conj(H) * g_u, inverse=True)  # u = H . F x  =>  g_x = F^H (conj(H) g_u)
    gH_ref = np.sum(np.conj(X) * g_u, axis=0)
    for got, want in ((gx, gx_ref), (gH, gH_ref)):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"gradient rel-L2 against the reference-built gradient: {err:.3g}")
        assert err < 3e-7 * 16 * 4, err  # four FP32 transforms of 2^16 points on each side


def test_adjoint_identity():
    """<F x, y> = <x, F^H y> with F^H the unnormalised inverse, at several sizes."""
    rng = np.random.default_rng(3)
    for L in (4, 10, 12, 16, 20):
        n = 1 << L
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        y = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
        fx = F.fft(xt).cpu().numpy().astype(np.complex128)
        fhy = F.fft(yt, inverse=True).cpu().numpy().astype(np.complex128)
        lhs = np.vdot(y.astype(np.complex128), fx)
        rhs = np.vdot(fhy, x.astype(np.complex128))
        bound = 1e-6 * L * np.linalg.norm(x) * np.linalg.norm(y) * np.sqrt(n)
        assert abs(lhs - rhs) <= bound, (L, lhs, rhs)


def test_autograd_gradcheck_small():
    """torch.autograd.gradcheck-style check in FP32 on a small size: the backward is F^H."""
    n = 256
    x = torch.randn(3, n, dtype=torch.complex64, device="cuda", requires_grad=True)
    v = torch.randn(3, n, dtype=torch.complex64, device="cuda")
    y = F.fft_autograd(x)
    (g,) = torch.autograd.grad(y, x, v)
    want = torch.from_numpy(np.fft.ifft(v.cpu().numpy().astype(np.complex128), axis=1) * n).cuda()
    assert torch.allclose(g.to(torch.complex128), want, rtol=1e-5, atol=1e-4)
    y = F.fft_autograd(x, inverse=True)
    (g,) = torch.autograd.grad(y, x, v)
    want = torch.from_numpy(np.fft.fft(v.cpu().numpy().astype(np.complex128), axis=1)).cuda()
    assert torch.allclose(g.to(torch.complex128), want, rtol=1e-5, atol=1e-4)

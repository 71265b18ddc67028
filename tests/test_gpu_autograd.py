"""GPU: torch.autograd over the aclfft C API with the hand-written adjoints.

Differentiation configuration: 2D C2C 2048x2048 batch 4 and 1D R2C N=8192 batch 1024, N(0,1) inputs,
loss = sum |W * FFT(x)|^2 with W ~ U[0,1) fixed by seed.  The GPU gradient (FP32 forward +
adjoint kernels) is checked against
  - the analytic FP64 gradient 2 F^H (W^2 * F x) (numpy FP64 FFTs), and
  - FP64 central finite differences of the loss at random coordinates; the perturbed losses are
    evaluated in FP64 from y = F x and the exact response F e_j of a unit coordinate.
Also: ifft/irfft/fft2 norms and the C2R adjoint against torch's own FP64 CPU autograd.
"""
import numpy as np
import pytest

import aclfft
from aclfft import autograd as AG

pytestmark = pytest.mark.gpu


def _fd_coords(rng, shape, k):
    return [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(k)]


def test_grad_fft2_2048x2048_b4(cuda):
    import torch
    B, NX, NY = 4, 2048, 2048
    g = torch.Generator().manual_seed(5)
    x = torch.complex(torch.randn(B, NX, NY, generator=g), torch.randn(B, NX, NY, generator=g)).to(torch.complex64)
    W = torch.rand(B, NX, NY, generator=g)
    xd = x.to(cuda).requires_grad_(True)
    Wd = W.to(cuda)
    y = AG.fft2(xd)
    loss = (Wd * y).abs().pow(2).sum()
    loss.backward()
    grad = xd.grad.cpu().numpy().astype(np.complex128)

    x64 = x.numpy().astype(np.complex128)
    W64 = W.numpy().astype(np.float64)
    y64 = np.fft.fft2(x64)
    ref = 2 * np.fft.ifft2(W64 ** 2 * y64) * (NX * NY)  # 2 F^H (W^2 F x)
    err = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    print("fft2 grad rel_l2 vs analytic FP64: %.3e" % err)
    assert err < 2e-5

    L0 = np.sum(W64 ** 2 * np.abs(y64) ** 2)
    rng = np.random.default_rng(0)
    eps = 1e-3
    for (b, i, j) in _fd_coords(rng, (B, NX, NY), 4):
        # response of a unit impulse at (i, j): exp(-2 pi i (k1 i/NX + k2 j/NY))
        ph = np.exp(-2j * np.pi * (np.arange(NX)[:, None] * i / NX + np.arange(NY)[None, :] * j / NY))
        for unit, comp in ((1.0, "re"), (1j, "im")):
            d = unit * ph
            Lp = np.sum(W64[b] ** 2 * np.abs(y64[b] + eps * d) ** 2)
            Lm = np.sum(W64[b] ** 2 * np.abs(y64[b] - eps * d) ** 2)
            fd = (Lp - Lm) / (2 * eps)
            got = grad[b, i, j].real if comp == "re" else grad[b, i, j].imag
            assert abs(fd - got) <= 1e-4 * max(1.0, abs(fd)) + 1e-3 * np.sqrt(L0 / x64.size), (b, i, j, comp, fd, got)


def test_grad_rfft_8192_b1024(cuda):
    import torch
    N, B = 8192, 1024
    g = torch.Generator().manual_seed(9)
    x = torch.randn(B, N, generator=g)
    W = torch.rand(B, N // 2 + 1, generator=g)
    xd = x.to(cuda).requires_grad_(True)
    y = AG.rfft(xd)
    loss = (W.to(cuda) * y).abs().pow(2).sum()
    loss.backward()
    grad = xd.grad.cpu().numpy().astype(np.float64)

    x64 = x.numpy().astype(np.float64)
    W64 = W.numpy().astype(np.float64)
    y64 = np.fft.rfft(x64)
    gy = 2 * W64 ** 2 * y64
    full = np.zeros((B, N), complex)
    full[:, : N // 2 + 1] = gy
    ref = np.real(np.fft.ifft(full) * N)  # Re sum_{k<=N/2} g_k e^{+2 pi i k m / N}
    err = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    print("rfft grad rel_l2 vs analytic FP64: %.3e" % err)
    assert err < 2e-5

    rng = np.random.default_rng(1)
    eps = 1e-3
    k = np.arange(N // 2 + 1)
    for (b, m) in _fd_coords(rng, (B, N), 6):
        d = np.exp(-2j * np.pi * k * m / N)
        Lp = np.sum(W64[b] ** 2 * np.abs(y64[b] + eps * d) ** 2)
        Lm = np.sum(W64[b] ** 2 * np.abs(y64[b] - eps * d) ** 2)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - grad[b, m]) <= 1e-4 * max(1.0, abs(fd)), (b, m, fd, grad[b, m])


@pytest.mark.parametrize("norm", ["backward", "forward", "ortho"])
@pytest.mark.parametrize("n", [1024, 3000, 8192])
def test_irfft_and_ifft_grads_vs_torch_fp64(cuda, norm, n):
    import torch
    B = 16
    g = torch.Generator().manual_seed(n)
    X = torch.complex(torch.randn(B, n // 2 + 1, generator=g), torch.randn(B, n // 2 + 1, generator=g))
    w = torch.rand(B, n, generator=g)
    # ours
    Xd = X.to(torch.complex64).to(cuda).requires_grad_(True)
    out = AG.irfft(Xd, n, norm)
    (w.to(cuda) * out).pow(2).sum().backward()
    # torch FP64 on the CPU (third-party reference autograd)
    X64 = X.to(torch.complex128).requires_grad_(True)
    out64 = torch.fft.irfft(X64, n, norm=norm)
    (w.double() * out64).pow(2).sum().backward()
    assert torch.allclose(out.detach().cpu().double(), out64.detach(), rtol=1e-4, atol=1e-4 * out64.abs().max().item())
    e = (Xd.grad.cpu().to(torch.complex128) - X64.grad).norm() / X64.grad.norm()
    print("irfft n=%d norm=%s grad rel_l2 %.3e" % (n, norm, e))
    assert e < 1e-5
    # complex ifft
    xc = torch.complex(torch.randn(B, n, generator=g), torch.randn(B, n, generator=g))
    xd = xc.to(torch.complex64).to(cuda).requires_grad_(True)
    AG.ifft(xd, norm).abs().pow(2).mul(w.to(cuda)).sum().backward()
    x64 = xc.to(torch.complex128).requires_grad_(True)
    torch.fft.ifft(x64, norm=norm).abs().pow(2).mul(w.double()).sum().backward()
    e = (xd.grad.cpu().to(torch.complex128) - x64.grad).norm() / x64.grad.norm()
    assert e < 1e-5


def test_adjoint_entry_gates(cuda):
    import torch
    p = aclfft.Plan.plan1d(64, aclfft.C2C, 1)
    a = torch.zeros(64, dtype=torch.complex64, device=cuda)
    with pytest.raises(aclfft.AclfftError) as e:
        p.exec_r2c_adjoint(a, a)
    assert e.value.code == aclfft.INVALID_TYPE
